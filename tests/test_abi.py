"""The C-ABI boundary without a GPU: the library loads, exports every entry
point that include/dsph_abi.h declares (and nothing the bindings do not
know), and the product path refuses to run without it (no CPU fallback)."""
import ctypes as C
import re
from pathlib import Path

import pytest

from dualsphysics_amd import _native as N

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "dsph_abi.h"


def declared_functions() -> set[str]:
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return set(re.findall(r"\b(dsph_[a-z0-9_]+)\s*\(", text))


def test_header_and_bindings_agree():
    decl = declared_functions()
    assert decl, "no declarations parsed"
    assert decl == set(N.SIGNATURES), (sorted(decl - set(N.SIGNATURES)),
                                       sorted(set(N.SIGNATURES) - decl))


def test_library_exports_every_declared_symbol():
    if not N.LIB_PATH.exists():
        pytest.fail(f"{N.LIB_PATH} not built (python -m dualsphysics_amd.build)")
    lib = C.CDLL(str(N.LIB_PATH))
    missing = [name for name in sorted(declared_functions()) if not hasattr(lib, name)]
    assert not missing, missing


def declared_parameters() -> dict[str, list[str]]:
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    out = {}
    for m in re.finditer(r"\b(dsph_[a-z0-9_]+)\s*\(([^()]*)\)\s*;", text):
        params = [p.strip() for p in m.group(2).split(",")]
        out[m.group(1)] = [] if params in ([""], ["void"]) else params
    return out


def test_binding_arity_and_pointer_kinds_match_header():
    """Each binding has the header's parameter count, and pointers where the
    header has pointers (a missing or extra argument would shift every
    argument after it)."""
    decl = declared_parameters()
    assert set(decl) == set(N.SIGNATURES)
    for name, params in decl.items():
        args = N.SIGNATURES[name]
        assert len(args) == len(params), (name, len(args), len(params))
        for p, a in zip(params, args):
            is_ptr = "*" in p or "[" in p
            a_ptr = a in (C.c_void_p, C.c_char_p) or issubclass(a, C._Pointer)
            assert is_ptr == a_ptr, (name, p, a)


def test_bindings_load_with_signatures():
    lib = N.load()
    for name, args in N.SIGNATURES.items():
        fn = getattr(lib, name)
        assert list(fn.argtypes) == list(args), name


STRUCTS = {"dsph_cte": N.DsphCte, "dsph_divdata": N.DsphDivData,
           "dsph_inter_parms": N.DsphInterParms, "dsph_case": N.DsphCase,
           "dsph_stats": N.DsphStats, "dsph_motion": N.DsphMotion,
           "dsph_mlp": N.DsphMlp, "dsph_ens_state": N.DsphEnsState,
           "dsph_ens_shape": N.DsphEnsShape,
           "dsph_floating": N.DsphFloating, "dsph_floating_state": N.DsphFloatingState}


def test_struct_layouts_match_the_c_compiler(tmp_path):
    """sizeof/offsetof of every ABI struct as gcc lays it out from the header
    equal the ctypes mirrors used by the bindings."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "dsph_abi.h"',
             "int main(void) {"]
    for sname, cls in STRUCTS.items():
        lines.append(f'  printf("{sname} size %zu\\n", sizeof({sname}));')
        for fname, _ in cls._fields_:
            lines.append(f'  printf("{sname} {fname} %zu\\n", offsetof({sname}, {fname}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    r = subprocess.run([cc, "-std=c99", f"-I{ROOT / 'include'}", str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for sname, cls in STRUCTS.items():
        assert got[(sname, "size")] == C.sizeof(cls), sname
        for fname, _ in cls._fields_:
            assert got[(sname, fname)] == getattr(cls, fname).offset, (sname, fname)


def test_product_path_fails_loudly_without_library(monkeypatch, tmp_path):
    monkeypatch.setattr(N, "_lib", None)
    monkeypatch.setattr(N, "LIB_PATH", tmp_path / "missing.so")
    with pytest.raises(RuntimeError, match="no non-HIP fallback"):
        N.load()
    from dualsphysics_amd import sim
    from dualsphysics_amd import case as casemod
    c = casemod.dam_break(dp=0.05)
    with pytest.raises(RuntimeError):
        sim.SphGpuSingle(c)


def test_particle_counts_beyond_the_32bit_offset_limit_are_refused():
    """The interaction kernels address float4 arrays with 32-bit byte
    offsets: simulations and ensembles of 2^28 particles or more are refused
    with DSPH_E_ARG before any device call (so this runs without a GPU)."""
    from dualsphysics_amd import case as casemod
    lib = N.load()
    c = casemod.dam_break(dp=0.05)
    nc = c.to_native()
    dummy = (C.c_double * 16)()
    h = C.c_void_p()
    for np_ in (N.NP_LIMIT, N.NP_LIMIT + 5, 0xFFFFFFFF):
        rc = lib.dsph_sim_create(0, C.byref(nc), np_, dummy, dummy, dummy, None, C.byref(h))
        assert rc == N.DSPH_E_ARG and not h.value
    cases = (N.DsphCase * 2)(nc, nc)
    npk = (C.c_uint32 * 2)(N.NP_LIMIT // 2, N.NP_LIMIT // 2)
    tend = (C.c_double * 2)(1.0, 1.0)
    rc = lib.dsph_ens_create(0, 2, cases, npk, dummy, dummy, dummy, tend, C.byref(h))
    assert rc == N.DSPH_E_ARG and not h.value
    with open(HEADER) as f:
        text = f.read()
    assert re.search(r"#define DSPH_NP_LIMIT 268435456u", text)


def test_noalloc_limit_matches_header():
    """The particle limit of the no-alloc interaction entry point is the one
    the header states, and its scratch (24 bytes per particle) is 24 MiB."""
    text = HEADER.read_text()
    m = re.search(r"#define DSPH_NOALLOC_NP_MAX (\d+)u", text)
    assert m and int(m.group(1)) == N.NOALLOC_NP_MAX
    assert N.NOALLOC_NP_MAX * (8 + 16) == 24 << 20
