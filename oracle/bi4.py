"""Reader and writer of the reference's binary particle files (.bi4) and the
case XML, to run the reference CPU build (oracle/_ref) on the exact particles
this package simulates.

TEST INFRASTRUCTURE ONLY: used by tests/golden/make_reference_fixtures.py
and the ref-marked tests; the product never imports it.

The .bi4 container (JBinaryData, JBinaryData.cpp:1280-1297 header,
:1163-1198 items/arrays, :700-1040 value encoding):
  header  64 bytes: "#FileJBD <root name>" padded with ' ' to 58, '\\n', 0,
          byte order (0 little endian), si64 flag, 2 unused bytes;
  item    u32 size of its definition, definition = str "\\nITEM\\n", str name,
          i32 hide, i32 hidevalues, str fmtfloat, str fmtdouble, u32 narrays,
          u32 nitems, u32 size of the values block; values block = str
          "\\nVALUES", u32 count, per value str name, i32 type, payload; then
          the arrays, then the sub-items;
  array   u32 size of its definition, definition = str "\\nARRAY", str name,
          i32 hide, i32 type, u32 count, u32 data size; then the raw data;
  str     u32 length + bytes (no terminator).
The particle file layout (root "JPartDataBi4", item "PART_%04u" with arrays
Idp, Posd|Pos, Vel, Rhop) follows JPartDataBi4.cpp:57-455 and what
JPartsLoad4::LoadParticles reads (JPartsLoad4.cpp:150-290).
"""
from __future__ import annotations

import io
import struct
from pathlib import Path

import numpy as np

# JBinaryDataDef::TpData (JBinaryData.h:83-87).
T_TEXT, T_BOOL, T_CHAR, T_UCHAR, T_SHORT, T_USHORT, T_INT, T_UINT, T_LLONG, T_ULLONG = range(1, 11)
T_FLOAT, T_DOUBLE = 11, 12
T_INT3, T_UINT3, T_FLOAT3, T_DOUBLE3 = 20, 21, 22, 23
_SCALAR = {T_BOOL: "<i", T_CHAR: "<b", T_UCHAR: "<B", T_SHORT: "<h", T_USHORT: "<H", T_INT: "<i",
           T_UINT: "<I", T_LLONG: "<q", T_ULLONG: "<Q", T_FLOAT: "<f", T_DOUBLE: "<d",
           T_INT3: "<3i", T_UINT3: "<3I", T_FLOAT3: "<3f", T_DOUBLE3: "<3d"}
_NPTYPE = {T_CHAR: (np.int8, 1), T_UCHAR: (np.uint8, 1), T_SHORT: (np.int16, 1),
           T_USHORT: (np.uint16, 1), T_INT: (np.int32, 1), T_UINT: (np.uint32, 1),
           T_LLONG: (np.int64, 1), T_ULLONG: (np.uint64, 1), T_FLOAT: (np.float32, 1),
           T_DOUBLE: (np.float64, 1), T_INT3: (np.int32, 3), T_UINT3: (np.uint32, 3),
           T_FLOAT3: (np.float32, 3), T_DOUBLE3: (np.float64, 3)}


def _str(s: str) -> bytes:
    b = s.encode()
    return struct.pack("<I", len(b)) + b


class Item:
    """One JBinaryData item: ordered typed values, arrays and sub-items."""

    def __init__(self, name: str):
        self.name = name
        self.values: list[tuple[str, int, object]] = []
        self.arrays: list[tuple[str, int, np.ndarray]] = []
        self.items: list[Item] = []

    # ---- building
    def set(self, name, typ, v):
        self.values = [x for x in self.values if x[0] != name] + [(name, typ, v)]
        return self

    def array(self, name, typ, a):
        dt, k = _NPTYPE[typ]
        a = np.ascontiguousarray(a, dt)
        assert a.size % k == 0
        self.arrays.append((name, typ, a))
        return self

    def item(self, name) -> "Item":
        it = Item(name)
        self.items.append(it)
        return it

    # ---- access
    def get(self, name, default=None):
        for n, _, v in self.values:
            if n == name:
                return v
        return default

    def get_array(self, name):
        for n, _, a in self.arrays:
            if n == name:
                return a
        return None

    def get_item(self, name):
        for it in self.items:
            if it.name == name:
                return it
        return None

    # ---- encoding
    def _values_block(self) -> bytes:
        out = [_str("\nVALUES"), struct.pack("<I", len(self.values))]
        for n, t, v in self.values:
            out.append(_str(n) + struct.pack("<i", t))
            if t == T_TEXT:
                out.append(_str(v))
            elif t == T_BOOL:
                out.append(struct.pack("<i", 1 if v else 0))
            elif t in (T_INT3, T_UINT3, T_FLOAT3, T_DOUBLE3):
                out.append(struct.pack(_SCALAR[t], *v))
            else:
                out.append(struct.pack(_SCALAR[t], v))
        return b"".join(out)

    def encode(self) -> bytes:
        vals = self._values_block()
        d = (_str("\nITEM\n") + _str(self.name) + struct.pack("<ii", 0, 0) + _str("%.7E") +
             _str("%.15E") + struct.pack("<III", len(self.arrays), len(self.items), len(vals)))
        out = [struct.pack("<I", len(d)), d, vals]
        for n, t, a in self.arrays:
            k = _NPTYPE[t][1]
            raw = a.tobytes()
            ad = (_str("\nARRAY") + _str(n) + struct.pack("<ii", 0, t) +
                  struct.pack("<II", a.size // k, len(raw)))
            out += [struct.pack("<I", len(ad)), ad, raw]
        for it in self.items:
            out.append(it.encode())
        return b"".join(out)


def _rstr(f) -> str:
    (n,) = struct.unpack("<I", f.read(4))
    return f.read(n).decode()


def _read_item(f) -> Item:
    (sdef,) = struct.unpack("<I", f.read(4))
    d = io.BytesIO(f.read(sdef))
    if _rstr(d) != "\nITEM\n":
        raise ValueError("bad item code")
    it = Item(_rstr(d))
    struct.unpack("<ii", d.read(8))
    _rstr(d), _rstr(d)
    narr, nit, svals = struct.unpack("<III", d.read(12))
    if svals:
        v = io.BytesIO(f.read(svals))
        if _rstr(v) != "\nVALUES":
            raise ValueError("bad values code")
        (num,) = struct.unpack("<I", v.read(4))
        for _ in range(num):
            name = _rstr(v)
            (t,) = struct.unpack("<i", v.read(4))
            if t == T_TEXT:
                val = _rstr(v)
            else:
                fmt = _SCALAR[t]
                val = struct.unpack(fmt, v.read(struct.calcsize(fmt)))
                val = val if len(val) > 1 else val[0]
                if t == T_BOOL:
                    val = bool(val)
            it.values.append((name, t, val))
    for _ in range(narr):
        (sad,) = struct.unpack("<I", f.read(4))
        a = io.BytesIO(f.read(sad))
        if _rstr(a) != "\nARRAY":
            raise ValueError("bad array code")
        name = _rstr(a)
        _, t = struct.unpack("<ii", a.read(8))
        count, size = struct.unpack("<II", a.read(8))
        raw = f.read(size)
        if t == T_TEXT:
            raise ValueError("text arrays are not used in particle files")
        dt, k = _NPTYPE[t]
        arr = np.frombuffer(raw, dt).copy()
        it.arrays.append((name, t, arr.reshape(count, k) if k > 1 else arr))
    for _ in range(nit):
        it.items.append(_read_item(f))
    return it


def write_file(path, root: Item):
    titu = ("#FileJBD " + root.name)[:58].ljust(58) + "\n"
    head = titu.encode() + b"\0" + bytes([0, 0, 0, 0])
    assert len(head) == 64
    Path(path).write_bytes(head + root.encode())


def read_file(path) -> Item:
    with open(path, "rb") as f:
        head = f.read(64)
        if not head.startswith(b"#FileJBD "):
            raise ValueError(f"{path}: not a JBinaryData file")
        if head[60] != 0 or head[61] != 0:
            raise ValueError(f"{path}: big-endian or 64-bit-size files are not supported")
        return _read_item(f)


# ---------------------------------------------------------------- case files
def write_case(case, directory, name: str, timemax: float = 1.0, timeout: float = 1.0):
    """Writes <name>.xml and <name>.bi4 for the reference solver: the
    <execution> section GenCase would produce (constants, parameters,
    particle blocks mk 0 = fixed walls, mk 1 = fluid) and the particle file
    with FP64 positions."""
    d = Path(directory)
    d.mkdir(parents=True, exist_ok=True)
    p = case.params
    n, nb, nf = case.np, case.nbound, case.nfluid
    pmin, pmax = case.pos.min(0), case.pos.max(0)
    root = Item("JPartDataBi4")
    root.set("Piece", T_UINT, 0).set("Npiece", T_UINT, 1)
    root.set("RunCode", T_TEXT, "00000000").set("Date", T_TEXT, "???")
    root.set("AppName", T_TEXT, "dualsphysics_amd case writer").set("CaseName", T_TEXT, name)
    root.set("Data2d", T_BOOL, False).set("Data2dPosY", T_DOUBLE, 0.0)
    root.set("MapPosMin", T_DOUBLE3, (0.0, 0.0, 0.0)).set("MapPosMax", T_DOUBLE3, (0.0, 0.0, 0.0))
    root.set("PeriMode", T_INT, 0)
    for k in ("PeriXinc", "PeriYinc", "PeriZinc"):
        root.set(k, T_DOUBLE3, (0.0, 0.0, 0.0))
    root.set("AxisDiv", T_INT, 0)
    nmov = getattr(case, "nmoving", 0)
    nfix = nb - nmov
    root.set("CaseNp", T_ULLONG, n).set("CaseNfixed", T_ULLONG, nfix)
    root.set("CaseNmoving", T_ULLONG, nmov).set("CaseNfloat", T_ULLONG, 0)
    root.set("CaseNfluid", T_ULLONG, nf)
    root.set("CasePosMin", T_DOUBLE3, tuple(map(float, pmin)))
    root.set("CasePosMax", T_DOUBLE3, tuple(map(float, pmax)))
    root.set("NpDynamic", T_BOOL, False).set("ReuseIds", T_BOOL, False)
    root.set("Dp", T_DOUBLE, case.dp).set("H", T_DOUBLE, case.h).set("B", T_DOUBLE, case.b)
    root.set("Rhop0", T_DOUBLE, case.rhop0).set("Gamma", T_DOUBLE, case.gamma)
    root.set("MassBound", T_DOUBLE, case.massbound).set("MassFluid", T_DOUBLE, case.massfluid)
    root.set("Symmetry", T_BOOL, False).set("Splitting", T_BOOL, False)
    part = root.item("PART_0000")
    part.set("Cpart", T_UINT, 0).set("TimeStep", T_DOUBLE, 0.0).set("Npok", T_UINT, n)
    part.set("Nout", T_UINT, 0).set("Step", T_UINT, 0).set("RunTime", T_DOUBLE, 0.0)
    part.set("DomainMin", T_DOUBLE3, tuple(map(float, pmin)))
    part.set("DomainMax", T_DOUBLE3, tuple(map(float, pmax)))
    part.array("Idp", T_UINT, case.idp)
    part.array("Posd", T_DOUBLE3, case.pos)
    part.array("Vel", T_FLOAT3, case.velrho[:, :3])
    part.array("Rhop", T_FLOAT, case.velrho[:, 3])
    if p.tboundary == 2:
        if case.boundnor is None:
            raise ValueError("mDBC needs boundary normals")
        part.array("BoundNor", T_FLOAT3, case.boundnor)
    write_file(d / f"{name}.bi4", root)

    r = repr
    g = case.gravity
    params = {
        "SavePosDouble": 1, "StepAlgorithm": p.tstep, "VerletSteps": p.verletsteps,
        "Kernel": 2, "ViscoTreatment": p.tvisco, "Visco": p.visco,
        "ViscoBoundFactor": p.viscoboundfactor, "DensityDT": p.tdensity,
        "DensityDTvalue": p.ddtvalue, "Shifting": 0, "RigidAlgorithm": 1, "FtPause": 0.0,
        "CoefDtMin": 0.05, "DtIni": p.dtini, "DtMin": p.dtmin, "DtFixed": 0,
        "DtAllParticles": 0, "TimeMax": timemax, "TimeOut": timeout,
        "RhopOutMin": p.rhopoutmin, "RhopOutMax": p.rhopoutmax,
        "Boundary": p.tboundary,
    }
    if p.tboundary == 2:
        params["SlipMode"] = getattr(p, "slipmode", 1)
    moving, motion = "", ""
    if nmov:
        m = case.motion
        moving = (f'            <moving mkbound="1" mk="12" begin="{nfix}" count="{nmov}" '
                  f'refmotion="0" />\n')
        v3 = lambda tag, x: f'<{tag} x="{r(x[0])}" y="{r(x[1])}" z="{r(x[2])}" />'
        kind = m.get("type", "rectsinu")
        dur = r(min(m.get('finish', 1e30), 1e6))
        if kind == "rectsinu":
            mov = f"""<mvrectsinu id="1" duration="{dur}" anglesunits="radians">
                    {v3('freq', m['freq'])}
                    {v3('ampl', m['ampl'])}
                    {v3('phase', m['phase'])}
                </mvrectsinu>"""
        elif kind == "rotsinu":
            # Degrees: amplitude as given; the phase (radians here) is
            # written in degrees and converted back by the reader.
            import math
            mov = f"""<mvrotsinu id="1" duration="{dur}" anglesunits="degrees">
                    {v3('axisp1', m['axisp1'])}
                    {v3('axisp2', m['axisp2'])}
                    <freq v="{r(m['angfreq'])}" />
                    <ampl v="{r(m['angampl'])}" />
                    <phase v="{r(math.degrees(m['angphase']))}" />
                </mvrotsinu>"""
        else:
            mov = f"""<mvrot id="1" duration="{dur}" anglesunits="degrees">
                    <vel ang="{r(m['angvel'])}" />
                    {v3('axisp1', m['axisp1'])}
                    {v3('axisp2', m['axisp2'])}
                </mvrot>"""
        motion = f"""        <motion>
            <objreal ref="0">
                <begin mov="1" start="{r(m.get('start', 0.0))}" />
                {mov}
            </objreal>
        </motion>
"""
    plines = "\n".join(f'            <parameter key="{k}" value="{r(v)}" />'
                       for k, v in params.items())
    xml = f"""<?xml version="1.0" encoding="UTF-8" ?>
<case app="dualsphysics_amd case writer" date="???">
    <execution>
        <constants>
            <data2d value="false" />
            <gravity x="{r(g[0])}" y="{r(g[1])}" z="{r(g[2])}" />
            <cflnumber value="{r(p.cflnumber)}" />
            <gamma value="{r(case.gamma)}" />
            <rhop0 value="{r(case.rhop0)}" />
            <dp value="{r(case.dp)}" />
            <h value="{r(case.h)}" />
            <b value="{r(case.b)}" />
            <massbound value="{r(case.massbound)}" />
            <massfluid value="{r(case.massfluid)}" />
        </constants>
        <particles np="{n}" nb="{nb}" nbf="{nfix}" mkboundfirst="11" mkfluidfirst="1">
            <fixed mkbound="0" mk="11" begin="0" count="{nfix}" />
{moving}            <fluid mkfluid="0" mk="1" begin="{nb}" count="{nf}" />
        </particles>
{motion}
        <parameters>
{plines}
            <simulationdomain>
                <posmin x="default" y="default" z="default" />
                <posmax x="default" y="default" z="default + 50%" />
            </simulationdomain>
        </parameters>
    </execution>
</case>
"""
    (d / f"{name}.xml").write_text(xml)


def read_part(path) -> dict:
    """Particles of a PART_xxxx file, sorted by id; ``idp_saved`` keeps the
    ids in the order the solver saved them (its particle order)."""
    root = read_file(path)
    part = next(it for it in root.items if it.name.startswith("PART_"))
    idp = part.get_array("Idp")
    pos = part.get_array("Posd")
    if pos is None:
        pos = part.get_array("Pos").astype(np.float64)
    vel = part.get_array("Vel")
    rho = part.get_array("Rhop")
    o = np.argsort(idp, kind="stable")
    return {"idp": idp[o], "pos": pos[o], "vel": vel[o], "rho": rho[o], "idp_saved": idp,
            "time": part.get("TimeStep"), "step": part.get("Step"), "npok": part.get("Npok")}
