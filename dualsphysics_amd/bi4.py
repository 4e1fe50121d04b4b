"""Case and particle files of the reference: the case XML (<execution>
section: constants, particle blocks, motion, parameters) with its .bi4
particle file, and the PART_xxxx.bi4 files a run saves (every PART is a
restart point). Reading a case written by GenCase or by write_case gives a
:class:`dualsphysics_amd.case.Case` for the GPU solver; write_part /
read_part save and reload a run's state.

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
def write_case(case, directory, name: str, timemax: float | None = None,
               timeout: float | None = None, gauges=None):
    """Writes <name>.xml and <name>.bi4 for the reference solver: the
    <execution> section GenCase would produce (constants, parameters,
    particle blocks mk 0 = fixed walls, mk 1 = fluid) and the particle file
    with FP64 positions."""
    d = Path(directory)
    d.mkdir(parents=True, exist_ok=True)
    p = case.params
    timemax = p.timemax if timemax is None else timemax
    timeout = (p.timeout or timemax) if timeout is None else timeout
    n, nb, nf = case.np, case.nbound, case.nfluid
    pmin, pmax = case.pos.min(0), case.pos.max(0)
    root = Item("JPartDataBi4")
    root.set("Piece", T_UINT, 0).set("Npiece", T_UINT, 1)
    root.set("RunCode", T_TEXT, "00000000").set("Date", T_TEXT, "???")
    root.set("AppName", T_TEXT, "dualsphysics_amd case writer").set("CaseName", T_TEXT, name)
    root.set("Data2d", T_BOOL, bool(case.data2d))
    root.set("Data2dPosY", T_DOUBLE, float(case.data2dposy) if case.data2d else 0.0)
    root.set("MapPosMin", T_DOUBLE3, (0.0, 0.0, 0.0)).set("MapPosMax", T_DOUBLE3, (0.0, 0.0, 0.0))
    root.set("PeriMode", T_INT, 0)
    for k in ("PeriXinc", "PeriYinc", "PeriZinc"):
        root.set(k, T_DOUBLE3, (0.0, 0.0, 0.0))
    root.set("AxisDiv", T_INT, 0)
    nmov = getattr(case, "nmoving", 0)
    nflt = case.nfloat
    nfix = nb - nflt - nmov
    root.set("CaseNp", T_ULLONG, n).set("CaseNfixed", T_ULLONG, nfix)
    root.set("CaseNmoving", T_ULLONG, nmov).set("CaseNfloat", T_ULLONG, nflt)
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
        "DensityDTvalue": p.ddtvalue, "Shifting": getattr(p, "shifting", 0),
        "ShiftCoef": getattr(p, "shiftcoef", -2.0), "ShiftTFS": getattr(p, "shifttfs", 0.0),
        "RigidAlgorithm": 1, "FtPause": 0.0,
        "CoefDtMin": 0.05, "DtIni": p.dtini, "DtMin": p.dtmin, "DtFixed": 0,
        "DtAllParticles": 0, "TimeMax": timemax, "TimeOut": timeout,
        "RhopOutMin": p.rhopoutmin, "RhopOutMax": p.rhopoutmax,
        "Boundary": p.tboundary,
    }
    if p.tboundary == 2:
        params["SlipMode"] = getattr(p, "slipmode", 1)
        if getattr(p, "slipmode", 1) >= 2 and getattr(p, "nopenetration", 0):
            params["NoPenetration"] = 1
    # Periodic axes (JSph.cpp:810-826): a key of the axis' increments turns it on.
    peri = getattr(p, "periodic", 0)
    for a, keys in enumerate(_PERI_KEYS):
        if peri & (1 << a):
            for key, comp in keys:
                params[key] = float(p.perinc[a][comp])
    moving, motion = "", ""
    # Floating bodies (JCasePartBlock_Floating::ReadXml, JCaseParts.cpp:274-297):
    # mkbound 2.. after the fixed (0) and moving (1) blocks.
    floating = ""
    for k, f in enumerate(case.floatings or []):
        iv = " ".join(f'v{i + 1}{j + 1}="{r(float(f.inertia[3 * i + j]))}"'
                      for i in range(3) for j in range(3))
        c3 = lambda tag, x: f'<{tag} x="{r(float(x[0]))}" y="{r(float(x[1]))}" z="{r(float(x[2]))}" />'
        floating += (f'            <floating mkbound="{2 + k}" mk="{13 + k}" begin="{f.begin}" '
                     f'count="{f.count}">\n'
                     f'                <massbody value="{r(float(f.massbody))}" />\n'
                     f'                <masspart value="{r(float(f.masspart))}" />\n'
                     f'                {c3("center", f.center)}\n'
                     f'                <inertia {iv} />\n'
                     f'                {c3("linearvelini", f.velini)}\n'
                     f'                {c3("angularvelini", f.omegaini)}\n'
                     f'            </floating>\n')
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
    d2 = (f'<data2d value="true" />\n            <data2dposy value="{r(case.data2dposy)}" />'
          if case.data2d else '<data2d value="false" />')
    xml = f"""<?xml version="1.0" encoding="UTF-8" ?>
<case app="dualsphysics_amd case writer" date="???">
    <execution>
        <constants>
            {d2}
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
{moving}{floating}            <fluid mkfluid="0" mk="1" begin="{nb}" count="{nf}" />
        </particles>
{motion}
        <parameters>
{plines}
            <simulationdomain>
                <posmin x="default" y="default" z="default" />
                <posmax x="default" y="default" z="default + 50%" />
            </simulationdomain>
        </parameters>
{_gauges_xml(gauges)}    </execution>
</case>
"""
    (d / f"{name}.xml").write_text(xml)


def _gauges_xml(gs) -> str:
    """<special><gauges> of a dualsphysics_amd.gauges.GaugeSystem (the keys
    JGaugeSystem::ReadXml reads, JDsGaugeSystem.cpp:244-310); CSV output at
    every computation."""
    if gs is None or not gs.gauges:
        return ""
    from . import gauges as G
    r = repr
    v3 = lambda tag, x: f'<{tag} x="{r(float(x[0]))}" y="{r(float(x[1]))}" z="{r(float(x[2]))}" />'
    ind = " " * 16
    out = ["        <special>", "            <gauges>", "                <default>",
           f'                    <computedt value="{r(float(gs.computedt))}" />',
           '                    <output value="true" />',
           f'                    <outputdt value="{r(float(gs.computedt))}" />',
           "                </default>"]
    for g in gs.gauges:
        cd = ([f'{ind}    <computedt value="{r(float(g.computedt))}" />',
               f'{ind}    <outputdt value="{r(float(g.computedt))}" />']
              if g.computedt is not None else [])
        if isinstance(g, G.Velocity):
            out += [f'{ind}<velocity name="{g.name}">', f"{ind}    {v3('point', g.point)}"] + cd + \
                   [f"{ind}</velocity>"]
        elif isinstance(g, G.Swl):
            ml = ([f'{ind}    <masslimit value="{r(float(g.masslimit))}" />']
                  if g.masslimit is not None else [])
            out += [f'{ind}<swl name="{g.name}">', f"{ind}    {v3('point0', g.point0)}",
                    f"{ind}    {v3('point2', g.point2)}",
                    f'{ind}    <pointdp value="{r(float(g.pointdp))}" />'] + ml + cd + \
                   [f"{ind}</swl>"]
        elif isinstance(g, G.Force):
            out += [f'{ind}<force name="{g.name}">', f'{ind}    <target mkbound="{g.mkbound}" />'] + \
                   cd + [f"{ind}</force>"]
        else:
            out += [f'{ind}<maxz name="{g.name}">', f"{ind}    {v3('point0', g.point0)}",
                    f'{ind}    <height value="{r(float(g.height))}" />',
                    f'{ind}    <distlimit value="{r(float(g.distlimit))}" />'] + cd + \
                   [f"{ind}</maxz>"]
    out += ["            </gauges>", "        </special>", ""]
    return "\n".join(out)


def read_part(path, sort_by_id: bool = True) -> dict:
    """Particles of a PART_xxxx file, sorted by id (``sort_by_id``) or in the
    order the solver saved them; ``idp_saved`` keeps the ids in the saved
    order (the solver's particle order)."""
    root = read_file(path)
    part = next(it for it in root.items if it.name.startswith("PART_"))
    idp = part.get_array("Idp")
    pos = part.get_array("Posd")
    if pos is None:
        pos = part.get_array("Pos").astype(np.float64)
    vel = part.get_array("Vel")
    rho = part.get_array("Rhop")
    o = np.argsort(idp, kind="stable") if sort_by_id else np.arange(idp.shape[0])
    return {"idp": idp[o], "pos": pos[o], "vel": vel[o], "rho": rho[o], "idp_saved": idp,
            "time": part.get("TimeStep"), "step": part.get("Step"), "npok": part.get("Npok"),
            "cpart": part.get("Cpart"), "symplecticdtpre": part.get("SymplecticDtPre")}


# Case-XML keys of the periodic increments per axis: (key, component).
_PERI_KEYS = ((("XPeriodicIncY", 1), ("XPeriodicIncZ", 2)),
              (("YPeriodicIncX", 0), ("YPeriodicIncZ", 2)),
              (("ZPeriodicIncX", 0), ("ZPeriodicIncY", 1)))


def periodic_increments(case) -> list:
    """PeriXinc, PeriYinc, PeriZinc of a case as the solver holds them
    (JSph.cpp:2183-2188): the case's increments with each periodic axis' own
    component set to -MapRealSize; zero for axes that are not periodic."""
    p = case.params
    peri = int(getattr(p, "periodic", 0))
    size = [float(case.maprealposmax[k]) - float(case.maprealposmin[k]) for k in range(3)]
    out = []
    for a in range(3):
        if peri & (1 << a):
            inc = [float(p.perinc[a][k]) for k in range(3)]
            inc[a] = -size[a]
        else:
            inc = [0.0, 0.0, 0.0]
        out.append(tuple(inc))
    return out


def _set_periodic(root: "Item", case) -> None:
    """PeriMode (TpPeri = the axis bits of PeriActive, JPeriodicDef.h:33-42)
    and the increments of a PART or Part_Head header (JPartDataHead::
    ConfigSimPeri, JSph.cpp:2743): the reference's JPartsLoad4::CheckConfig
    refuses a resume whose periodic mode differs from the case's."""
    root.set("PeriMode", T_INT, int(getattr(case.params, "periodic", 0)) & 7)
    for k, inc in zip(("PeriXinc", "PeriYinc", "PeriZinc"), periodic_increments(case)):
        root.set(k, T_DOUBLE3, inc)


def _root_header(case, name: str, app: str = "dualsphysics_amd") -> "Item":
    """Root values of a particle file (JPartDataBi4::ConfigBasic/ConfigParticles,
    JPartDataBi4.cpp:105-190)."""
    n, nb, nf = case.np, case.nbound, case.nfluid
    nmov = getattr(case, "nmoving", 0)
    nflt = case.nfloat
    pmin, pmax = case.pos.min(0), case.pos.max(0)
    root = Item("JPartDataBi4")
    root.set("Piece", T_UINT, 0).set("Npiece", T_UINT, 1)
    root.set("RunCode", T_TEXT, "00000000").set("Date", T_TEXT, "???")
    root.set("AppName", T_TEXT, app).set("CaseName", T_TEXT, name)
    root.set("Data2d", T_BOOL, bool(case.data2d))
    root.set("Data2dPosY", T_DOUBLE, float(case.data2dposy) if case.data2d else 0.0)
    root.set("MapPosMin", T_DOUBLE3, tuple(map(float, case.maprealposmin)))
    root.set("MapPosMax", T_DOUBLE3, tuple(map(float, case.maprealposmax)))
    _set_periodic(root, case)
    root.set("AxisDiv", T_INT, 0)
    root.set("CaseNp", T_ULLONG, n).set("CaseNfixed", T_ULLONG, nb - nflt - nmov)
    root.set("CaseNmoving", T_ULLONG, nmov).set("CaseNfloat", T_ULLONG, nflt)
    root.set("CaseNfluid", T_ULLONG, nf)
    root.set("CasePosMin", T_DOUBLE3, tuple(map(float, pmin)))
    root.set("CasePosMax", T_DOUBLE3, tuple(map(float, pmax)))
    root.set("NpDynamic", T_BOOL, False).set("ReuseIds", T_BOOL, False)
    root.set("Dp", T_DOUBLE, case.dp).set("H", T_DOUBLE, case.h).set("B", T_DOUBLE, case.b)
    root.set("Rhop0", T_DOUBLE, case.rhop0).set("Gamma", T_DOUBLE, case.gamma)
    root.set("MassBound", T_DOUBLE, case.massbound).set("MassFluid", T_DOUBLE, case.massfluid)
    root.set("Symmetry", T_BOOL, False).set("Splitting", T_BOOL, False)
    return root


def write_part(path, case, idp, pos, velrho, timestep: float, step: int, cpart: int,
               symplecticdtpre: float | None = None, posdouble: bool = True):
    """Saves a run's particles as PART_<cpart> (JSph::SavePartData,
    JSph.cpp:3060-3110; JPartDataBi4::AddPartData): Idp, Posd (FP64) or Pos
    (FP32), Vel, Rhop in the given (the solver's) order, TimeStep, Step,
    Npok, and SymplecticDtPre for a resume."""
    root = _root_header(case, case.name)
    part = root.item(f"PART_{cpart:04d}")
    n = int(np.asarray(idp).shape[0])
    part.set("Cpart", T_UINT, cpart).set("TimeStep", T_DOUBLE, float(timestep))
    part.set("Npok", T_UINT, n).set("Nout", T_UINT, case.np - n).set("Step", T_UINT, int(step))
    part.set("RunTime", T_DOUBLE, 0.0)
    pos = np.asarray(pos, np.float64)
    part.set("DomainMin", T_DOUBLE3, tuple(map(float, pos.min(0))))
    part.set("DomainMax", T_DOUBLE3, tuple(map(float, pos.max(0))))
    if symplecticdtpre is not None:
        part.set("SymplecticDtPre", T_DOUBLE, float(symplecticdtpre))
    part.array("Idp", T_UINT, idp)
    if posdouble:
        part.array("Posd", T_DOUBLE3, pos)
    else:
        part.array("Pos", T_FLOAT3, pos.astype(np.float32))
    part.array("Vel", T_FLOAT3, np.asarray(velrho)[:, :3])
    part.array("Rhop", T_FLOAT, np.asarray(velrho)[:, 3])
    write_file(path, root)


def write_part_head(path, case, app: str = "dualsphysics_amd"):
    """Part_Head.ibi4 of a run's data directory (JPartDataHead,
    JPartDataHead.cpp): the case's constants and particle blocks, which a
    run resumed from one of its PARTs reads first (JPartsLoad4.cpp:150-190)."""
    p = case.params
    nmov = getattr(case, "nmoving", 0)
    nflt = case.nfloat
    nb = case.nbound
    pmin, pmax = case.pos.min(0), case.pos.max(0)
    root = Item("JPartDataHead")
    root.set("FmtVersion", T_UINT, 180324).set("AppName", T_TEXT, app)
    root.set("Date", T_TEXT, "???").set("RunCode", T_TEXT, "00000000")
    root.set("CaseName", T_TEXT, case.name).set("Data2d", T_BOOL, bool(case.data2d))
    root.set("Data2dPosY", T_DOUBLE, float(case.data2dposy) if case.data2d else 0.0)
    root.set("Npiece", T_UINT, 1).set("FirstPart", T_UINT, 0)
    root.set("CasePosMin", T_DOUBLE3, tuple(map(float, pmin)))
    root.set("CasePosMax", T_DOUBLE3, tuple(map(float, pmax)))
    root.set("NpDynamic", T_BOOL, False).set("ReuseIds", T_BOOL, False)
    root.set("MapPosMin", T_DOUBLE3, tuple(map(float, case.maprealposmin)))
    root.set("MapPosMax", T_DOUBLE3, tuple(map(float, case.maprealposmax)))
    _set_periodic(root, case)
    root.set("ViscoType", T_UINT, p.tvisco).set("ViscoValue", T_FLOAT, p.visco)
    root.set("ViscoBoundFactor", T_FLOAT, p.viscoboundfactor)
    root.set("Symmetry", T_BOOL, False).set("Splitting", T_BOOL, False)
    root.set("Dp", T_DOUBLE, case.dp).set("H", T_DOUBLE, case.h).set("B", T_DOUBLE, case.b)
    root.set("Gamma", T_DOUBLE, case.gamma).set("RhopZero", T_DOUBLE, case.rhop0)
    root.set("MassBound", T_DOUBLE, case.massbound).set("MassFluid", T_DOUBLE, case.massfluid)
    root.set("Gravity", T_FLOAT3, tuple(map(float, case.gravity)))
    root.set("CaseNp", T_ULLONG, case.np).set("CaseNfixed", T_ULLONG, nb - nflt - nmov)
    root.set("CaseNmoving", T_ULLONG, nmov).set("CaseNfloat", T_ULLONG, nflt)
    root.set("CaseNfluid", T_ULLONG, case.nfluid)
    blocks = [("Fixed", 11, 0, nb - nflt - nmov)] + ([("Moving", 12, 1, nmov)] if nmov else []) + \
             [("Floating", 13 + k, 2 + k, f.count) for k, f in enumerate(case.floatings or [])] + \
             [("Fluid", 1, 0, case.nfluid)]
    mk = root.item("MkBlocks")
    mk.set("Count", T_UINT, len(blocks))
    for k, (typ, mkv, mkt, cnt) in enumerate(blocks):
        b = mk.item(f"MkBlock_{k:03d}")
        b.set("Type", T_TEXT, typ).set("Mk", T_UINT, mkv).set("MkType", T_UINT, mkt)
        b.set("Count", T_UINT, cnt)
    write_file(path, root)


# ---------------------------------------------------------------- case XML reader
_NOT_BUILT_PARAMS = {
    "RigidAlgorithm": None, "FtPause": None,
    "DtFixed": "fixed dt", "PeriXinc": "periodic boundaries", "PeriYinc": "periodic boundaries",
    "PeriZinc": "periodic boundaries", "PeriX": "periodic boundaries", "PeriY": "periodic boundaries",
    "PeriZ": "periodic boundaries",
}


def _domain_limit(expr: str, base: float, size: float) -> float:
    """"default", "default + X%" or "default + X" (m), or a number
    (JSpaceEParms domain expressions; JSph::LoadConfigParameters)."""
    e = expr.replace(" ", "")
    if not e.startswith("default"):
        return float(e)
    rest = e[len("default"):]
    if not rest:
        return base
    sign = 1.0 if rest[0] == "+" else -1.0
    rest = rest[1:]
    if rest.endswith("%"):   # percent of the box size, away from the box (size < 0 for posmin)
        return base + sign * abs(size) * float(rest[:-1]) / 100.0 * (1.0 if size >= 0 else -1.0)
    return base + sign * float(rest)


def read_case(directory, name: str):
    """Reads <name>.xml and <name>.bi4 (GenCase output or write_case) into a
    Case: constants, parameters, fixed/moving/fluid particle blocks, the
    moving object's rectilinear sinusoid, rotation sinusoid or rotation
    (JMotion.cpp:705-800), mDBC normals ("BoundNor"). Options outside the
    built path (floating bodies, periodic boundaries, shifting, several
    moving objects, other kernels or motions) raise ValueError."""
    import math
    import xml.etree.ElementTree as ET
    from . import case as cm
    d = Path(directory)
    tree = ET.parse(d / f"{name}.xml")
    ex = tree.getroot().find("execution")
    if ex is None:
        raise ValueError("case XML without <execution> section")
    ct = ex.find("constants")
    val = lambda tag, a="value": float(ct.find(tag).get(a))
    g = ct.find("gravity")
    gravity = (float(g.get("x")), float(g.get("y")), float(g.get("z")))
    if ct.find("data2d") is not None and ct.find("data2d").get("value") in ("true", "1"):
        data2d, posy = True, float(ct.find("data2dposy").get("value"))
    else:
        data2d, posy = False, 0.0
    dp, h, b = val("dp"), val("h"), val("b")
    rhop0, gamma = val("rhop0"), val("gamma")
    massb, massf = val("massbound"), val("massfluid")
    # Parameters marked active="false" are skipped, as the solver does
    # (JCaseEParms::ReadXml with onlyactivated, JCaseEParms.cpp:454-466).
    prm = {e.get("key"): e.get("value") for e in ex.find("parameters").iter("parameter")
           if e.get("active", "true").strip().lower() not in ("false", "0")}
    for k, what in _NOT_BUILT_PARAMS.items():
        if what and k in prm and prm[k] not in ("0", "false", "0.0", "0:0:0"):
            raise ValueError(f"parameter {k}={prm[k]}: {what} not built")
    gi = lambda k, dflt: int(float(prm.get(k, dflt)))
    gf = lambda k, dflt: float(prm.get(k, dflt))
    if gi("Kernel", 2) != 2:
        raise ValueError("only the Wendland kernel is built")
    p = cm.Parameters(tstep=gi("StepAlgorithm", 1), verletsteps=gi("VerletSteps", 40),
                      tvisco=gi("ViscoTreatment", 1), visco=gf("Visco", 0.01),
                      viscoboundfactor=gf("ViscoBoundFactor", 1.0),
                      tdensity=gi("DensityDT", 0), ddtvalue=gf("DensityDTvalue", 0.1),
                      tboundary=gi("Boundary", 1), slipmode=gi("SlipMode", 1),
                      nopenetration=int(str(prm.get("NoPenetration", "0")).strip().upper()
                                        in ("1", "TRUE")),
                      cflnumber=float(ct.find("cflnumber").get("value"))
                      if ct.find("cflnumber") is not None else gf("CFLnumber", 0.2),
                      dtini=gf("DtIni", 0.0), dtmin=gf("DtMin", 0.0),
                      rhopoutmin=gf("RhopOutMin", 700.0), rhopoutmax=gf("RhopOutMax", 1300.0),
                      timemax=gf("TimeMax", 1.0), timeout=gf("TimeOut", 0.0),
                      shifting=gi("Shifting", 0), shiftcoef=gf("ShiftCoef", -2.0),
                      shifttfs=gf("ShiftTFS", 0.0))
    # Periodic boundaries (JSph.cpp:812-826): an increment key turns its axis
    # on with that component; XYPeriodic / XZPeriodic / YZPeriodic set the
    # two axes with zero increments (and turn the third off).
    peri, perinc = 0, [[0.0, 0.0, 0.0] for _ in range(3)]
    for a, keys in enumerate(_PERI_KEYS):
        for key, comp in keys:
            if key in prm:
                perinc[a][comp] = float(prm[key])
                peri |= 1 << a
    for key, bits in (("XYPeriodic", 3), ("XZPeriodic", 5), ("YZPeriodic", 6)):
        if key in prm:
            peri, perinc = bits, [[0.0, 0.0, 0.0] for _ in range(3)]
    unknown = [k for k in prm if "Periodic" in k and k not in
               {key for keys in _PERI_KEYS for key, _ in keys} | {"XYPeriodic", "XZPeriodic",
                                                                   "YZPeriodic"}]
    if unknown:
        raise ValueError(f"periodic parameter(s) {unknown} not recognised")
    p.periodic = peri
    p.perinc = tuple(tuple(v) for v in perinc)
    parts = ex.find("particles")
    blocks = list(parts)
    # Header counts against the blocks (JCaseParts::ReadXml, JCaseParts.cpp:566-568):
    # np all particles, nb all but fluid, nbf the fixed ones.
    cnt = {t: sum(int(bk.get("count")) for bk in blocks if bk.tag == t)
           for t in ("fixed", "moving", "floating", "fluid")}
    if (int(parts.get("np")) != sum(cnt.values()) or
            int(parts.get("nb")) != int(parts.get("np")) - cnt["fluid"] or
            int(parts.get("nbf")) != cnt["fixed"]):
        raise ValueError("The amount of particles does not match the header.")
    floatings = []
    v3 = lambda e, d=(0.0, 0.0, 0.0): d if e is None else \
        (float(e.get("x")), float(e.get("y")), float(e.get("z")))
    for bk in blocks:
        if bk.tag != "floating":
            continue
        if any(bk.find(t) is not None for t in ("translation", "translationDOF", "rotation",
                                                 "rotationDOF", "linearvel", "angularvel",
                                                 "linearforce", "angularforce")):
            raise ValueError("floating constraints, imposed velocities or forces are not built")
        ine = bk.find("inertia")
        if ine.get("x") is not None:
            d3 = v3(ine)
            inertia = (d3[0], 0.0, 0.0, 0.0, d3[1], 0.0, 0.0, 0.0, d3[2])
        else:
            inertia = tuple(float(ine.get(f"v{i}{j}", 0.0)) for i in (1, 2, 3) for j in (1, 2, 3))
        vel = bk.find("linearvelini")
        vel = bk.find("velini") if vel is None else vel
        om = bk.find("angularvelini")
        om = bk.find("omegaini") if om is None else om
        floatings.append(cm.Floating(
            begin=int(bk.get("begin")), count=int(bk.get("count")),
            massbody=float(bk.find("massbody").get("value")),
            masspart=float(bk.find("masspart").get("value")), center=v3(bk.find("center")),
            inertia=inertia, velini=v3(vel), omegaini=v3(om)))
    nmov = sum(int(bk.get("count")) for bk in blocks if bk.tag == "moving")
    if sum(1 for bk in blocks if bk.tag == "moving") > 1:
        raise ValueError("several moving objects are not built")
    root = read_file(d / f"{name}.bi4")
    part = next(it for it in root.items if it.name.startswith("PART_"))
    idp = part.get_array("Idp").astype(np.uint32)
    pos = part.get_array("Posd")
    pos = part.get_array("Pos").astype(np.float64) if pos is None else pos.astype(np.float64)
    vel = part.get_array("Vel")
    rho = part.get_array("Rhop")
    o = np.argsort(idp, kind="stable")
    idp, pos, vel, rho = idp[o], pos[o], vel[o], rho[o]
    nb = int(root.get("CaseNfixed", 0)) + int(root.get("CaseNmoving", 0))
    nflt = int(root.get("CaseNfloat", 0))
    if nflt != sum(f.count for f in floatings):
        raise ValueError("CaseNfloat does not match the floating blocks")
    nb += nflt
    velrho = np.concatenate([vel, rho[:, None]], 1).astype(np.float32)
    bn = part.get_array("BoundNor")
    if bn is not None:   # per particle (file order) or per boundary particle (id order)
        bn = bn[o] if bn.shape[0] == o.shape[0] else bn
    boundnor = None if bn is None else np.ascontiguousarray(bn[:nb], np.float32)
    motion = None
    if nmov:
        mo = ex.find("motion")
        mv = None if mo is None else next((e for e in mo.iter() if e.tag.startswith("mv")), None)
        if mv is None:
            raise ValueError("moving particles without a motion")
        beg = mo.find(".//begin")
        start = float(beg.get("start", 0.0)) if beg is not None else 0.0
        finish = start + float(mv.get("duration", 1e30))
        deg = mv.get("anglesunits", "degrees") == "degrees"
        v3 = lambda e: (float(e.get("x")), float(e.get("y")), float(e.get("z")))
        if mv.tag == "mvrectsinu":
            ph = v3(mv.find("phase")) if mv.find("phase") is not None else (0.0, 0.0, 0.0)
            if deg:
                ph = tuple(math.radians(x) for x in ph)
            motion = {"freq": v3(mv.find("freq")), "ampl": v3(mv.find("ampl")), "phase": ph,
                      "start": start, "finish": finish}
        elif mv.tag == "mvrotsinu":
            ampl = float(mv.find("ampl").get("v"))
            ph = float(mv.find("phase").get("v")) if mv.find("phase") is not None else 0.0
            motion = {"type": "rotsinu", "axisp1": v3(mv.find("axisp1")),
                      "axisp2": v3(mv.find("axisp2")), "angfreq": float(mv.find("freq").get("v")),
                      "angampl": ampl if deg else math.degrees(ampl),
                      "angphase": math.radians(ph) if deg else ph, "start": start,
                      "finish": finish}
        elif mv.tag == "mvrot":
            va = float(mv.find("vel").get("ang"))
            motion = {"type": "rot", "axisp1": v3(mv.find("axisp1")),
                      "axisp2": v3(mv.find("axisp2")), "angvel": va if deg else math.degrees(va),
                      "start": start, "finish": finish}
        else:
            raise ValueError(f"motion <{mv.tag}> is not built")
    # Domain: the particles' box widened by h*0.05 ("default"), then the
    # expressions of <simulationdomain> ("default + X%" of that box's size).
    bmin, bmax = (np.array(v) for v in cm.map_limits(pos, h, zextra=0.0, periodic=p.periodic,
                                                     dp=dp))
    size = bmax - bmin
    mn, mx = bmin.copy(), bmax.copy()
    sd = ex.find("parameters").find("simulationdomain")
    if sd is not None:
        lo, hi = sd.find("posmin"), sd.find("posmax")
        for k, a in enumerate("xyz"):
            if lo is not None:
                mn[k] = _domain_limit(lo.get(a), bmin[k], -size[k])
            if hi is not None:
                mx[k] = _domain_limit(hi.get(a), bmax[k], size[k])
    return cm.Case(name=name, dp=dp, h=h, rhop0=rhop0, gamma=gamma,
                   cs0=math.sqrt(b * gamma / rhop0), b=b, massfluid=massf, massbound=massb,
                   gravity=gravity, hswl=0.0, params=p, idp=idp, pos=pos, velrho=velrho,
                   nbound=nb, nfluid=int(idp.shape[0]) - nb, maprealposmin=tuple(map(float, mn)),
                   maprealposmax=tuple(map(float, mx)), boundnor=boundnor, nmoving=nmov,
                   motion=motion, floatings=floatings or None, data2d=data2d,
                   data2dposy=posy)
