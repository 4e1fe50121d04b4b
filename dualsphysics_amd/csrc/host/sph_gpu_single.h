// Host driver state of the single-GPU SPH step (struct dsph_sim): the
// re-implementation of the reference's JSphGpuSingle / JSphGpu /
// JCellDivGpuSingle flow over the C ABI. Shared by the dsph_sim_* entry points
// and by the ensemble driver, which reuses its constant and cell-map setup.
#pragma once
#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>
#include "dsph_abi.h"

namespace dsph_host {

struct u3 { unsigned x, y, z; };
inline bool operator!=(const u3& a, const u3& b) { return a.x != b.x || a.y != b.y || a.z != b.z; }

// DCEL helpers (JDsDcellDef.h:28-45; JDsDcell.cpp:30-67).
inline unsigned calc_bits(unsigned v, unsigned minbits) {
  unsigned n = (minbits < 1 || minbits > 30) ? 1 : minbits;
  for (; v >> n; n++) {}
  return n;
}
inline unsigned calc_cell_code(u3 nc) {
  unsigned sx = calc_bits(nc.x, 2), sy = calc_bits(nc.y, 2), sz = calc_bits(nc.z, 2);
  const unsigned smin = sx + sy + sz;
  const unsigned maxbits = 31;
  if (smin > maxbits) return 0;
  unsigned rest = maxbits - smin;
  while (rest) {
    if (rest) { sx++; rest--; }
    if (rest) { sy++; rest--; }
    if (rest) { sz++; rest--; }
  }
  return ((sx + 1) << 25) | (sy << 20) | (sz << 15) | ((sy + sz) << 10) | ((sx + 1 + sz) << 5) |
         (sx + 1 + sy);
}
inline unsigned dcel_cell(unsigned dcc, unsigned cx, unsigned cy, unsigned cz) {
  return (cx << ((dcc >> 10) & 31u)) | (cy << ((dcc >> 15) & 31u)) | cz;
}

struct Err : std::runtime_error {
  int code;
  Err(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

template <typename T>
struct DevArr {
  T* p = nullptr;
  size_t n = 0;
};

}  // namespace dsph_host

namespace dsph {
void rect_sinu_mov_host(const dsph_motion* m, double t, double dt, double mv[3]);
}

using namespace dsph_host;

struct dsph_sim {
  dsph_ctx* ctx = nullptr;
  dsph_case cs{};
  std::string err;

  // ---- constants (JSph members)
  float KernelH = 0, KernelSize = 0, KernelSize2 = 0, Scell = 0, MovLimit = 0, PosCellSize = 0;
  float Eta2 = 0, awen = 0, bwenh = 0, DDTkh = 0, DDTgz = 0, CteB = 0, Gamma = 0, RhopZero = 0;
  float MassFluid = 0, MassBound = 0, Dp = 0, Visco = 0, ViscoBoundFactor = 1;
  float RhopOutMin = 700, RhopOutMax = 1300, SpsSmag = 0, SpsBlin = 0;
  float Gravity[3] = {0, 0, 0};
  double Cs0 = 0, DtIni = 0, DtMin = 0, CFLnumber = 0.2;
  int TStep = 2, VerletSteps = 40, TVisco = 1, TDensity = 0, TBoundary = 1, ScellDiv = 1;
  int SlipMode = DSPH_SLIP_VEL0;
  bool NoPen = false;  // mDBC2 NoPenetration (TMdbc2 == MDBC2_NoPen)
  bool Simulate2D = false;  // particles in the plane y = Simulate2DPosY (JSph.cpp:646-647)
  double Simulate2DPosY = 0;
  // mDBC2: no-slip or free slip (MdbcCorrector, JSph.cpp:897).
  bool mdbc2_on() const { return TBoundary == DSPH_BC_MDBC && SlipMode >= DSPH_SLIP_NOSLIP; }
  int ShiftMode = DSPH_SHIFT_NONE;
  double ShiftCoef = -2.0, ShiftTFS = 0.0;
  bool CellDomFixed = false;
  double MapRealPosMin[3], MapRealPosMax[3], MapRealSize[3];
  double Map_PosMin[3], Map_PosMax[3], Map_Size[3];
  u3 Map_Cells{}, DomCells{};
  double DomPosMin[3];
  unsigned DomCellCode = 0;
  unsigned CaseNbound = 0, CaseNpb = 0;
  // Periodic boundaries (JSph.cpp:810-826, 2183-2194).
  unsigned PeriActive = 0;
  double PeriInc[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  unsigned NpbPer = 0, NpfPer = 0;
  unsigned* perilist = nullptr;   // NpCap entries
  unsigned* periwork = nullptr;   // block offsets of the list
  float* acemod = nullptr;        // |ace|^2 of normal particles

  // ---- particle state
  unsigned Np = 0, Npb = 0, NpbOk = 0, NpCap = 0;
  unsigned *idp = nullptr, *dcell = nullptr, *idp2 = nullptr, *dcell2 = nullptr;
  uint16_t *code = nullptr, *code2 = nullptr;
  double *posxy = nullptr, *posz = nullptr, *posxy2 = nullptr, *posz2 = nullptr;
  float *velrho = nullptr, *velrho2 = nullptr, *poscell = nullptr;
  double *posxypre = nullptr, *poszpre = nullptr, *posxypre2 = nullptr, *poszpre2 = nullptr;
  float *velrhopre = nullptr, *velrhopre2 = nullptr, *velrhom1 = nullptr, *velrhom1b = nullptr;
  float *ace = nullptr, *ar = nullptr, *viscdt = nullptr;
  double *movxy = nullptr, *movz = nullptr;
  float *sps2strain = nullptr, *spstau = nullptr, *spstau2 = nullptr;
  float *boundnor = nullptr, *boundnor2 = nullptr;
  // mDBC2: boundmode (DSPH_BMODE_*) and tangential velocity of [0,Npb),
  // rewritten by every correction (JSphGpuSingle.cpp:700-707).
  uint8_t* boundmode = nullptr;
  float* tangenvel = nullptr;
  // mDBC2 with moving boundaries: motion velocity and acceleration per
  // particle (MotionVel_g, MotionAce_g; zero at the start, JSphGpu.cpp:834-835).
  float *motionvel = nullptr, *motionvel2 = nullptr, *motionace = nullptr, *motionace2 = nullptr;
  // mDBC2 no penetration (MDBC2_NoPen): float4 per particle written by every
  // interaction for the fluid, used by the corrector / Verlet step.
  float* nopenshift = nullptr;
  // Shifting: float4 per particle (sums, then the displacement), fluid part
  // cleared before every interaction (JSphShifting::InitGpu without zones).
  float* shiftposfs = nullptr;
  bool pre_active = false;
  // device scalars: red[0..2] (velmax^2, viscdtmax, acemax^2), dt slots, flags
  float* d_red = nullptr;
  double* d_dt = nullptr;   // [0]=dt_p [1]=dt_c [2]=SymplecticDtPre [3]=time
  unsigned* d_flags = nullptr;

  // ---- cell division (JCellDivGpuSingle state)
  unsigned *cellpart = nullptr, *sortpart = nullptr;
  int* begincell = nullptr;
  size_t begincell_cap = 0;
  u3 CellDomainMin{}, CellDomainMax{};
  unsigned Ncx = 0, Ncy = 0, Ncz = 0, Nsheet = 0, Nct = 0, Nctt = 0;
  unsigned BoxBoundIgnore = 0, BoxFluid = 0, BoxBoundOut = 0;
  bool BoundLimitOk = false, BoundDivideOk = false, DivideFull = false;
  u3 BoundLimitCellMin{}, BoundLimitCellMax{}, BoundDivideCellMin{}, BoundDivideCellMax{};
  unsigned NpFinal = 0, NpbFinal = 0, NpbIgnore = 0, NpbOut = 0, NpfOut = 0;
  bool BoundChanged = true;
  // Learned SPS closure (optional): network and the velocity-gradient buffer.
  dsph_mlp Closure{};
  bool UseClosure = false;
  float* velgrad = nullptr;
  float* closure_nu = nullptr;

  // ---- floating bodies (StFloatingData, JSph::FtObjs; JSph.cpp:1123-1167)
  struct FtObj {
    unsigned begin = 0, count = 0;
    float mass = 0, massp = 0;
    double center[3] = {0, 0, 0};
    float angles[3] = {0, 0, 0}, fvel[3] = {0, 0, 0}, fomega[3] = {0, 0, 0};
    float facelin[3] = {0, 0, 0}, faceang[3] = {0, 0, 0};
    float inertia[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  };
  std::vector<FtObj> ft;
  float* d_ftomassp = nullptr;    // FtoMasspg: particle mass per body
  unsigned* d_ftbody = nullptr;   // {first index into d_ridp, count} per body
  double* d_ftcenter = nullptr;   // FtoCenterg
  float* d_ftace = nullptr;       // FtoAceg: 6 sums per body
  unsigned* d_ridp = nullptr;     // RidpMotg for the floating ids

  // ---- run state
  double SymplecticDtPre = 0, TimeStep = 0, DtLast = 0;
  double VelMax = 0, AceMax = 0;
  float ViscDtMax = 0;
  unsigned Nstep = 0, VerletStep = 0, Ndiv = 0, NdivFull = 0, NpfOutTotal = 0;
  bool device_dt = false;

  ~dsph_sim() { release(); }

  void chk(int rc, const char* what) {
    if (rc != DSPH_OK) throw Err(rc, std::string(what) + ": " + dsph_last_error(ctx));
  }
  template <typename T>
  T* alloc(size_t n) {
    void* p = nullptr;
    chk(dsph_malloc(ctx, sizeof(T) * std::max<size_t>(n, 1), &p), "dsph_malloc");
    return static_cast<T*>(p);
  }
  void release() {
    if (!ctx) return;
    void* ptrs[] = {idp, dcell, idp2, dcell2, code, code2, posxy, posz, posxy2, posz2, velrho,
                    velrho2, poscell, posxypre, poszpre, posxypre2, poszpre2, velrhopre,
                    velrhopre2, velrhom1, velrhom1b, ace, ar, viscdt, movxy, movz, sps2strain,
                    spstau, spstau2, boundnor, boundnor2, boundmode, tangenvel, nopenshift, shiftposfs,
                    motionvel, motionvel2, motionace, motionace2,
                    d_red, d_dt,
                    d_flags, cellpart,
                    sortpart, begincell, velgrad, closure_nu, d_ftomassp, d_ftbody,
                    d_ftcenter, d_ftace, d_ridp};
    for (void* p : ptrs) dsph_free(ctx, p);
    dsph_ctx_destroy(ctx);
    ctx = nullptr;
  }

  // ---- configuration (JSph::ConfigConstants1/2, JSph.cpp:1481-1546)
  void config(const dsph_case& c) {
    cs = c;
    Dp = float(c.dp);
    KernelH = float(c.kernelh);
    CteB = float(c.cteb);
    Gamma = float(c.gamma);
    RhopZero = float(c.rhopzero);
    MassFluid = float(c.massfluid);
    MassBound = float(c.massbound);
    for (int k = 0; k < 3; k++) Gravity[k] = float(c.gravity[k]);
    CFLnumber = c.cflnumber;
    TStep = c.tstep;
    VerletSteps = c.verletsteps ? c.verletsteps : 40;
    TVisco = c.tvisco;
    TDensity = c.tdensity;
    TBoundary = c.tboundary;
    SlipMode = c.slipmode ? c.slipmode : DSPH_SLIP_VEL0;
    if (TBoundary == DSPH_BC_MDBC && (SlipMode < DSPH_SLIP_VEL0 || SlipMode > DSPH_SLIP_FREESLIP))
      throw Err(DSPH_E_ARG, "SlipMode: 1 (vel=0), 2 (no-slip) or 3 (free slip)");
    // NoPenetration is read for SlipMode >= no-slip only (JSph.cpp:715-716).
    NoPen = mdbc2_on() && c.nopenetration != 0;
    ShiftMode = c.shiftmode;
    ShiftCoef = c.shiftcoef ? c.shiftcoef : -2.0;
    ShiftTFS = c.shifttfs;
    if (ShiftMode < DSPH_SHIFT_NONE || ShiftMode > DSPH_SHIFT_FULL)
      throw Err(DSPH_E_ARG, "Shifting: mode 0..3");
    if (ShiftMode && NoPen)
      throw Err(DSPH_E_ARG, "shifting with mDBC2 no penetration is not built");
    Visco = float(c.visco);
    ViscoBoundFactor = float(c.viscoboundfactor);
    RhopOutMin = float(c.rhopoutmin);
    RhopOutMax = float(c.rhopoutmax);
    ScellDiv = (c.cellmode == 2 ? 2 : 1);
    CellDomFixed = c.celldomfixed != 0;
    if (c.tkernel != DSPH_KERNEL_WENDLAND) throw Err(DSPH_E_ARG, "only Wendland kernel is built");
    // 2-D: every particle lies in the plane y = Simulate2DPosY, so every pair
    // distance has dy = 0 and the y row of each pair term is a product with
    // 0; the y component of ace therefore stays 0 (the reference zeroes it,
    // JSphGpuSingle.cpp:646, and the fluid's y velocity stays 0 as gravity.y
    // is the only other source). Linear motions lose their y part
    // (JDsMotion.cpp:132).
    Simulate2D = c.simulate2d != 0;
    Simulate2DPosY = c.simulate2dposy;
    if (Simulate2D) {
      if (c.periactive & 2)
        throw Err(DSPH_E_ARG, "Cannot use periodic conditions in Y with 2D simulations");
      cs.motion.ampl[1] = 0;
    }
    const double h = KernelH;
    KernelSize = float(h * 2.0);
    KernelSize2 = KernelSize * KernelSize;
    // Wendland constants, 2-D or 3-D (FunSphKernel.h:198-209).
    awen = Simulate2D ? float(0.557 / (h * h)) : float(0.41778 / (h * h * h));
    bwenh = Simulate2D ? float(-2.7852 / (h * h * h * h)) : float(-2.08891 / (h * h * h * h * h));
    Cs0 = std::sqrt(double(Gamma) * double(CteB) / double(RhopZero));
    Eta2 = float((h * 0.1) * (h * 0.1));
    if (TVisco == DSPH_VISCO_LAMINARSPS) {
      // dp_sps (JSph.cpp:1529).
      const double dp_sps = Simulate2D ? std::sqrt(double(Dp) * double(Dp) * 2.) / 2.
                                       : std::sqrt(double(Dp) * double(Dp) * 3.) / 3.;
      SpsSmag = float((2. / 3.) * std::pow(0.12 * dp_sps, 2));
      SpsBlin = float((2. / 3.) * 0.0066 * dp_sps * dp_sps);
    }
    DDTkh = KernelSize * float(c.ddtvalue);
    DDTgz = float(double(RhopZero) * double(std::fabs(Gravity[2])) / double(CteB));
    DtIni = c.dtini ? c.dtini : double(KernelH) / Cs0;
    DtMin = c.dtmin ? c.dtmin : (double(KernelH) / Cs0) * double(0.05f);
    CaseNbound = c.casenbound;
    CaseNpb = c.casenfixed + c.casenmoving;
    PeriActive = unsigned(c.periactive) & 7u;
    if (PeriActive && c.casenmoving)
      throw Err(DSPH_E_ARG, "periodic boundaries with moving boundaries are not built");
    if (PeriActive && mdbc2_on() && cs.casenmoving)
      throw Err(DSPH_E_ARG, "periodic boundaries with moving boundaries are not built");
    if (PeriActive && ShiftMode)
      throw Err(DSPH_E_ARG, "periodic boundaries with shifting are not built");
    ft.clear();
    if (c.casenfloat) {
      if (!c.ftcount || !c.floatings) throw Err(DSPH_E_ARG, "floating particles without bodies");
      if (TBoundary != DSPH_BC_DBC || ShiftMode || PeriActive)
        throw Err(DSPH_E_ARG, "floating bodies are built with DBC boundaries, without shifting "
                              "or periodic axes");
      unsigned id = CaseNpb;
      for (unsigned cf = 0; cf < c.ftcount; cf++) {
        const dsph_floating& f = c.floatings[cf];
        if (f.begin != id || !f.count)
          throw Err(DSPH_E_ARG, "floating bodies must tile the floating ids in order");
        id += f.count;
        FtObj o;
        o.begin = f.begin; o.count = f.count;
        o.mass = float(f.massbody); o.massp = float(f.masspart);
        for (int k = 0; k < 3; k++) {
          o.center[k] = f.center[k];
          o.fvel[k] = float(f.velini[k]);
          o.fomega[k] = float(f.omegaini[k]);
        }
        for (int k = 0; k < 9; k++) o.inertia[k] = float(f.inertia[k]);
        ft.push_back(o);
      }
      if (id != CaseNbound) throw Err(DSPH_E_ARG, "floating bodies must tile the floating ids");
      if (ft.size() > 2048) throw Err(DSPH_E_ARG, "more than 2048 floating bodies");
    }
    cs.floatings = nullptr;   // the caller's array is read here only
    for (int k = 0; k < 3; k++) {
      MapRealPosMin[k] = c.maprealposmin[k];
      MapRealPosMax[k] = c.maprealposmax[k];
      MapRealSize[k] = MapRealPosMax[k] - MapRealPosMin[k];
      Map_PosMin[k] = MapRealPosMin[k];
      Map_PosMax[k] = MapRealPosMax[k];
    }
    // LoadCaseParticles (JSph.cpp:2183-2195): the own component of each
    // periodic increment spans the real map; the map grows by KernelSize on
    // both sides of a periodic axis.
    for (int a = 0; a < 3; a++) {
      for (int k = 0; k < 3; k++) PeriInc[a][k] = (PeriActive & (1u << a)) ? c.perinc[a][k] : 0.0;
      if (PeriActive & (1u << a)) {
        PeriInc[a][a] = -MapRealSize[a];
        Map_PosMin[a] = Map_PosMin[a] - KernelSize;
        Map_PosMax[a] = Map_PosMax[a] + KernelSize;
      }
    }
    for (int k = 0; k < 3; k++) Map_Size[k] = Map_PosMax[k] - Map_PosMin[k];
    // ConfigPosCellGpu (JSph.cpp:1949-2014): 13/10/9 bits for PosCell.
    PosCellSize = KernelSize;
    {
      int nks = 1;
      auto ncell = [&](double s) {
        return u3{unsigned(std::ceil(Map_Size[0] / s)), unsigned(std::ceil(Map_Size[1] / s)),
                  unsigned(std::ceil(Map_Size[2] / s))};
      };
      u3 n = ncell(PosCellSize);
      while (n.x > 8192 || n.y > 1024 || n.z > 512) {
        nks = (nks == 1 ? 2 : nks + 2);
        PosCellSize = KernelSize * nks;
        n = ncell(PosCellSize);
      }
    }
    // ConfigCellDivision (JSph.cpp:1877-1895).
    Scell = KernelSize / ScellDiv;
    MovLimit = Scell * 0.9f;
    Map_Cells = u3{unsigned(std::ceil(Map_Size[0] / Scell)), unsigned(std::ceil(Map_Size[1] / Scell)),
                   unsigned(std::ceil(Map_Size[2] / Scell))};
    // SelecDomain(0, Map_Cells) (JSph.cpp:1901-1943).
    DomCells = Map_Cells;
    for (int k = 0; k < 3; k++) DomPosMin[k] = Map_PosMin[k];
    DomCellCode = calc_cell_code(u3{DomCells.x + 1, DomCells.y + 1, DomCells.z + 1});
    if (!DomCellCode) throw Err(DSPH_E_ARG, "no valid cell code for the domain");
  }

  dsph_cte make_cte() const {
    dsph_cte c{};
    c.massb = MassBound; c.massf = MassFluid;
    c.kernelh = KernelH; c.kernelsize2 = KernelSize2; c.poscellsize = PosCellSize;
    c.awen = awen; c.bwenh = bwenh; c.cs0 = float(Cs0); c.eta2 = Eta2;
    c.ddtkh = DDTkh; c.ddtgz = DDTgz; c.scell = Scell; c.kernelsize = KernelSize; c.dp = Dp;
    c.cteb = CteB; c.gamma = Gamma; c.rhopzero = RhopZero; c.ovrhopzero = 1.f / RhopZero;
    c.movlimit = MovLimit; c.tboundary = unsigned(TBoundary); c.cellcode = DomCellCode;
    c.domposminx = DomPosMin[0]; c.domposminy = DomPosMin[1]; c.domposminz = DomPosMin[2];
    c.maprealposminx = MapRealPosMin[0]; c.maprealposminy = MapRealPosMin[1];
    c.maprealposminz = MapRealPosMin[2];
    c.maprealsizex = MapRealSize[0]; c.maprealsizey = MapRealSize[1]; c.maprealsizez = MapRealSize[2];
    c.periactive = PeriActive;
    c.simulate2d = Simulate2D ? 1u : 0u;
    for (int k = 0; k < 3; k++) {
      c.xperinc[k] = PeriInc[0][k]; c.yperinc[k] = PeriInc[1][k]; c.zperinc[k] = PeriInc[2][k];
    }
    return c;
  }

  // ---- particle upload (JSphGpuSingle::ConfigDomain, JSphGpuSingle.cpp:151-232)
  void load(unsigned np, const unsigned* hidp, const double* hpos, const float* hvr,
            const float* hbn) {
    Np = np; Npb = CaseNpb; NpbOk = Npb;
    NpCap = np;
    if (PeriActive) {
      // Room for the periodic copies: until the divide drops them, the
      // copies of the last divide stay beside the new ones, so twice the
      // copies are held; each is a particle within KernelSize of a periodic
      // face (counted at load), given twice that again as headroom.
      size_t near = 0;
      for (unsigned p = 0; p < np; p++)
        for (int a = 0; a < 3; a++)
          if (PeriActive & (1u << a)) {
            const double x = hpos[3 * size_t(p) + a];
            if (x < MapRealPosMin[a] + KernelSize || x >= MapRealPosMax[a] - KernelSize) near++;
          }
      NpCap = unsigned(std::min<size_t>(size_t(np) + 4 * near + 4096, DSPH_NP_LIMIT - 1u));
    }
    std::vector<uint16_t> hcode(np);
    std::vector<unsigned> hdcell(np);
    std::vector<double> hxy(2 * size_t(np)), hz(np);
    for (unsigned p = 0; p < np; p++) {
      const unsigned id = hidp[p];
      // LoadCodeParticles: fixed boundary 0x0000, moving boundary
      // CODE_TYPE_MOVING|1 (mk 1), floating CODE_TYPE_FLOATING|body, fluid
      // CODE_TYPE_FLUID 0x1800.
      uint16_t cd = 0x1800;
      if (id < cs.casenfixed) cd = 0x0000;
      else if (id < CaseNpb) cd = 0x0801;
      else if (id < CaseNbound)
        for (unsigned cf = 0; cf < ft.size(); cf++)
          if (id - ft[cf].begin < ft[cf].count) cd = uint16_t(0x1000u | cf);
      hcode[p] = cd;
      const double x = hpos[3 * size_t(p)], y = hpos[3 * size_t(p) + 1], z = hpos[3 * size_t(p) + 2];
      hxy[2 * size_t(p)] = x; hxy[2 * size_t(p) + 1] = y; hz[p] = z;
      // LoadDcellParticles (JSph.cpp:1794-1816).
      const bool in = x >= MapRealPosMin[0] && y >= MapRealPosMin[1] && z >= MapRealPosMin[2] &&
                      x < MapRealPosMax[0] && y < MapRealPosMax[1] && z < MapRealPosMax[2];
      if (!in) throw Err(DSPH_E_ARG, "Found new particles out.");
      hdcell[p] = dcel_cell(DomCellCode, unsigned((x - DomPosMin[0]) / Scell),
                            unsigned((y - DomPosMin[1]) / Scell), unsigned((z - DomPosMin[2]) / Scell));
    }
    const size_t n = NpCap;
    if (PeriActive) {
      perilist = alloc<unsigned>(n);
      periwork = alloc<unsigned>((n + 255) / 256 + 1);
      acemod = alloc<float>(n);
    }
    idp = alloc<unsigned>(n); idp2 = alloc<unsigned>(n);
    dcell = alloc<unsigned>(n); dcell2 = alloc<unsigned>(n);
    code = alloc<uint16_t>(n); code2 = alloc<uint16_t>(n);
    posxy = alloc<double>(2 * n); posxy2 = alloc<double>(2 * n);
    posz = alloc<double>(n); posz2 = alloc<double>(n);
    velrho = alloc<float>(4 * n); velrho2 = alloc<float>(4 * n);
    poscell = alloc<float>(4 * n);
    ace = alloc<float>(3 * n); ar = alloc<float>(n); viscdt = alloc<float>(n);
    movxy = alloc<double>(2 * n); movz = alloc<double>(n);
    cellpart = alloc<unsigned>(n); sortpart = alloc<unsigned>(n);
    if (TStep == 2) {
      posxypre = alloc<double>(2 * n); posxypre2 = alloc<double>(2 * n);
      poszpre = alloc<double>(n); poszpre2 = alloc<double>(n);
      velrhopre = alloc<float>(4 * n); velrhopre2 = alloc<float>(4 * n);
    } else {
      velrhom1 = alloc<float>(4 * n); velrhom1b = alloc<float>(4 * n);
    }
    if (TVisco == DSPH_VISCO_LAMINARSPS) {
      sps2strain = alloc<float>(6 * n);
      spstau = alloc<float>(6 * n); spstau2 = alloc<float>(6 * n);
    }
    if (ShiftMode) shiftposfs = alloc<float>(4 * n);
    if (TBoundary == DSPH_BC_MDBC) {
      boundnor = alloc<float>(3 * n); boundnor2 = alloc<float>(3 * n);
      if (mdbc2_on()) {
        boundmode = alloc<uint8_t>(n);
        tangenvel = alloc<float>(3 * n);
        if (NoPen) nopenshift = alloc<float>(4 * n);
        if (cs.casenmoving) {
          motionvel = alloc<float>(3 * n); motionvel2 = alloc<float>(3 * n);
          motionace = alloc<float>(3 * n); motionace2 = alloc<float>(3 * n);
        }
      }
    }
    d_red = alloc<float>(4);
    d_dt = alloc<double>(8);
    d_flags = alloc<unsigned>(4);
    if (!ft.empty()) {
      const unsigned nft = unsigned(ft.size());
      d_ftomassp = alloc<float>(nft);
      d_ftbody = alloc<unsigned>(2 * nft);
      d_ftcenter = alloc<double>(3 * nft);
      d_ftace = alloc<float>(6 * nft);
      d_ridp = alloc<unsigned>(cs.casenfloat);
      std::vector<float> mp(nft);
      std::vector<unsigned> body(2 * nft);
      for (unsigned cf = 0; cf < nft; cf++) {
        mp[cf] = ft[cf].massp;
        body[2 * cf] = ft[cf].begin - CaseNpb;
        body[2 * cf + 1] = ft[cf].count;
      }
      chk(dsph_memcpy_h2d(ctx, d_ftomassp, mp.data(), 4 * nft), "upload ftomassp");
      chk(dsph_memcpy_h2d(ctx, d_ftbody, body.data(), 8 * nft), "upload ftbody");
    }
    begincell_cap = 2ull * size_t(DomCells.x) * DomCells.y * DomCells.z + 5;
    begincell = alloc<int>(2 * begincell_cap);
    chk(dsph_ctx_reserve(ctx, NpCap, unsigned(begincell_cap)), "dsph_ctx_reserve");
    const size_t n0 = np;
    chk(dsph_memcpy_h2d(ctx, idp, hidp, 4 * n0), "upload idp");
    chk(dsph_memcpy_h2d(ctx, code, hcode.data(), 2 * n0), "upload code");
    chk(dsph_memcpy_h2d(ctx, dcell, hdcell.data(), 4 * n0), "upload dcell");
    chk(dsph_memcpy_h2d(ctx, posxy, hxy.data(), 16 * n0), "upload posxy");
    chk(dsph_memcpy_h2d(ctx, posz, hz.data(), 8 * n0), "upload posz");
    chk(dsph_memcpy_h2d(ctx, velrho, hvr, 16 * n0), "upload velrho");
    if (velrhom1) chk(dsph_memcpy_h2d(ctx, velrhom1, hvr, 16 * n0), "upload velrhom1");
    if (spstau) chk(dsph_memset_async(ctx, spstau, 0, 24 * n), "memset tau");
    if (motionvel) {
      chk(dsph_memset_async(ctx, motionvel, 0, 12 * n), "memset motionvel");
      chk(dsph_memset_async(ctx, motionace, 0, 12 * n), "memset motionace");
    }
    if (boundnor) {
      // ConfigBoundNormals: normals are stored doubled (JSph.cpp:1385-1434).
      // hbn is indexed by id (boundary ids 0..CaseNbound-1); the particles
      // may come in any order with the boundary block first (a resumed PART
      // keeps the solver's order).
      std::vector<float> bn(3 * n0, 0.f);
      if (hbn)
        for (unsigned p = 0; p < np; p++)
          if (hidp[p] < CaseNbound)
            for (int k = 0; k < 3; k++) bn[3 * size_t(p) + k] = hbn[3 * size_t(hidp[p]) + k] * 2.f;
      chk(dsph_memcpy_h2d(ctx, boundnor, bn.data(), 12 * n0), "upload boundnor");
    }
    std::vector<double> dt0(8, 0.0);
    dt0[2] = DtIni;
    chk(dsph_memcpy_h2d(ctx, d_dt, dt0.data(), 8 * 8), "upload dt");
    unsigned fl[4] = {0, 0, 0, 0};
    chk(dsph_memcpy_h2d(ctx, d_flags, fl, 16), "upload flags");
    const dsph_cte cte = make_cte();
    chk(dsph_set_constants(ctx, &cte), "dsph_set_constants");
    SymplecticDtPre = DtIni;
    BoundChanged = true;
    run_cell_divide();
  }

  // ---- JCellDivGpuSingle::CalcCellDomain / MergeMapCellBoundFluid
  void limits(unsigned n, unsigned pini, u3& cmin, u3& cmax) {
    unsigned mn[3], mx[3];
    chk(dsph_cell_limits_sync(ctx, n, pini, DomCellCode, dcell, code, mn, mx), "cell limits");
    cmin = u3{mn[0], mn[1], mn[2]};
    cmax = u3{mx[0], mx[1], mx[2]};
  }
  // Bound limits over [0,npb1) and the bound copies [npb1+npf1, +npb2);
  // fluid over [npb1, +npf1) and the fluid copies after them
  // (JCellDivGpu::CalcCellDomainBound/Fluid, JCellDivGpu.cpp:250-300).
  void limits2(unsigned n, unsigned pini, unsigned n2, unsigned pini2, u3& cmin, u3& cmax) {
    u3 mn, mx;
    limits(n, pini, mn, mx);
    if (mn.x > mx.x) { mn = DomCells; mx = u3{0, 0, 0}; }
    if (n2) {
      u3 mn2, mx2;
      limits(n2, pini2, mn2, mx2);
      if (mn2.x > mx2.x) { mn2 = DomCells; mx2 = u3{0, 0, 0}; }
      mn = u3{std::min(mn.x, mn2.x), std::min(mn.y, mn2.y), std::min(mn.z, mn2.z)};
      mx = u3{std::max(mx.x, mx2.x), std::max(mx.y, mx2.y), std::max(mx.z, mx2.z)};
    }
    cmin = mn; cmax = mx;
  }
  void calc_cell_domain(unsigned npb1, unsigned npf1, unsigned npb2 = 0, unsigned npf2 = 0) {
    if (CellDomFixed) {
      CellDomainMin = u3{0, 0, 0};
      CellDomainMax = u3{DomCells.x - 1, DomCells.y - 1, DomCells.z - 1};
      return;
    }
    u3 cbmin, cbmax;
    if (!BoundLimitOk) {
      u3 mn, mx;
      limits2(npb1, 0, npb2, npb1 + npf1, mn, mx);
      const bool none = mn.x > mx.x;
      cbmin = none ? DomCells : mn;
      cbmax = none ? u3{0, 0, 0} : mx;
      BoundLimitOk = true; BoundLimitCellMin = cbmin; BoundLimitCellMax = cbmax;
    } else {
      cbmin = BoundLimitCellMin; cbmax = BoundLimitCellMax;
    }
    u3 cfmin, cfmax;
    {
      u3 mn, mx;
      limits2(npf1, npb1, npf2, npb1 + npf1 + npb2, mn, mx);
      const bool none = mn.x > mx.x;
      cfmin = none ? DomCells : mn;
      cfmax = none ? u3{0, 0, 0} : mx;
    }
    const unsigned s = unsigned(ScellDiv);
    auto lo = [&](unsigned b, unsigned f) { return std::max(std::min(b, f), f >= s ? f - s : 0u); };
    auto hi = [&](unsigned b, unsigned f) { return std::min(std::max(b, f), f + s); };
    CellDomainMin = u3{lo(cbmin.x, cfmin.x), lo(cbmin.y, cfmin.y), lo(cbmin.z, cfmin.z)};
    CellDomainMax = u3{hi(cbmax.x, cfmax.x), hi(cbmax.y, cfmax.y), hi(cbmax.z, cfmax.z)};
    if (CellDomainMax.x >= DomCells.x) CellDomainMax.x = DomCells.x - 1;
    if (CellDomainMax.y >= DomCells.y) CellDomainMax.y = DomCells.y - 1;
    if (CellDomainMax.z >= DomCells.z) CellDomainMax.z = DomCells.z - 1;
    if (CellDomainMin.x > CellDomainMax.x || CellDomainMin.y > CellDomainMax.y ||
        CellDomainMin.z > CellDomainMax.z)
      CellDomainMin = CellDomainMax = u3{0, 0, 0};
  }

  unsigned cell_size(unsigned box) {
    int be[2];
    chk(dsph_cell_begin_end_sync(ctx, begincell, box, 1, be), "cell size");
    return unsigned(be[1] - be[0]);
  }

  // ---- JCellDivGpuSingle::Divide (JCellDivGpuSingle.cpp:152-236)
  void divide(unsigned npb1, unsigned npf1, unsigned npb2, unsigned npf2, bool boundchanged) {
    const unsigned nptot = npb1 + npf1 + npb2 + npf2;
    if (boundchanged || PeriActive) {
      BoundLimitOk = BoundDivideOk = false;
      BoundLimitCellMin = BoundLimitCellMax = u3{0, 0, 0};
      BoundDivideCellMin = BoundDivideCellMax = u3{0, 0, 0};
    }
    calc_cell_domain(npb1, npf1, npb2, npf2);
    Ncx = CellDomainMax.x - CellDomainMin.x + 1;
    Ncy = CellDomainMax.y - CellDomainMin.y + 1;
    Ncz = CellDomainMax.z - CellDomainMin.z + 1;
    Nsheet = Ncx * Ncy;
    Nct = Nsheet * Ncz;
    Nctt = 2 * Nct + 5;
    if (size_t(Nctt) > begincell_cap) throw Err(DSPH_E_NOMEM, "begincell capacity exceeded");
    BoxBoundIgnore = Nct;
    BoxFluid = BoxBoundIgnore + 1;
    BoxBoundOut = BoxFluid + Nct;
    if (!BoundDivideOk || BoundDivideCellMin != CellDomainMin || BoundDivideCellMax != CellDomainMax) {
      DivideFull = true;
      BoundDivideOk = true; BoundDivideCellMin = CellDomainMin; BoundDivideCellMax = CellDomainMax;
      NdivFull++;
    } else {
      DivideFull = false;
    }
    const unsigned cmin[3] = {CellDomainMin.x, CellDomainMin.y, CellDomainMin.z};
    const unsigned ncel[3] = {Ncx, Ncy, Ncz};
    if (DivideFull)
      chk(dsph_presort_full(ctx, nptot, DomCellCode, dcell, code, cmin, ncel, cellpart, sortpart),
          "presort full");
    else
      chk(dsph_presort_fluid(ctx, nptot - npb1, npb1, DomCellCode, dcell, code, cmin, ncel,
                             cellpart, sortpart),
          "presort fluid");
    if (DivideFull)
      chk(dsph_sort_by_cell(ctx, cellpart, sortpart, nptot, Nctt), "sort");
    else
      chk(dsph_sort_by_cell(ctx, cellpart + npb1, sortpart + npb1, nptot - npb1, Nctt), "sort");
    chk(dsph_begin_end_cell(ctx, DivideFull ? 1 : 0, nptot, npb1, Nctt, BoxFluid, cellpart, begincell),
        "begin end cell");
    NpbIgnore = cell_size(BoxBoundIgnore);
    int be[8];
    chk(dsph_cell_begin_end_sync(ctx, begincell, BoxBoundOut, 4, be), "out boxes");
    NpbOut = unsigned(be[1] - be[0]);
    NpfOut = unsigned(be[3] - be[2]);
    const unsigned npbOutIgnore = unsigned(be[5] - be[4]);
    const unsigned npfOutIgnore = unsigned(be[7] - be[6]);
    NpFinal = nptot - NpbOut - NpfOut - npbOutIgnore - npfOutIgnore;
    NpbFinal = npb1 + npb2 - npbOutIgnore;
    if (NpbOut != 0 && DivideFull) NpbFinal = UINT_MAX;
    Ndiv++;
  }

  dsph_divdata divdata() const {
    dsph_divdata d{};
    d.scelldiv = ScellDiv;
    d.ncx = int(Ncx); d.ncy = int(Ncy); d.ncz = int(Ncz); d.ncw = int(Ncx * Ncy);
    d.cellfluid = Ncx * Ncy * Ncz + 1;
    d.cellzerox = int(CellDomainMin.x); d.cellzeroy = int(CellDomainMin.y); d.cellzeroz = int(CellDomainMin.z);
    d.d_beginendcell = begincell;
    d.scell = Scell;
    d.domcellcode = DomCellCode;
    d.domposminx = DomPosMin[0]; d.domposminy = DomPosMin[1]; d.domposminz = DomPosMin[2];
    d.kernelsize2 = KernelSize2;
    d.poscellsize = PosCellSize;
    return d;
  }

  template <typename T>
  static void swp(T*& a, T*& b) { std::swap(a, b); }

  // ---- JSphGpuSingle::RunCellDivide (JSphGpuSingle.cpp:372-499)
  // ---- JSphGpuSingle::RunPeriodic (JSphGpuSingle.cpp:276-366): the copies
  // of the last divide are marked to be ignored and new ones are appended
  // at Np, boundary copies first; for each periodic axis the copies made
  // for earlier axes are examined before the original particles, so that
  // corners get copies too.
  void run_periodic() {
    chk(dsph_periodic_ignore(ctx, Np, code), "periodic ignore");
    const unsigned npb0 = Npb, npf0 = Np - Npb;
    NpbPer = NpfPer = 0;
    BoundChanged = true;
    for (unsigned ctype = 0; ctype < 2; ctype++) {
      const unsigned pini = ctype ? npb0 : 0u, num = ctype ? npf0 : npb0;
      for (unsigned a = 0; a < 3; a++) {
        if (!(PeriActive & (1u << a))) continue;
        for (unsigned cblock = 0; cblock < 2; cblock++) {
          const unsigned nper = ctype ? NpfPer : NpbPer;
          const unsigned pini2 = cblock ? pini : Np - nper, num2 = cblock ? num : nper;
          if (!num2) continue;
          unsigned nmax = NpCap - Np;
          unsigned count = 0;
          chk(dsph_periodic_make_list(ctx, num2, pini2, nmax, Map_PosMin, Map_PosMax, PeriInc[a],
                                      posxy, posz, code, periwork, perilist, &count),
              "periodic list");
          if (count > nmax) {
            // More copies than room: grow every particle array (contents of
            // [0,Np) kept, as JSphGpuSingle::ResizeParticlesSize) and make
            // the list again; nothing has changed the state of this block yet.
            grow_particles(size_t(Np) + count + (size_t(Np) + count) / 4 + 4096);
            nmax = NpCap - Np;
            chk(dsph_periodic_make_list(ctx, num2, pini2, nmax, Map_PosMin, Map_PosMax,
                                        PeriInc[a], posxy, posz, code, periwork, perilist, &count),
                "periodic list");
            if (count > nmax) throw Err(DSPH_E_STATE, "periodic list changed after growing");
          }
          chk(dsph_periodic_duplicate(ctx, count, Np, &DomCells.x, PeriInc[a], perilist, idp, code,
                                      dcell, posxy, posz, velrho, spstau,
                                      pre_active ? posxypre : nullptr,
                                      pre_active ? poszpre : nullptr,
                                      pre_active ? velrhopre : nullptr,
                                      TStep == 1 ? velrhom1 : nullptr, boundnor),
              "periodic duplicate");
          Np += count;
          if (!ctype) NpbPer += count; else NpfPer += count;
        }
      }
    }
  }

  // Reallocates every per-particle array for `cap` particles, keeping the
  // contents of [0,Np) (JSphGpuSingle::ResizeParticlesSize,
  // JSphGpuSingle.cpp:240-262); the context scratch follows.
  void grow_particles(size_t cap) {
    cap = std::min<size_t>(cap, DSPH_NP_LIMIT - 1u);
    if (cap <= NpCap) throw Err(DSPH_E_NOMEM, "particle arrays at DSPH_NP_LIMIT");
    chk(dsph_synchronize(ctx), "synchronize");
    const size_t keep = Np;
    auto regrow = [&](auto*& ptr, size_t per) {
      using T = std::remove_reference_t<decltype(*ptr)>;
      if (!ptr) return;
      T* q = alloc<T>(per * cap);
      chk(dsph_memcpy_d2d_async(ctx, q, ptr, sizeof(T) * per * keep), "grow copy");
      chk(dsph_synchronize(ctx), "synchronize");
      chk(dsph_free(ctx, ptr), "grow free");
      ptr = q;
    };
    regrow(perilist, 1); regrow(acemod, 1);
    regrow(idp, 1); regrow(idp2, 1); regrow(dcell, 1); regrow(dcell2, 1);
    regrow(code, 1); regrow(code2, 1);
    regrow(posxy, 2); regrow(posxy2, 2); regrow(posz, 1); regrow(posz2, 1);
    regrow(velrho, 4); regrow(velrho2, 4); regrow(poscell, 4);
    regrow(ace, 3); regrow(ar, 1); regrow(viscdt, 1); regrow(movxy, 2); regrow(movz, 1);
    regrow(cellpart, 1); regrow(sortpart, 1);
    regrow(posxypre, 2); regrow(posxypre2, 2); regrow(poszpre, 1); regrow(poszpre2, 1);
    regrow(velrhopre, 4); regrow(velrhopre2, 4); regrow(velrhom1, 4); regrow(velrhom1b, 4);
    regrow(sps2strain, 6); regrow(spstau, 6); regrow(spstau2, 6); regrow(shiftposfs, 4);
    regrow(boundnor, 3); regrow(boundnor2, 3); regrow(boundmode, 1); regrow(tangenvel, 3);
    regrow(nopenshift, 4);
    regrow(motionvel, 3); regrow(motionvel2, 3); regrow(motionace, 3); regrow(motionace2, 3);
    regrow(velgrad, 9); regrow(closure_nu, 1);
    if (periwork) {
      chk(dsph_free(ctx, periwork), "grow free");
      periwork = alloc<unsigned>((cap + 255) / 256 + 1);
    }
    NpCap = unsigned(cap);
    chk(dsph_ctx_reserve(ctx, NpCap, unsigned(begincell_cap)), "dsph_ctx_reserve");
  }

  void run_cell_divide() {
    if (PeriActive) run_periodic();
    const unsigned nptot = Np;
    divide(Npb, Np - Npb - NpbPer - NpfPer, NpbPer, NpfPer, BoundChanged);
    const unsigned pini = DivideFull ? 0u : NpbFinal;
    chk(dsph_sort_data_basic_poscell(ctx, nptot, pini, sortpart, idp, code, dcell, posxy, posz,
                                     velrho, idp2, code2, dcell2, posxy2, posz2, velrho2,
                                     Map_PosMin, PosCellSize, poscell),
        "sort basic");
    swp(idp, idp2); swp(code, code2); swp(dcell, dcell2); swp(posxy, posxy2); swp(posz, posz2);
    swp(velrho, velrho2);
    if (TStep == 1) {
      chk(dsph_sort_data_f4(ctx, nptot, pini, sortpart, velrhom1, velrhom1b), "sort velrhom1");
      swp(velrhom1, velrhom1b);
    } else if (pre_active) {
      chk(dsph_sort_data_d2_d_f4(ctx, nptot, pini, sortpart, posxypre, poszpre, velrhopre,
                                 posxypre2, poszpre2, velrhopre2),
          "sort pre");
      swp(posxypre, posxypre2); swp(poszpre, poszpre2); swp(velrhopre, velrhopre2);
    }
    if (TVisco == DSPH_VISCO_LAMINARSPS) {
      chk(dsph_sort_data_sym6(ctx, nptot, pini, sortpart, spstau, spstau2), "sort tau");
      swp(spstau, spstau2);
    }
    if (TBoundary == DSPH_BC_MDBC) {
      chk(dsph_sort_data_f3(ctx, nptot, pini, sortpart, boundnor, boundnor2), "sort boundnor");
      swp(boundnor, boundnor2);
    }
    if (motionvel) {
      chk(dsph_sort_data_f3(ctx, nptot, pini, sortpart, motionvel, motionvel2), "sort motionvel");
      swp(motionvel, motionvel2);
      chk(dsph_sort_data_f3(ctx, nptot, pini, sortpart, motionace, motionace2), "sort motionace");
      swp(motionace, motionace2);
    }
    if (NpbOut) throw Err(DSPH_E_NUMERIC, "boundary particles excluded (AbortBoundOut)");
    Np = NpFinal;
    Npb = NpbFinal;
    NpbOk = Npb - NpbIgnore;
    NpfOutTotal += NpfOut;
    BoundChanged = false;
  }

  // ---- JSphGpu::PreInteraction_Forces (JSphGpu.cpp:846-913)
  void pre_interaction() {
    chk(dsph_memset_async(ctx, viscdt, 0, 4 * size_t(Np)), "memset viscdt");
    chk(dsph_memset_async(ctx, ar, 0, 4 * size_t(Np)), "memset ar");
    chk(dsph_memset_async(ctx, ace, 0, 12 * size_t(Np)), "memset ace");
    if (sps2strain)
      chk(dsph_memset_async(ctx, sps2strain + 6 * size_t(Npb), 0, 24 * size_t(Np - Npb)), "memset strain");
    if (shiftposfs)
      chk(dsph_memset_async(ctx, shiftposfs + 4 * size_t(Npb), 0, 16 * size_t(Np - Npb)),
          "memset shiftposfs");
    chk(dsph_velmax2(ctx, Np - Npb, Npb, velrho, d_red + 0), "velmax");
  }

  // ---- JSphGpuSingle::MdbcBoundCorrection (JSphGpuSingle.cpp:686-712)
  // SLIP_Vel0 corrects before the predictor only (MdbcCorrector=false);
  // no-slip before both interactions (MdbcCorrector=true, JSph.cpp:897).
  void mdbc(bool corrector) {
    if (TBoundary != DSPH_BC_MDBC) return;
    const dsph_divdata d = divdata();
    if (mdbc2_on()) {
      chk(dsph_memset_async(ctx, boundmode, DSPH_BMODE_DBC, Npb), "memset boundmode");
      chk(dsph_mdbc2_correction_slip(ctx, SlipMode, Npb, Gravity, &d, Map_PosMin, posxy, posz,
                                     poscell, code, boundnor, motionvel, motionace, velrho,
                                     boundmode, tangenvel),
          "mdbc2");
      return;
    }
    if (corrector) return;
    chk(dsph_mdbc_correction(ctx, NpbOk, &d, Map_PosMin, posxy, posz, poscell, code, boundnor, velrho),
        "mdbc");
  }

  // ---- JSphGpuSingle::Interaction_Forces (JSphGpuSingle.cpp:590-680)
  void interaction() {
    dsph_inter_parms t{};
    t.tkernel = DSPH_KERNEL_WENDLAND;
    t.tvisco = TVisco;
    t.tdensity = TDensity;
    t.viscob = Visco * ViscoBoundFactor;
    t.viscof = Visco;
    t.np = Np; t.npb = Npb; t.npbok = NpbOk;
    t.divdata = divdata();
    t.d_dcell = dcell; t.d_poscell = poscell; t.d_velrho = velrho; t.d_code = code;
    t.d_spstaurho2 = spstau;
    t.d_viscdt = viscdt; t.d_ar = ar; t.d_ace = ace; t.d_sps2strain = sps2strain;
    t.d_velgrad = UseClosure ? velgrad : nullptr;
    t.d_boundmode = boundmode;
    t.d_tangenvel = tangenvel;
    t.shiftmode = ShiftMode;
    t.d_shiftposfs = shiftposfs;
    t.d_ftomassp = d_ftomassp;
    t.d_boundnor = NoPen ? boundnor : nullptr;
    t.d_motionvel = NoPen ? motionvel : nullptr;
    t.d_nopenshift = nopenshift;
    chk(dsph_interaction_forces(ctx, &t), "interaction");
    if (TVisco == DSPH_VISCO_LAMINARSPS) {
      if (UseClosure)
        chk(dsph_sps_tau_learned(ctx, &Closure, Np, Npb, SpsBlin, float(Dp), velrho, velgrad,
                                 sps2strain, spstau, closure_nu),
            "learned sps tau");
      else
        chk(dsph_sps_tau(ctx, Np, Npb, SpsSmag, SpsBlin, velrho, sps2strain, spstau), "sps tau");
    }
    chk(dsph_redu_max_f32(ctx, Np, 0, viscdt, d_red + 1), "viscdtmax");
    if (PeriActive)
      chk(dsph_acemax2_normal(ctx, Np - Npb, Npb, ace, code, acemod, d_red + 2), "acemax");
    else
      chk(dsph_acemax2(ctx, Np - Npb, Npb, ace, d_red + 2), "acemax");
  }

  // ---- JSphGpu::DtVariable (JSphGpu.cpp:1103-1147)
  double dt_variable(unsigned slot) {
    if (device_dt) {
      chk(dsph_dt_variable(ctx, d_red, double(KernelH), Cs0, CFLnumber, DtMin, d_dt, slot, d_flags),
          "dt");
      return 0;
    }
    float red[4];
    chk(dsph_memcpy_d2h(ctx, red, d_red, 12), "dt readback");
    VelMax = double(std::sqrt(red[0]));
    ViscDtMax = red[1];
    AceMax = std::sqrt(double(red[2]));
    const double dt1 = (AceMax ? std::sqrt(double(KernelH) / AceMax) : DBL_MAX);
    const double dt2 = double(KernelH) / (std::max(Cs0, VelMax * 10.) + double(KernelH) * ViscDtMax);
    double dt = CFLnumber * std::min(dt1, dt2);
    if (std::isnan(dt) || std::isinf(dt)) throw Err(DSPH_E_NUMERIC, "dt is NaN or infinity");
    if (dt < DtMin) dt = DtMin;
    return dt;
  }

  void set_host_dt_slot(unsigned slot, double v) {
    chk(dsph_memcpy_h2d(ctx, d_dt + slot, &v, 8), "dt upload");
  }

  // ---- JSphGpu::RunShifting (JSphGpu.cpp:1153-1157) over [t, t + dt*scale]:
  // the shifting displacement of the fluid from the interaction's sums.
  void run_shifting(double dt, double scale) {
    if (!shiftposfs) return;
    chk(dsph_run_shifting(ctx, Np - Npb, Npb, dt * scale, device_dt ? d_dt : nullptr, 2, scale,
                          double(float(ShiftCoef)), double(KernelH), float(ShiftTFS), cs.dp, velrho,
                          shiftposfs),
        "shifting");
  }

  // ---- JSphGpu::ComputeSymplecticPre / Corr (JSphGpu.cpp:996-1097)
  void symplectic_pre(double dt) {
    swp(posxypre, posxy); swp(poszpre, posz); swp(velrhopre, velrho);
    pre_active = true;
    chk(dsph_step_symplectic_pre(ctx, Np, Npb, velrhopre, boundmode, ar, ace, shiftposfs, dt * .5,
                                 device_dt ? d_dt : nullptr, 2, RhopZero, RhopOutMin, RhopOutMax,
                                 Gravity, code, movxy, movz, velrho),
        "symplectic pre");
    chk(dsph_step_pos2(ctx, Np, Npb, posxypre, poszpre, movxy, movz, posxy, posz, dcell, code), "pos2");
    chk(dsph_memcpy_d2d_async(ctx, posxy, posxypre, 16 * size_t(Npb)), "copy bound xy");
    chk(dsph_memcpy_d2d_async(ctx, posz, poszpre, 8 * size_t(Npb)), "copy bound z");
  }
  void symplectic_cor(double dt) {
    chk(dsph_step_symplectic_cor(ctx, Np, Npb, velrhopre, boundmode, ar, ace, shiftposfs, dt * .5, dt,
                                 device_dt ? d_dt : nullptr, 2, RhopZero, RhopOutMin, RhopOutMax,
                                 Gravity, code, movxy, movz, velrho),
        "symplectic cor");
    if (NoPen)
      chk(dsph_mdbc2_nopen_step(ctx, Np, Npb, nopenshift, velrhopre, dt, device_dt ? d_dt : nullptr,
                                2, velrho, movxy, movz),
          "no penetration");
    chk(dsph_step_pos2(ctx, Np, Npb, posxypre, poszpre, movxy, movz, posxy, posz, dcell, code), "pos2");
    pre_active = false;
  }

  // ---- JSphGpu::RunMotion (JSphGpu.cpp:1173-1229) for one moving object:
  // after the step it is moved over [t, t+dt], by a displacement
  // (rectilinear sinusoid, MoveLinBound) or a matrix (rotations,
  // MoveMatBound, which also turns the mDBC normals), and takes the velocity
  // displacement/dt; the next divide is full (BoundChanged).
  void run_motion(double dt) {
    const dsph_motion& m = cs.motion;
    if (!cs.casenmoving || m.type == 0) return;
    move_boundaries(dt);
    // mDBC2: motion velocity and acceleration of the moving particles,
    // whether the motion is active or not (JSphGpu.cpp:1219-1223).
    if (motionvel)
      chk(dsph_mdbc2_motion_vel_ace(ctx, Npb, m.idbegin, m.count, dt, device_dt ? d_dt : nullptr,
                                    2, idp, velrho, motionvel, motionace),
          "motion vel ace");
  }
  void move_boundaries(double dt) {
    const dsph_motion& m = cs.motion;
    if (device_dt) {
      chk(dsph_move_motion_device(ctx, Npb, &m, 0.0, d_dt, idp, posxy, posz, dcell, velrho,
                                  code, boundnor),
          "motion");
    } else {
      if (!(TimeStep >= m.start && TimeStep < m.finish)) return;
      if (m.type == 1) {
        double mv[3];
        dsph::rect_sinu_mov_host(&m, TimeStep, dt, mv);
        const float vel[3] = {float(mv[0] / dt), float(mv[1] / dt), float(mv[2] / dt)};
        chk(dsph_move_lin_bound(ctx, Npb, m.idbegin, m.count, mv, vel, idp, posxy, posz, dcell,
                                velrho, code),
            "motion");
      } else {
        double mat[12];
        chk(dsph_motion_matrix(&m, TimeStep, dt, mat), "motion matrix");
        chk(dsph_move_mat_bound(ctx, Npb, m.idbegin, m.count, mat, dt, idp, posxy, posz, dcell,
                                velrho, code, boundnor),
            "motion");
      }
    }
    BoundChanged = true;
  }

  // ---- JSphGpuSingle::RunFloating (JSphGpuSingle.cpp:827-903) without
  // Chrono, moorings, external forces, imposed velocities, constraints,
  // periodic axes or mDBC normals, FtPause 0: the particle sums on the
  // device, the rigid-body update of each body on the host (JSph::
  // FtComputeAceVel, JSph.cpp:2565-2660), the particles moved on the device,
  // and after the corrector JSph::FtUpdateFloatings (JSph.cpp:2689-2704).
  void ft_compute_ace_vel(unsigned cf, double dt, const float* s6, float fvel[3],
                          float fomega[3], double fcenter[3]) const {
    const FtObj& f = ft[cf];
    float fforcelin[3], fforceang[3], acelin[3], aceang[3];
    for (int k = 0; k < 3; k++) {
      fforcelin[k] = s6[k] * f.massp;
      fforceang[k] = s6[3 + k] * f.massp;
      acelin[k] = (fforcelin[k] + Gravity[k] * f.mass) / f.mass;
    }
    // fmath::RotMatrix3x3 (FunctionsMath.h:343-352); I = (R I0) R^T; inverse
    // (FunctionsMath.h:132-150).
    const float cx = std::cos(f.angles[0]), cy = std::cos(f.angles[1]), cz = std::cos(f.angles[2]);
    const float sx = std::sin(f.angles[0]), sy = std::sin(f.angles[1]), sz = std::sin(f.angles[2]);
    const float r[9] = {cy * cz, -cy * sz, sy,
                        sx * sy * cz + cx * sz, -sx * sy * sz + cx * cz, -sx * cy,
                        -cx * sy * cz + sx * sz, cx * sy * sz + sx * cz, cx * cy};
    auto mul = [](const float* a, const float* b, float* o) {
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
          o[3 * i + j] = a[3 * i] * b[j] + a[3 * i + 1] * b[3 + j] + a[3 * i + 2] * b[6 + j];
    };
    float ri[9], rt[9], in[9];
    mul(r, f.inertia, ri);
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) rt[3 * i + j] = r[3 * j + i];
    mul(ri, rt, in);
    const float* d = in;
    const float det = d[0] * d[4] * d[8] + d[1] * d[5] * d[6] + d[2] * d[3] * d[7] -
                      d[6] * d[4] * d[2] - d[7] * d[5] * d[0] - d[8] * d[3] * d[1];
    float inv[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (det) {
      inv[0] = (d[4] * d[8] - d[5] * d[7]) / det;
      inv[1] = -(d[1] * d[8] - d[2] * d[7]) / det;
      inv[2] = (d[1] * d[5] - d[2] * d[4]) / det;
      inv[3] = -(d[3] * d[8] - d[5] * d[6]) / det;
      inv[4] = (d[0] * d[8] - d[2] * d[6]) / det;
      inv[5] = -(d[0] * d[5] - d[2] * d[3]) / det;
      inv[6] = (d[3] * d[7] - d[4] * d[6]) / det;
      inv[7] = -(d[0] * d[7] - d[1] * d[6]) / det;
      inv[8] = (d[0] * d[4] - d[1] * d[3]) / det;
    }
    for (int i = 0; i < 3; i++)
      aceang[i] = fforceang[0] * inv[3 * i] + fforceang[1] * inv[3 * i + 1] + fforceang[2] * inv[3 * i + 2];
    for (int k = 0; k < 3; k++) {
      fcenter[k] = f.center[k] + dt * f.fvel[k];
      fvel[k] = float(dt * acelin[k] + f.fvel[k]);
      fomega[k] = float(dt * aceang[k] + f.fomega[k]);
    }
    if (Simulate2D) {  // in-plane motion only (JSph.cpp:2628-2640)
      fcenter[1] = f.center[1];
      fvel[1] = 0;
      fomega[0] = fomega[2] = 0;
    }
  }
  void run_floating(double dt, bool predictor) {
    if (ft.empty()) return;
    const unsigned nft = unsigned(ft.size());
    chk(dsph_memset_async(ctx, d_ridp, 0xff, 4 * size_t(cs.casenfloat)), "clear ft ridp");
    chk(dsph_ft_calc_ridp(ctx, Np - Npb, Npb, CaseNpb, CaseNbound, idp, d_ridp), "ft ridp");
    std::vector<double> cen(3 * nft);
    for (unsigned cf = 0; cf < nft; cf++)
      for (int k = 0; k < 3; k++) cen[3 * cf + k] = ft[cf].center[k];
    chk(dsph_memcpy_h2d(ctx, d_ftcenter, cen.data(), 24 * nft), "upload ft centre");
    chk(dsph_ft_parts_sum_ace(ctx, nft, d_ftbody, d_ftcenter, d_ridp, posxy, posz, ace, d_ftace),
        "ft sum ace");
    std::vector<float> s6(6 * nft);
    chk(dsph_memcpy_d2h(ctx, s6.data(), d_ftace, 24 * nft), "ft sums readback");
    for (unsigned cf = 0; cf < nft; cf++) {
      float fvel[3], fomega[3];
      double fcenter[3];
      ft_compute_ace_vel(cf, dt, &s6[6 * cf], fvel, fomega, fcenter);
      chk(dsph_ft_parts_update(ctx, dt, ft[cf].count, ft[cf].begin - CaseNpb, fvel, fomega,
                               fcenter, d_ridp, posxy, posz, velrho, dcell, code),
          "ft update");
      if (!predictor) {
        FtObj& f = ft[cf];
        for (int k = 0; k < 3; k++) {
          f.center[k] = fcenter[k];
          f.angles[k] = float(double(f.angles[k]) + double(fomega[k]) * dt);
          f.facelin[k] = (fvel[k] - f.fvel[k]) / float(dt);
          f.faceang[k] = (fomega[k] - f.fomega[k]) / float(dt);
          f.fvel[k] = fvel[k];
          f.fomega[k] = fomega[k];
        }
      }
    }
  }

  // ---- JSphGpuSingle::ComputeStep_Sym (JSphGpuSingle.cpp:785-821)
  void step_sym() {
    const double dt = SymplecticDtPre;
    mdbc(false);
    pre_interaction();
    interaction();
    const double dt_p = dt_variable(0);
    run_shifting(dt, 0.5);
    symplectic_pre(dt);
    run_floating(dt * .5, true);
    run_cell_divide();
    mdbc(true);
    pre_interaction();
    interaction();
    const double dt_c = dt_variable(1);
    run_shifting(dt, 1.0);
    symplectic_cor(dt);
    run_floating(dt, false);
    run_motion(dt);
    if (device_dt) {
      chk(dsph_dt_variable_min_advance(), "dt advance");
    } else {
      SymplecticDtPre = std::min(dt_p, dt_c);
      TimeStep += dt;
      DtLast = dt;
    }
  }
  int dsph_dt_variable_min_advance();

  // ---- JSphGpuSingle::ComputeStep_Ver (JSphGpuSingle.cpp:761-781) + JSphGpu::ComputeVerlet
  void step_ver() {
    mdbc(false);
    pre_interaction();
    interaction();
    device_dt = false;
    const double dt = dt_variable(0);
    run_shifting(dt, 1.0);
    VerletStep++;
    if (VerletStep < unsigned(VerletSteps)) {
      chk(dsph_step_verlet(ctx, Np, Npb, velrho, velrhom1, boundmode, ar, ace, shiftposfs, dt, dt + dt, RhopZero,
                           RhopOutMin, RhopOutMax, Gravity, code, movxy, movz, velrhom1),
          "verlet");
    } else {
      chk(dsph_step_verlet(ctx, Np, Npb, velrho, velrho, boundmode, ar, ace, shiftposfs, dt, dt, RhopZero, RhopOutMin,
                           RhopOutMax, Gravity, code, movxy, movz, velrhom1),
          "verlet euler");
      VerletStep = 0;
    }
    if (NoPen)
      chk(dsph_mdbc2_nopen_step(ctx, Np, Npb, nopenshift, velrho, dt, nullptr, 0, velrhom1, movxy,
                                movz),
          "no penetration");
    chk(dsph_step_pos(ctx, Np, Npb, movxy, movz, posxy, posz, dcell, code), "pos");
    swp(velrho, velrhom1);
    run_floating(dt, false);
    run_motion(dt);
    TimeStep += dt;
    DtLast = dt;
  }

  void run(unsigned nsteps, bool sync_dt) {
    // Floating bodies need dt on the host (their update runs there).
    device_dt = (!sync_dt && TStep == 2 && ft.empty());
    if (device_dt) {
      // The device carries the absolute time so that it is summed step by
      // step as on the host (time += dt), whatever the batching of runs.
      set_host_dt_slot(2, SymplecticDtPre);
      set_host_dt_slot(3, TimeStep);
    }
    for (unsigned s = 0; s < nsteps; s++) {
      if (TStep == 2) step_sym(); else step_ver();
      run_cell_divide();
      Nstep++;
    }
    if (device_dt) {
      double h[8];
      chk(dsph_memcpy_d2h(ctx, h, d_dt, 64), "dt readback");
      unsigned fl[4];
      chk(dsph_memcpy_d2h(ctx, fl, d_flags, 16), "flags readback");
      if (fl[0] & 1u) throw Err(DSPH_E_NUMERIC, "dt is NaN or infinity");
      SymplecticDtPre = h[2];
      TimeStep = h[3];
      DtLast = h[4];
    }
    chk(dsph_synchronize(ctx), "synchronize");
  }
};
