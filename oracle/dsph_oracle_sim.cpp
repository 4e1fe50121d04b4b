// Oracle whole-step driver: the reference CPU solver flow (JSphCpuSingle::
// ComputeStep_Sym / ComputeStep_Ver + RunCellDivide, JSphCpuSingle.cpp:
// 510-587, 685-932; JCellDivCpuSingle::Divide, JCellDivCpuSingle.cpp:280-350)
// over the oracle functions of dsph_oracle.cpp.
//
// TEST INFRASTRUCTURE ONLY (see dsph_oracle.cpp). Used as the end-to-end
// parity checker and, timed on the host cores, as the CPU baseline.
//
// Pair semantics follow the GPU path (SEM_GPU, FP32) unless the simulation is
// created with prec=1 (FP64 yardstick with FP64 distances) or prec=2 (FP32
// with the reference CPU build's semantics, SEM_CPU).
#include <omp.h>
#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>
#include "dsph_abi.h"

typedef uint16_t typecode;

extern "C" {
struct orc_cte {
  float massb, massf, kernelh, kernelsize2, poscellsize, awen, bwenh, cs0, eta2, ddtkh, ddtgz;
  float scell, kernelsize, dp, cteb, gamma, rhopzero, ovrhopzero, movlimit;
  unsigned cellcode;
  double domposmin[3], maprealposmin[3], maprealsize[3];
  double mapposmin[3];
  unsigned tboundary;
  unsigned periactive;        // periodic axes, bits 1 X, 2 Y, 4 Z
  double perinc[3][3];         // PeriXinc, PeriYinc, PeriZinc
  unsigned sim2d;              // 2-D simulation
};
struct orc_div {
  int scelldiv, ncx, ncy, ncz, ncw;
  unsigned cellfluid;
  int czx, czy, czz;
  const int* begincell;
};
unsigned orc_cell_code(unsigned, unsigned, unsigned);
void orc_limits(unsigned, unsigned, unsigned, const unsigned*, const typecode*, unsigned*, unsigned*);
void orc_presort(int, unsigned, unsigned, unsigned, const unsigned*, const typecode*,
                 const unsigned*, const unsigned*, unsigned*);
void orc_sort_cells(unsigned, unsigned, unsigned, unsigned, unsigned*, unsigned*, int*);
void orc_poscell(unsigned, const double*, const double*, float, float*);
void orc_interaction(int, int, int, int, float, float, const orc_cte*, const orc_div*, unsigned,
                     unsigned, unsigned, const unsigned*, const float*, const double*,
                     const float*, const float*, float*, float*, float*, float*, float*);
void orc_sps_tau(unsigned, unsigned, float, float, const float*, const float*, float*);
void orc_sym_pre(unsigned, unsigned, const float*, const float*, const float*, double, float, float,
                 float, const float*, typecode*, double*, float*);
void orc_sym_cor(unsigned, unsigned, const float*, const float*, const float*, double, double,
                 float, float, float, const float*, typecode*, double*, float*);
void orc_verlet(unsigned, unsigned, const float*, const float*, const float*, const float*, double,
                double, float, float, float, const float*, typecode*, double*, float*);
void orc_step_pos(unsigned, unsigned, const orc_cte*, const double*, const double*, double*,
                  unsigned*, typecode*);
void orc_mdbc(int, int, const orc_cte*, const orc_div*, unsigned, const double*, const float*,
              const typecode*, const float*, float*);
void orc_interaction_mdbc2(int, int, int, int, float, float, const orc_cte*, const orc_div*,
                           unsigned, unsigned, unsigned, const unsigned*, const float*,
                           const double*, const float*, const float*, float*, float*, float*,
                           float*, float*, const unsigned char*, const float*);
void orc_mdbc2(int, int, const orc_cte*, const orc_div*, unsigned, const double*, const float*,
               const typecode*, const float*, const float*, const float*, const float*, double,
               double, float*, unsigned char*, float*);
void orc_mdbc2_keep(unsigned, const unsigned char*, const float*, float*);
void orc_mdbc2_slip(int, int, int, const orc_cte*, const orc_div*, unsigned, const double*,
                    const float*, const typecode*, const float*, const float*, const float*,
                    const float*, double, double, float*, unsigned char*, float*);
void orc_interaction_np(int, int, int, int, float, float, const orc_cte*, const orc_div*,
                        unsigned, unsigned, unsigned, const unsigned*, const float*,
                        const double*, const float*, const float*, float*, float*, float*,
                        float*, float*, const unsigned char*, const float*, int,
                        const typecode*, float*, const float*, const float*, float*);
void orc_nopen_step(unsigned, unsigned, const float*, const float*, double, float*, double*);
void orc_interaction_range_mv(int, int, int, int, float, float, const orc_cte*, const orc_div*,
                              unsigned, unsigned, unsigned, const unsigned*, const float*,
                              const double*, const float*, const float*, float*, float*, float*,
                              float*, float*, const unsigned char*, const float*, int,
                              const typecode*, float*, const float*, const float*, float*,
                              const float*, long, long, long, long);
void orc_motion_vel_ace(unsigned, unsigned, unsigned, double, const unsigned*, const float*,
                        float*, float*);
void orc_interaction_ext(int, int, int, int, float, float, const orc_cte*, const orc_div*,
                         unsigned, unsigned, unsigned, const unsigned*, const float*,
                         const double*, const float*, const float*, float*, float*, float*,
                         float*, float*, const unsigned char*, const float*, int,
                         const typecode*, float*);
void orc_run_shifting(unsigned, unsigned, double, float, float, float, double, const float*,
                      float*, unsigned);
void orc_add_shift(unsigned, unsigned, const float*, double*);
void orc_interaction_ft(int, int, int, int, float, float, const orc_cte*, const orc_div*,
                        unsigned, unsigned, unsigned, const unsigned*, const float*,
                        const double*, const float*, const float*, float*, float*, float*,
                        float*, float*, const unsigned char*, const float*, int,
                        const typecode*, float*, const float*);
void orc_pips(const orc_cte*, const orc_div*, unsigned, unsigned, unsigned, const unsigned*,
              const float*, unsigned long long*);
}

namespace {

struct Sim {
  dsph_case cs{};
  int prec = 0;
  std::string err;
  orc_cte c{};
  float Scell = 0, KernelH = 0, SpsSmag = 0, SpsBlin = 0, Visco = 0, ViscoB = 0;
  float RhopOutMin = 0, RhopOutMax = 0, g[3] = {0, 0, 0};
  double Cs0 = 0, DtIni = 0, DtMin = 0, CFL = 0.2;
  int ScellDiv = 1, TStep = 2, TVisco = 1, TDDT = 0, VerletSteps = 40;
  unsigned DomCells[3] = {0, 0, 0};
  unsigned CaseNbound = 0;
  // particles
  unsigned Np = 0, Npb = 0, NpbOk = 0;
  std::vector<unsigned> idp, dcell;
  std::vector<typecode> code;
  std::vector<double> pos, pospre, mov;
  std::vector<float> velrho, velrhopre, velrhom1, poscell, ace, ar, viscdt, strain, tau;
  std::vector<float> boundnor;  // doubled normals (ConfigBoundNormals), mDBC only
  int TBoundary = 1, SlipMode = 1;
  bool NoPen = false;                // mDBC2 NoPenetration (MDBC2_NoPen)
  std::vector<float> nopenshift;     // float4 per particle, rewritten by every interaction
  int ShiftMode = 0;
  float ShiftCoef = -2.f, ShiftTFS = 0.f;
  std::vector<float> shiftposfs;     // shifting float4 per particle
  std::vector<unsigned char> bmode;  // mDBC2 boundary modes of [0,Npb)
  std::vector<float> tangenvel;      // mDBC2 tangential velocities of [0,Npb)
  // mDBC2 with moving boundaries: motion velocity and acceleration per
  // particle (MotionVel, MotionAce), zero at the start (JSphCpu.cpp:476).
  std::vector<float> motionvel, motionace;
  bool pre_active = false;
  // cell division
  std::vector<unsigned> cellpart, sortpart;
  std::vector<int> begincell;
  unsigned cmin[3] = {0, 0, 0}, cmax[3] = {0, 0, 0}, Ncx = 0, Ncy = 0, Ncz = 0, Nct = 0;
  bool BoundLimitOk = false, BoundDivideOk = false, DivideFull = false, BoundChanged = true;
  unsigned blmin[3], blmax[3], bdmin[3], bdmax[3];
  unsigned NpfOutTotal = 0, Nstep = 0, VerletStep = 0;
  double SymplecticDtPre = 0, TimeStep = 0, DtLast = 0, VelMax = 0, AceMax = 0;
  float ViscDtMax = 0;
  unsigned NpbPer = 0, NpfPer = 0;
  double MapPosMax[3] = {0, 0, 0};
  // Floating bodies (StFloatingData, DualSphDef.h; JSph::LoadCaseConfig,
  // JSph.cpp:1123-1167): body data, particle mass per body and the index of
  // each floating particle by id (RidpMot).
  struct FtObj {
    unsigned begin = 0, count = 0;
    float mass = 0, massp = 0;
    double center[3] = {0, 0, 0};
    float angles[3] = {0, 0, 0}, fvel[3] = {0, 0, 0}, fomega[3] = {0, 0, 0};
    float facelin[3] = {0, 0, 0}, faceang[3] = {0, 0, 0};
    float inertia[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  };
  std::vector<FtObj> ft;
  std::vector<float> ftmassp;
  std::vector<unsigned> ridp;
  unsigned CaseNpb = 0;

  void config(const dsph_case& s) {
    cs = s;
    KernelH = float(s.kernelh);
    const double h = KernelH;
    const float ks = float(h * 2.0);
    c.kernelh = KernelH;
    c.kernelsize = ks;
    c.kernelsize2 = ks * ks;
    // Wendland constants (FunSphKernel.h:198-210), 2-D or 3-D.
    c.sim2d = s.simulate2d ? 1u : 0u;
    c.awen = c.sim2d ? float(0.557 / (h * h)) : float(0.41778 / (h * h * h));
    c.bwenh = c.sim2d ? float(-2.7852 / (h * h * h * h)) : float(-2.08891 / (h * h * h * h * h));
    c.cteb = float(s.cteb); c.gamma = float(s.gamma); c.rhopzero = float(s.rhopzero);
    c.ovrhopzero = 1.f / c.rhopzero;
    Cs0 = std::sqrt(double(c.gamma) * double(c.cteb) / double(c.rhopzero));
    c.cs0 = float(Cs0);
    c.eta2 = float((h * 0.1) * (h * 0.1));
    c.massf = float(s.massfluid); c.massb = float(s.massbound); c.dp = float(s.dp);
    for (int k = 0; k < 3; k++) g[k] = float(s.gravity[k]);
    c.ddtkh = ks * float(s.ddtvalue);
    c.ddtgz = float(double(c.rhopzero) * double(std::fabs(g[2])) / double(c.cteb));
    TStep = s.tstep; TVisco = s.tvisco; TDDT = s.tdensity; TBoundary = s.tboundary;
    SlipMode = s.slipmode ? s.slipmode : 1;
    // NoPenetration is read for SlipMode >= no-slip only (JSph.cpp:715-716).
    NoPen = TBoundary == 2 && SlipMode >= 2 && s.nopenetration != 0;
    ShiftMode = s.shiftmode;
    ShiftCoef = float(s.shiftcoef ? s.shiftcoef : -2.0);
    ShiftTFS = float(s.shifttfs);
    VerletSteps = s.verletsteps ? s.verletsteps : 40;
    Visco = float(s.visco); ViscoB = Visco * float(s.viscoboundfactor);
    RhopOutMin = float(s.rhopoutmin); RhopOutMax = float(s.rhopoutmax);
    CFL = s.cflnumber;
    DtIni = s.dtini ? s.dtini : double(KernelH) / Cs0;
    DtMin = s.dtmin ? s.dtmin : (double(KernelH) / Cs0) * double(0.05f);
    // 2-D: linear motions lose their y part (JDsMotion.cpp:132).
    // 2-D: no periodic y (JSph.cpp:827).
    if (c.sim2d && (s.periactive & 2))
      throw std::runtime_error("Cannot use periodic conditions in Y with 2D simulations");
    if (c.sim2d) cs.motion.ampl[1] = 0;
    if (TVisco == 2) {
      // dp_sps (JSph.cpp:1529).
      const double dps = c.sim2d ? std::sqrt(double(c.dp) * double(c.dp) * 2.) / 2.
                                 : std::sqrt(double(c.dp) * double(c.dp) * 3.) / 3.;
      SpsSmag = float((2. / 3.) * std::pow(0.12 * dps, 2));
      SpsBlin = float((2. / 3.) * 0.0066 * dps * dps);
    }
    ScellDiv = s.cellmode == 2 ? 2 : 1;
    double msize[3];
    c.periactive = unsigned(s.periactive) & 7u;
    if (c.periactive && s.casenmoving)
      throw std::runtime_error("periodic boundaries with moving boundaries are not built");
    for (int k = 0; k < 3; k++) {
      c.maprealposmin[k] = s.maprealposmin[k];
      c.maprealsize[k] = s.maprealposmax[k] - s.maprealposmin[k];
      c.mapposmin[k] = s.maprealposmin[k];
      MapPosMax[k] = s.maprealposmax[k];
    }
    // JSph.cpp:2183-2195: periodic increments and the map grown by
    // KernelSize on both sides of each periodic axis.
    for (int a = 0; a < 3; a++) {
      const bool on = (c.periactive >> a) & 1u;
      for (int k = 0; k < 3; k++) c.perinc[a][k] = on ? s.perinc[a][k] : 0.0;
      if (on) {
        c.perinc[a][a] = -c.maprealsize[a];
        c.mapposmin[a] = c.mapposmin[a] - double(ks);
        MapPosMax[a] = MapPosMax[a] + double(ks);
      }
    }
    for (int k = 0; k < 3; k++) {
      c.domposmin[k] = c.mapposmin[k];
      msize[k] = MapPosMax[k] - c.mapposmin[k];
    }
    c.poscellsize = ks;
    {
      int nks = 1;
      auto fits = [&](double pcs) {
        return unsigned(std::ceil(msize[0] / pcs)) <= 8192 && unsigned(std::ceil(msize[1] / pcs)) <= 1024 &&
               unsigned(std::ceil(msize[2] / pcs)) <= 512;
      };
      while (!fits(c.poscellsize)) { nks = (nks == 1 ? 2 : nks + 2); c.poscellsize = ks * nks; }
    }
    Scell = ks / ScellDiv;
    c.scell = Scell;
    c.movlimit = Scell * 0.9f;
    c.tboundary = unsigned(TBoundary);
    for (int k = 0; k < 3; k++) DomCells[k] = unsigned(std::ceil(msize[k] / Scell));
    c.cellcode = orc_cell_code(DomCells[0] + 1, DomCells[1] + 1, DomCells[2] + 1);
    CaseNbound = s.casenbound;
    CaseNpb = s.casenfixed + s.casenmoving;
    ft.clear();
    ftmassp.clear();
    if (s.casenfloat) {
      if (!s.ftcount || !s.floatings) throw std::runtime_error("floating particles without bodies");
      if (TBoundary != 1 || ShiftMode || c.periactive)
        throw std::runtime_error("floating bodies are built with DBC boundaries, without "
                                 "shifting or periodic axes");
      unsigned id = CaseNpb;
      for (unsigned cf = 0; cf < s.ftcount; cf++) {
        const dsph_floating& f = s.floatings[cf];
        if (f.begin != id || !f.count) throw std::runtime_error("floating bodies must tile the floating ids");
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
        ftmassp.push_back(o.massp);
      }
      if (id != CaseNbound) throw std::runtime_error("floating bodies must tile the floating ids");
    }
  }

  void load(unsigned np, const unsigned* hidp, const double* hpos, const float* hvr) {
    Np = np; Npb = cs.casenfixed + cs.casenmoving; NpbOk = Npb;
    idp.assign(hidp, hidp + np);
    pos.assign(hpos, hpos + 3 * size_t(np));
    velrho.assign(hvr, hvr + 4 * size_t(np));
    if (mdbc2_on() && cs.casenmoving) {
      motionvel.assign(3 * size_t(np), 0.f);
      motionace.assign(3 * size_t(np), 0.f);
    }
    code.resize(np);
    dcell.resize(np);
    for (unsigned p = 0; p < np; p++) {
      // LoadCodeParticles: fixed 0x0000, moving CODE_TYPE_MOVING|1, floating
      // CODE_TYPE_FLOATING|body, fluid CODE_TYPE_FLUID.
      const unsigned id = idp[p];
      typecode cd = 0x1800;
      if (id < cs.casenfixed) cd = 0x0000;
      else if (id < CaseNpb) cd = 0x0801;
      else if (id < CaseNbound)
        for (unsigned cf = 0; cf < ft.size(); cf++)
          if (id - ft[cf].begin < ft[cf].count) cd = typecode(0x1000u | cf);
      code[p] = cd;
      const double* x = &pos[3 * size_t(p)];
      const unsigned cx = unsigned((x[0] - c.domposmin[0]) / Scell);
      const unsigned cy = unsigned((x[1] - c.domposmin[1]) / Scell);
      const unsigned cz = unsigned((x[2] - c.domposmin[2]) / Scell);
      const unsigned dcc = c.cellcode;
      dcell[p] = (cx << ((dcc >> 10) & 31u)) | (cy << ((dcc >> 15) & 31u)) | cz;
    }
    if (TStep == 1) velrhom1 = velrho;
    if (TVisco == 2) { tau.assign(6 * size_t(np), 0.f); strain.assign(6 * size_t(np), 0.f); }
    SymplecticDtPre = DtIni;
    BoundChanged = true;
    cell_divide();
  }

  orc_div divdata() const {
    orc_div d;
    d.scelldiv = ScellDiv;
    d.ncx = int(Ncx); d.ncy = int(Ncy); d.ncz = int(Ncz); d.ncw = int(Ncx * Ncy);
    d.cellfluid = Ncx * Ncy * Ncz + 1;
    d.czx = int(cmin[0]); d.czy = int(cmin[1]); d.czz = int(cmin[2]);
    d.begincell = begincell.data();
    return d;
  }

  template <typename T>
  void gather(std::vector<T>& a, unsigned w, unsigned n, unsigned pini) {
    std::vector<T> b(a.size());
    for (unsigned p = 0; p < n; p++) {
      const unsigned o = p < pini ? p : sortpart[p];
      for (unsigned k = 0; k < w; k++) b[size_t(p) * w + k] = a[size_t(o) * w + k];
    }
    for (size_t k = size_t(n) * w; k < a.size(); k++) b[k] = a[k];
    a.swap(b);
  }

  // JSphCpuSingle::RunPeriodic (JSphCpuSingle.cpp:420-500): the copies of
  // the last divide are ignored; new ones are appended at Np (boundary
  // copies first), for each periodic axis from the copies made so far and
  // then from the original particles; the list is in particle order, the
  // plus copy of a particle before its minus copy (PeriodicMakeList,
  // JSphCpuSingle.cpp:210-240); a copy takes the shifted position, the cell
  // from DomPosMin clamped to the domain, a periodic code and the particle's
  // state (PeriodicDuplicate*, JSphCpuSingle.cpp:245-330).
  template <typename T>
  void grow(std::vector<T>& a, unsigned w, unsigned n) { if (!a.empty()) a.resize(size_t(n) * w); }
  void copy_row(std::vector<float>& a, unsigned w, unsigned dst, unsigned src) {
    if (a.empty()) return;
    for (unsigned k = 0; k < w; k++) a[size_t(dst) * w + k] = a[size_t(src) * w + k];
  }
  void run_periodic() {
    for (unsigned p = 0; p < Np; p++)
      if ((code[p] & 0xe000u) == 0x2000u) code[p] = typecode((code[p] & ~0xe000u) | 0x4000u);
    const unsigned npb0 = Npb, npf0 = Np - Npb;
    NpbPer = NpfPer = 0;
    BoundChanged = true;
    for (unsigned ctype = 0; ctype < 2; ctype++) {
      const unsigned pini = ctype ? npb0 : 0u, num = ctype ? npf0 : npb0;
      for (unsigned a = 0; a < 3; a++) {
        if (!((c.periactive >> a) & 1u)) continue;
        const double* inc = c.perinc[a];
        for (unsigned cblock = 0; cblock < 2; cblock++) {
          const unsigned nper = ctype ? NpfPer : NpbPer;
          const unsigned pini2 = cblock ? pini : Np - nper, num2 = cblock ? num : nper;
          std::vector<unsigned> list;
          for (unsigned p2 = pini2; p2 < pini2 + num2; p2++) {
            if ((code[p2] & 0xe000u) > 0x2000u) continue;
            const double* x = &pos[3 * size_t(p2)];
            for (int sg = 0; sg < 2; sg++) {
              const double q[3] = {sg ? x[0] - inc[0] : x[0] + inc[0],
                                   sg ? x[1] - inc[1] : x[1] + inc[1],
                                   sg ? x[2] - inc[2] : x[2] + inc[2]};
              bool in = true;
              for (int k = 0; k < 3; k++) in = in && c.mapposmin[k] <= q[k] && q[k] < MapPosMax[k];
              if (in) list.push_back(sg ? (p2 | 0x80000000u) : p2);
            }
          }
          const unsigned count = unsigned(list.size());
          if (!count) continue;
          const unsigned n1 = Np + count;
          idp.resize(n1); code.resize(n1); dcell.resize(n1);
          pos.resize(3 * size_t(n1)); velrho.resize(4 * size_t(n1));
          grow(velrhom1, 4, n1); grow(tau, 6, n1); grow(strain, 6, n1); grow(boundnor, 3, n1);
          if (pre_active) { grow(pospre, 3, n1); grow(velrhopre, 4, n1); }
          for (unsigned i = 0; i < count; i++) {
            const unsigned pn = Np + i, pc = list[i] & 0x7fffffffu;
            const bool inv = list[i] >= 0x80000000u;
            unsigned cl[3];
            for (int k = 0; k < 3; k++) {
              const double v = pos[3 * size_t(pc) + k] + (inv ? -inc[k] : inc[k]);
              pos[3 * size_t(pn) + k] = v;
              const unsigned cc = unsigned((v - c.domposmin[k]) / Scell);
              cl[k] = std::min(cc, DomCells[k] - 1);
            }
            dcell[pn] = mkcell_host(cl[0], cl[1], cl[2]);
            idp[pn] = idp[pc];
            code[pn] = typecode((code[pc] & ~0xe000u) | 0x2000u);
            copy_row(velrho, 4, pn, pc);
            copy_row(velrhom1, 4, pn, pc);
            copy_row(tau, 6, pn, pc);
            copy_row(boundnor, 3, pn, pc);
            if (pre_active) {
              for (int k = 0; k < 3; k++) pospre[3 * size_t(pn) + k] = pospre[3 * size_t(pc) + k];
              copy_row(velrhopre, 4, pn, pc);
            }
          }
          Np = n1;
          if (!ctype) NpbPer += count; else NpfPer += count;
        }
      }
    }
  }

  // JCellDivCpuSingle::Divide + JSphCpuSingle::RunCellDivide.
  void cell_divide() {
    if (c.periactive) run_periodic();
    const unsigned npb2 = NpbPer, npf2 = NpfPer;
    const unsigned npb1 = Npb, npf1 = Np - Npb - npb2 - npf2, nptot = Np;
    if (BoundChanged || c.periactive) { BoundLimitOk = BoundDivideOk = false; }
    if (cs.celldomfixed) {
      for (int k = 0; k < 3; k++) { cmin[k] = 0; cmax[k] = DomCells[k] - 1; }
    } else {
      unsigned bmin[3], bmax[3], fmin[3], fmax[3];
      if (!BoundLimitOk) {
        orc_limits(npb1, 0, c.cellcode, dcell.data(), code.data(), bmin, bmax);
        if (bmin[0] > bmax[0]) for (int k = 0; k < 3; k++) { bmin[k] = DomCells[k]; bmax[k] = 0; }
        merge_limits(npb2, npb1 + npf1, bmin, bmax);
        BoundLimitOk = true;
        std::memcpy(blmin, bmin, 12); std::memcpy(blmax, bmax, 12);
      } else { std::memcpy(bmin, blmin, 12); std::memcpy(bmax, blmax, 12); }
      orc_limits(npf1, npb1, c.cellcode, dcell.data(), code.data(), fmin, fmax);
      if (fmin[0] > fmax[0]) for (int k = 0; k < 3; k++) { fmin[k] = DomCells[k]; fmax[k] = 0; }
      merge_limits(npf2, npb1 + npf1 + npb2, fmin, fmax);
      const unsigned s = unsigned(ScellDiv);
      for (int k = 0; k < 3; k++) {
        cmin[k] = std::max(std::min(bmin[k], fmin[k]), fmin[k] >= s ? fmin[k] - s : 0u);
        cmax[k] = std::min(std::max(bmax[k], fmax[k]), fmax[k] + s);
        if (cmax[k] >= DomCells[k]) cmax[k] = DomCells[k] - 1;
      }
      if (cmin[0] > cmax[0] || cmin[1] > cmax[1] || cmin[2] > cmax[2])
        for (int k = 0; k < 3; k++) cmin[k] = cmax[k] = 0;
    }
    Ncx = cmax[0] - cmin[0] + 1; Ncy = cmax[1] - cmin[1] + 1; Ncz = cmax[2] - cmin[2] + 1;
    Nct = Ncx * Ncy * Ncz;
    const unsigned nctt = 2 * Nct + 5;
    if (!BoundDivideOk || std::memcmp(bdmin, cmin, 12) || std::memcmp(bdmax, cmax, 12)) {
      DivideFull = true; BoundDivideOk = true;
      std::memcpy(bdmin, cmin, 12); std::memcpy(bdmax, cmax, 12);
    } else DivideFull = false;
    const unsigned nc[3] = {Ncx, Ncy, Ncz};
    cellpart.resize(nptot);
    sortpart.resize(nptot);
    begincell.resize(2 * size_t(nctt));
    orc_presort(DivideFull, nptot, npb1, c.cellcode, dcell.data(), code.data(), cmin, nc,
                cellpart.data());
    const unsigned key0 = DivideFull ? 0 : Nct + 1;
    const unsigned pini = DivideFull ? 0 : npb1;
    orc_sort_cells(nptot, pini, nctt, key0, cellpart.data(), sortpart.data(), begincell.data());
    auto csize = [&](unsigned box) { return unsigned(begincell[2 * box + 1] - begincell[2 * box]); };
    const unsigned npbignore = csize(Nct);
    const unsigned bbo = 2 * Nct + 1;
    const unsigned npbout = csize(bbo), npfout = csize(bbo + 1), npboi = csize(bbo + 2),
                   npfoi = csize(bbo + 3);
    const unsigned npfinal = nptot - npbout - npfout - npboi - npfoi;
    unsigned npbfinal = npb1 + npb2 - npboi;
    if (npbout && DivideFull) npbfinal = UINT_MAX;
    const unsigned gpini = DivideFull ? 0 : npbfinal;
    gather(idp, 1, nptot, gpini); gather(code, 1, nptot, gpini); gather(dcell, 1, nptot, gpini);
    gather(pos, 3, nptot, gpini); gather(velrho, 4, nptot, gpini);
    if (TStep == 1) gather(velrhom1, 4, nptot, gpini);
    else if (pre_active) { gather(pospre, 3, nptot, gpini); gather(velrhopre, 4, nptot, gpini); }
    if (TVisco == 2) gather(tau, 6, nptot, gpini);
    if (TBoundary == 2) gather(boundnor, 3, nptot, gpini);
    if (!motionvel.empty()) { gather(motionvel, 3, nptot, gpini); gather(motionace, 3, nptot, gpini); }
    if (npbout) throw std::runtime_error("boundary particles excluded");
    Np = npfinal; Npb = npbfinal; NpbOk = Npb - npbignore;
    NpfOutTotal += npfout;
    poscell.resize(4 * size_t(Np));
    orc_poscell(Np, pos.data(), c.mapposmin, c.poscellsize, poscell.data());
    BoundChanged = false;
  }

  // Second range of CalcCellDomainBound/Fluid (JCellDivGpu.cpp:250-300).
  void merge_limits(unsigned n2, unsigned pini2, unsigned mn[3], unsigned mx[3]) {
    if (!n2) return;
    unsigned a[3], b[3];
    orc_limits(n2, pini2, c.cellcode, dcell.data(), code.data(), a, b);
    if (a[0] > b[0]) for (int k = 0; k < 3; k++) { a[k] = DomCells[k]; b[k] = 0; }
    for (int k = 0; k < 3; k++) { mn[k] = std::min(mn[k], a[k]); mx[k] = std::max(mx[k], b[k]); }
  }

  void interaction() {
    ace.assign(3 * size_t(Np), 0.f);
    ar.assign(Np, 0.f);
    viscdt.assign(Np, 0.f);
    if (TVisco == 2) std::fill(strain.begin() + 6 * size_t(Npb), strain.begin() + 6 * size_t(Np), 0.f);
    float v2 = 0;
    for (unsigned p = Npb; p < Np; p++) {
      const float* v = &velrho[4 * size_t(p)];
      v2 = std::max(v2, v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    }
    VelMax = double(std::sqrt(v2));
    const orc_div d = divdata();
    // prec 0: FP32 GPU semantics, 1: FP64 yardstick, 2: FP32 CPU semantics.
    const bool m2 = mdbc2_on();
    if (ShiftMode) shiftposfs.assign(4 * size_t(Np), 0.f);  // JSphShifting::InitCpu, no zones
    if (NoPen) nopenshift.assign(4 * size_t(Np), 0.f);  // JSphCpu.cpp:519
    orc_interaction_range_mv(prec == 1 ? 1 : 0, prec == 1 ? 2 : (prec == 2 ? 1 : 0), TVisco,
                             TDDT, ViscoB, Visco, &c, &d, Np, Npb, NpbOk,
                             dcell.data(), poscell.data(), pos.data(), velrho.data(),
                             TVisco == 2 ? tau.data() : nullptr, ace.data(), ar.data(),
                             viscdt.data(), TVisco == 2 ? strain.data() : nullptr, nullptr,
                             m2 ? bmode.data() : nullptr, m2 ? tangenvel.data() : nullptr,
                             ShiftMode, code.data(), ShiftMode ? shiftposfs.data() : nullptr,
                             ft.empty() ? nullptr : ftmassp.data(),
                             NoPen ? boundnor.data() : nullptr,
                             NoPen ? nopenshift.data() : nullptr,
                             motionvel.empty() ? nullptr : motionvel.data(), 0, LONG_MAX, 0,
                             LONG_MAX);
    if (TVisco == 2) orc_sps_tau(Np, Npb, SpsSmag, SpsBlin, velrho.data(), strain.data(), tau.data());
    float vd = 0, a2 = 0;
    for (unsigned p = 0; p < Np; p++) vd = std::max(vd, viscdt[p]);
    for (unsigned p = Npb; p < Np; p++) {
      if (code[p] & 0xe000u) continue;  // ComputeAceMod: normal particles only
      const float* a = &ace[3 * size_t(p)];
      a2 = std::max(a2, a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    }
    ViscDtMax = vd;
    AceMax = std::sqrt(double(a2));
  }

  double dt_variable() const {
    const double dt1 = (AceMax ? std::sqrt(double(KernelH) / AceMax) : DBL_MAX);
    const double dt2 = double(KernelH) / (std::max(Cs0, VelMax * 10.) + double(KernelH) * ViscDtMax);
    double dt = CFL * std::min(dt1, dt2);
    if (std::isnan(dt) || std::isinf(dt)) throw std::runtime_error("dt is NaN or infinity");
    if (dt < DtMin) dt = DtMin;
    return dt;
  }

  // JSphCpu::RunMotion for one moving object, applied after the step; the
  // next divide is full. Rectilinear sinusoid (JMotionObj.cpp:558-575,
  // JMotionList.cpp:43-50): MoveLinBound with displacement over [t, t+dt] and
  // velocity displacement/dt. Rotations (sinusoidal JMotionObj.cpp:577-587,
  // constant angular velocity :518-524): MoveMatBound (JSphCpu.cpp
  // MoveMatBound; GPU KerMoveMatBound JSphGpu_ker.cu:1705-1740) with the
  // matrix JMatrix4::MatrixRot builds (JMatrix4.h:465-497): translate the
  // axis point to the origin, turn the axis onto z (about x, then about y),
  // rotate about z, undo; x' = M x, velocity (x'-x)/dt, mDBC normals
  // M(x+n) - M(x).
  unsigned mkcell_host(unsigned cx, unsigned cy, unsigned cz) const {
    const unsigned dcc = c.cellcode;
    return (cx << ((dcc >> 10) & 31u)) | (cy << ((dcc >> 15) & 31u)) | cz;
  }
  static void mat_mul(const double a[16], const double b[16], double out[16]) {
    double r[16];
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++) {
        double v = 0;
        for (int k = 0; k < 4; k++) v += a[4 * i + k] * b[4 * k + j];
        r[4 * i + j] = v;
      }
    std::memcpy(out, r, sizeof(r));
  }
  static void mat_rot(double ang, const double p1[3], const double p2[3], double t[16]) {
    const double rad = ang * (3.14159265358979323846 / 180.0);
    const double v[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
    const double L = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    const double L1 = std::sqrt(v[1] * v[1] + v[2] * v[2]);
    auto ident = [](double m[16]) { for (int k = 0; k < 16; k++) m[k] = (k % 5 == 0) ? 1 : 0; };
    double tr[16], tm1[16], rx[16], rxm1[16], ry[16], rym1[16], rz[16];
    ident(tr); ident(tm1); ident(rx); ident(rxm1); ident(ry); ident(rym1); ident(rz);
    for (int k = 0; k < 3; k++) { tr[4 * k + 3] = p1[k]; tm1[4 * k + 3] = -p1[k]; }
    if (L1 == 0) {
      rx[5] = 0; rx[6] = 1; rx[9] = -1; rx[10] = 0;
      rxm1[5] = 0; rxm1[6] = -1; rxm1[9] = 1; rxm1[10] = 0;
    } else {
      rx[5] = v[2] / L1; rx[6] = v[1] / L1; rx[9] = -v[1] / L1; rx[10] = v[2] / L1;
      rxm1[5] = v[2] / L1; rxm1[6] = -v[1] / L1; rxm1[9] = v[1] / L1; rxm1[10] = v[2] / L1;
    }
    ry[0] = L1 / L; ry[2] = v[0] / L; ry[8] = -v[0] / L; ry[10] = L1 / L;
    rym1[0] = L1 / L; rym1[2] = -v[0] / L; rym1[8] = v[0] / L; rym1[10] = L1 / L;
    const double cs = std::cos(rad), sn = std::sin(rad);
    rz[0] = cs; rz[1] = sn; rz[4] = -sn; rz[5] = cs;
    std::memcpy(t, tr, sizeof(tr));
    mat_mul(t, rx, t); mat_mul(t, ry, t); mat_mul(t, rz, t);
    mat_mul(t, rym1, t); mat_mul(t, rxm1, t); mat_mul(t, tm1, t);
  }
  void move_particle(unsigned p, const double d[3],
                     const char* what = "moving boundary particle left the domain") {
    double* x = &pos[3 * size_t(p)];
    for (int k = 0; k < 3; k++) x[k] += d[k];
    double dd[3];
    bool out = false;
    for (int k = 0; k < 3; k++) {
      dd[k] = x[k] - c.maprealposmin[k];
      if (dd[k] != dd[k] || dd[k] < 0 || dd[k] >= c.maprealsize[k]) out = true;
    }
    if (out) throw std::runtime_error(what);
    dcell[p] = mkcell_host(unsigned(dd[0] / c.scell), unsigned(dd[1] / c.scell),
                           unsigned(dd[2] / c.scell));
  }
  // JSphCpu::RunMotion (JSphCpu.cpp:2188-2234): the motion when active,
  // then, with mDBC2, the motion velocity and acceleration of the moving
  // particles (CopyMotionVelAce) whether the motion is active or not.
  void run_motion(double dt) {
    const dsph_motion& m = cs.motion;
    if (!cs.casenmoving || m.type == 0) return;
    move_boundaries(dt);
    if (!motionvel.empty())
      orc_motion_vel_ace(Npb, m.idbegin, m.count, dt, idp.data(), velrho.data(),
                         motionvel.data(), motionace.data());
  }
  void move_boundaries(double dt) {
    const dsph_motion& m = cs.motion;
    if (!(TimeStep >= m.start && TimeStep < m.finish)) return;
    const double twopi = 6.283185307179586476925286766559;
    if (m.type == 1) {
      double mv[3];
      for (int k = 0; k < 3; k++)
        mv[k] = m.ampl[k] != 0 ? m.ampl[k] * std::sin(m.phase[k] + twopi * m.freq[k] * (TimeStep + dt)) -
                                     m.ampl[k] * std::sin(m.phase[k] + twopi * m.freq[k] * TimeStep)
                               : 0.0;
      for (unsigned p = 0; p < Npb; p++) {
        if (idp[p] - m.idbegin >= m.count) continue;
        move_particle(p, mv);
        float* v = &velrho[4 * size_t(p)];
        for (int k = 0; k < 3; k++) v[k] = float(mv[k] / dt);
      }
    } else {
      double ang = 0;
      if (m.type == 2)
        ang = m.angampl * std::sin(m.angphase + twopi * m.angfreq * (TimeStep + dt)) -
              m.angampl * std::sin(m.angphase + twopi * m.angfreq * TimeStep);
      else
        ang = m.angvel * dt;
      double t[16];
      mat_rot(ang, m.axisp1, m.axisp2, t);
      auto apply = [&](const double x[3], double y[3]) {
        for (int i = 0; i < 3; i++)
          y[i] = x[0] * t[4 * i] + x[1] * t[4 * i + 1] + x[2] * t[4 * i + 2] + t[4 * i + 3];
      };
      for (unsigned p = 0; p < Npb; p++) {
        if (idp[p] - m.idbegin >= m.count) continue;
        const double x[3] = {pos[3 * size_t(p)], pos[3 * size_t(p) + 1], pos[3 * size_t(p) + 2]};
        double x2[3], d[3];
        apply(x, x2);
        if (c.sim2d) x2[1] = x[1];  // 2-D: y kept (JSphCpu.cpp:2161)
        for (int k = 0; k < 3; k++) d[k] = x2[k] - x[k];
        move_particle(p, d);
        float* v = &velrho[4 * size_t(p)];
        for (int k = 0; k < 3; k++) v[k] = float(d[k] / dt);
        if (!boundnor.empty()) {
          float* bn = &boundnor[3 * size_t(p)];
          const double gs[3] = {x[0] + bn[0], x[1] + bn[1], x[2] + bn[2]};
          double gs2[3];
          apply(gs, gs2);
          for (int k = 0; k < 3; k++) bn[k] = float(gs2[k] - x2[k]);
        }
      }
    }
    BoundChanged = true;
  }

  // JSphCpuSingle::MdbcBoundCorrection (JSphCpuSingle.cpp:754-775): mDBC
  // (SLIP_Vel0) before the predictor (Symplectic) or the Verlet interaction,
  // not before the corrector (MdbcCorrector=false); no-slip (mDBC2) before
  // every interaction (MdbcCorrector=true), modes cleared to DBC first.
  bool mdbc2_on() const { return TBoundary == 2 && SlipMode >= 2; }
  void mdbc(bool corrector = false) {
    if (TBoundary != 2) return;
    const orc_div d = divdata();
    if (mdbc2_on()) {
      bmode.assign(Npb, 0);
      tangenvel.assign(3 * size_t(Npb), 0.f);
      const bool mv = !motionvel.empty();
      orc_mdbc2_slip(prec == 1 ? 1 : 0, prec == 2 ? 1 : 0, SlipMode, &c, &d, Npb, pos.data(),
                     poscell.data(), code.data(), boundnor.data(), g,
                     mv ? motionvel.data() : nullptr, mv ? motionace.data() : nullptr, Cs0,
                     cs.dp, velrho.data(), bmode.data(), tangenvel.data());
      return;
    }
    if (corrector) return;
    orc_mdbc(prec == 1 ? 1 : 0, prec == 2 ? 1 : 0, &c, &d, NpbOk, pos.data(), poscell.data(),
             code.data(), boundnor.data(), velrho.data());
  }

  // JSphCpu::RunShifting: shifting displacement from the interaction's sums.
  void shifting(double dt) {
    if (!ShiftMode) return;
    orc_run_shifting(Np - Npb, Npb, dt, ShiftCoef, KernelH, ShiftTFS, cs.dp, velrho.data(),
                     shiftposfs.data(), c.sim2d);
  }

  // ---- floating bodies: JSphCpuSingle::RunFloating (JSphCpuSingle.cpp:938-993)
  // without Chrono, moorings, external forces, imposed velocities or
  // constraints, and FtPause 0.
  void calc_ridp() {
    ridp.assign(cs.casenfloat, UINT_MAX);
    for (unsigned p = Npb; p < Np; p++) {
      const unsigned id = idp[p];
      if (id >= CaseNpb && id < CaseNbound) ridp[id - CaseNpb] = p;
    }
    for (unsigned k = 0; k < cs.casenfloat; k++)
      if (ridp[k] == UINT_MAX) throw std::runtime_error("There are floating particles not found.");
  }
  // FtPartsSumAce: sum of ace and of (pos-center) x ace over the particles of
  // body cf. CPU build (prec 2): in id order (JSphCpuSingle.cpp:1014-1040);
  // GPU build (prec 0): KerFtPartsSumAce's order (JSphGpu_ker.cu:1938-2012),
  // 256 partial sums (partial t takes ids t, t+256, ...) then summed in
  // order; FP64 yardstick (prec 1): in id order in FP64.
  void ft_sum_ace(unsigned cf, float acelin[3], float aceang[3]) const {
    const FtObj& f = ft[cf];
    const unsigned fpini = f.begin - CaseNpb;
    auto term = [&](unsigned fp, float a[3], float dist[3]) {
      const unsigned p = ridp[fpini + fp];
      for (int k = 0; k < 3; k++) {
        a[k] = ace[3 * size_t(p) + k];
        dist[k] = float(pos[3 * size_t(p) + k] - f.center[k]);
      }
    };
    if (prec == 1) {
      double l[3] = {0, 0, 0}, g[3] = {0, 0, 0};
      for (unsigned fp = 0; fp < f.count; fp++) {
        float a[3], d[3];
        term(fp, a, d);
        for (int k = 0; k < 3; k++) l[k] += a[k];
        g[0] += double(a[2]) * d[1] - double(a[1]) * d[2];
        g[1] += double(a[0]) * d[2] - double(a[2]) * d[0];
        g[2] += double(a[1]) * d[0] - double(a[0]) * d[1];
      }
      for (int k = 0; k < 3; k++) { acelin[k] = float(l[k]); aceang[k] = float(g[k]); }
      return;
    }
    const unsigned nt = prec == 0 ? 256u : 1u;
    const unsigned ntid = std::min(f.count, nt);
    std::vector<float> pl(3 * size_t(ntid), 0.f), pg(3 * size_t(ntid), 0.f);
    for (unsigned fp = 0; fp < f.count; fp++) {
      const unsigned t = fp % nt;
      float a[3], d[3];
      term(fp, a, d);
      for (int k = 0; k < 3; k++) pl[3 * t + k] += a[k];
      pg[3 * t] += (a[2] * d[1] - a[1] * d[2]);
      pg[3 * t + 1] += (a[0] * d[2] - a[2] * d[0]);
      pg[3 * t + 2] += (a[1] * d[0] - a[0] * d[1]);
    }
    for (int k = 0; k < 3; k++) acelin[k] = aceang[k] = 0.f;
    for (unsigned t = 0; t < ntid; t++)
      for (int k = 0; k < 3; k++) { acelin[k] += pl[3 * t + k]; aceang[k] += pg[3 * t + k]; }
  }
  // FtComputeAceVel (JSph.cpp:2565-2660) for one body: accelerations from the
  // particle sums, gravity and the inertia tensor rotated to the body's
  // angles; new centre, linear and angular velocity over dt.
  void ft_compute_ace_vel(unsigned cf, double dt, const float sl[3], const float sa[3],
                          float fvel[3], float fomega[3], double fcenter[3]) const {
    const FtObj& f = ft[cf];
    float fforcelin[3], fforceang[3], acelin[3], aceang[3];
    for (int k = 0; k < 3; k++) {
      fforcelin[k] = sl[k] * f.massp;
      fforceang[k] = sa[k] * f.massp;
      acelin[k] = (fforcelin[k] + g[k] * f.mass) / f.mass;
    }
    // fmath::RotMatrix3x3 (FunctionsMath.h:343-352), I = (R I0) R^T, inverse.
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
    if (det) {  // fmath::InverseMatrix3x3 (FunctionsMath.h:132-150)
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
    if (c.sim2d) {  // in-plane motion only (JSph.cpp:2628-2640)
      fcenter[1] = f.center[1];
      fvel[1] = 0;
      fomega[0] = fomega[2] = 0;
    }
  }
  // FtPartsUpdate (JSphCpuSingle.cpp:1045-1105) without periodic axes and
  // mDBC normals: each particle moves by dt times its velocity (UpdatePos)
  // and takes the rigid velocity about the new centre.
  void ft_parts_update(unsigned cf, double dt, const float fvel[3], const float fomega[3],
                       const double fcenter[3]) {
    const FtObj& f = ft[cf];
    const unsigned fpini = f.begin - CaseNpb;
    for (unsigned fp = 0; fp < f.count; fp++) {
      const unsigned p = ridp[fpini + fp];
      float* vr = &velrho[4 * size_t(p)];
      const double dd[3] = {dt * double(vr[0]), dt * double(vr[1]), dt * double(vr[2])};
      move_particle(p, dd, "floating particle left the domain");
      float dist[3];
      for (int k = 0; k < 3; k++) dist[k] = float(pos[3 * size_t(p) + k] - fcenter[k]);
      vr[0] = fvel[0] + (fomega[1] * dist[2] - fomega[2] * dist[1]);
      vr[1] = fvel[1] + (fomega[2] * dist[0] - fomega[0] * dist[2]);
      vr[2] = fvel[2] + (fomega[0] * dist[1] - fomega[1] * dist[0]);
    }
  }
  void run_floating(double dt, bool predictor) {
    if (ft.empty()) return;
    calc_ridp();
    for (unsigned cf = 0; cf < ft.size(); cf++) {
      float sl[3], sa[3], fvel[3], fomega[3];
      double fcenter[3];
      ft_sum_ace(cf, sl, sa);
      ft_compute_ace_vel(cf, dt, sl, sa, fvel, fomega, fcenter);
      ft_parts_update(cf, dt, fvel, fomega, fcenter);
      if (!predictor) {  // FtUpdateFloatings (JSph.cpp:2689-2704)
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

  void step_sym() {
    const double dt = SymplecticDtPre;
    mdbc();
    interaction();
    const double dt_p = dt_variable();
    shifting(dt * .5);
    pospre = pos; velrhopre = velrho; pre_active = true;
    mov.assign(3 * size_t(Np), 0.0);
    orc_sym_pre(Np, Npb, velrhopre.data(), ar.data(), ace.data(), dt * .5, c.rhopzero, RhopOutMin,
                RhopOutMax, g, code.data(), mov.data(), velrho.data());
    if (ShiftMode) orc_add_shift(Np, Npb, shiftposfs.data(), mov.data());
    if (mdbc2_on()) orc_mdbc2_keep(Npb, bmode.data(), velrhopre.data(), velrho.data());
    orc_step_pos(Np, Npb, &c, pospre.data(), mov.data(), pos.data(), dcell.data(), code.data());
    run_floating(dt * .5, true);
    cell_divide();
    mdbc(true);
    interaction();
    const double dt_c = dt_variable();
    shifting(dt);
    mov.assign(3 * size_t(Np), 0.0);
    orc_sym_cor(Np, Npb, velrhopre.data(), ar.data(), ace.data(), dt * .5, dt, c.rhopzero,
                RhopOutMin, RhopOutMax, g, code.data(), mov.data(), velrho.data());
    if (NoPen) orc_nopen_step(Np, Npb, nopenshift.data(), velrhopre.data(), dt, velrho.data(),
                              mov.data());
    if (ShiftMode) orc_add_shift(Np, Npb, shiftposfs.data(), mov.data());
    if (mdbc2_on()) orc_mdbc2_keep(Npb, bmode.data(), velrhopre.data(), velrho.data());
    orc_step_pos(Np, Npb, &c, pospre.data(), mov.data(), pos.data(), dcell.data(), code.data());
    run_floating(dt, false);
    pre_active = false;
    run_motion(dt);
    SymplecticDtPre = std::min(dt_p, dt_c);
    TimeStep += dt;
    DtLast = dt;
  }

  void step_ver() {
    mdbc();
    interaction();
    const double dt = dt_variable();
    shifting(dt);
    VerletStep++;
    std::vector<float> nw(velrho.size());
    mov.assign(3 * size_t(Np), 0.0);
    if (VerletStep < unsigned(VerletSteps)) {
      orc_verlet(Np, Npb, velrho.data(), velrhom1.data(), ar.data(), ace.data(), dt, dt + dt,
                 c.rhopzero, RhopOutMin, RhopOutMax, g, code.data(), mov.data(), nw.data());
      if (mdbc2_on()) orc_mdbc2_keep(Npb, bmode.data(), velrhom1.data(), nw.data());
    } else {
      orc_verlet(Np, Npb, velrho.data(), velrho.data(), ar.data(), ace.data(), dt, dt, c.rhopzero,
                 RhopOutMin, RhopOutMax, g, code.data(), mov.data(), nw.data());
      if (mdbc2_on()) orc_mdbc2_keep(Npb, bmode.data(), velrho.data(), nw.data());
      VerletStep = 0;
    }
    if (NoPen) orc_nopen_step(Np, Npb, nopenshift.data(), velrho.data(), dt, nw.data(), mov.data());
    if (ShiftMode) orc_add_shift(Np, Npb, shiftposfs.data(), mov.data());
    std::vector<double> p0 = pos;
    orc_step_pos(Np, Npb, &c, p0.data(), mov.data(), pos.data(), dcell.data(), code.data());
    velrhom1 = velrho;
    velrho = nw;
    run_floating(dt, false);
    run_motion(dt);
    TimeStep += dt;
    DtLast = dt;
  }

  void run(unsigned n) {
    for (unsigned s = 0; s < n; s++) {
      if (TStep == 2) step_sym(); else step_ver();
      cell_divide();
      Nstep++;
    }
  }
};

}  // namespace

extern "C" {

void* orc_sim_create(const dsph_case* cs, int prec, unsigned np, const unsigned* idp,
                     const double* pos, const float* velrho, const float* boundnor) {
  Sim* s = new Sim();
  s->prec = prec;
  try {
    s->config(*cs);
    if (s->TBoundary == 2) {
      // Normals of the boundary particles (id order), stored doubled
      // (JSph::ConfigBoundNormals, JSph.cpp:1385-1434).
      if (!boundnor) throw std::runtime_error("mDBC needs boundary normals");
      s->boundnor.assign(3 * size_t(np), 0.f);
      for (unsigned p = 0; p < np; p++)
        if (idp[p] < cs->casenbound)
          for (int k = 0; k < 3; k++) s->boundnor[3 * size_t(p) + k] = boundnor[3 * size_t(idp[p]) + k] * 2.f;
    }
    s->load(np, idp, pos, velrho);
  } catch (const std::exception& e) {
    s->err = e.what();
  }
  return s;
}
void orc_sim_destroy(void* h) { delete static_cast<Sim*>(h); }
// The 4x4 matrix (row-major) of JMatrix4::MatrixRot(ang, p1, p2).
void orc_matrix_rot(double ang, const double* p1, const double* p2, double* m16) {
  Sim::mat_rot(ang, p1, p2, m16);
}
const char* orc_sim_error(void* h) { return static_cast<Sim*>(h)->err.c_str(); }
int orc_sim_run(void* h, unsigned nsteps) {
  Sim* s = static_cast<Sim*>(h);
  if (!s->err.empty()) return 1;
  try {
    s->run(nsteps);
  } catch (const std::exception& e) {
    s->err = e.what();
    return 1;
  }
  return 0;
}
// Downloads state (sorted order): np, idp, pos(double3), velrho, code.
unsigned orc_sim_download(void* h, unsigned* idp, double* pos, float* velrho, uint16_t* code) {
  Sim* s = static_cast<Sim*>(h);
  const size_t n = s->Np;
  if (idp) std::memcpy(idp, s->idp.data(), 4 * n);
  if (pos) std::memcpy(pos, s->pos.data(), 24 * n);
  if (velrho) std::memcpy(velrho, s->velrho.data(), 16 * n);
  if (code) std::memcpy(code, s->code.data(), 2 * n);
  return s->Np;
}
void orc_sim_forces(void* h, float* ace, float* ar) {
  Sim* s = static_cast<Sim*>(h);
  if (ace) std::memcpy(ace, s->ace.data(), 12 * size_t(s->Np));
  if (ar) std::memcpy(ar, s->ar.data(), 4 * size_t(s->Np));
}
// {np, npb, npbok, nct, nstep, npfouttotal}, {timestep, dtlast, symdtpre, velmax, acemax, viscdtmax}
void orc_sim_stats(void* h, unsigned* u6, double* d6) {
  Sim* s = static_cast<Sim*>(h);
  if (u6) {
    u6[0] = s->Np; u6[1] = s->Npb; u6[2] = s->NpbOk; u6[3] = s->Nct; u6[4] = s->Nstep;
    u6[5] = s->NpfOutTotal;
  }
  if (d6) {
    d6[0] = s->TimeStep; d6[1] = s->DtLast; d6[2] = s->SymplecticDtPre; d6[3] = s->VelMax;
    d6[4] = s->AceMax; d6[5] = s->ViscDtMax;
  }
}
// Cell-division view: cellmin/cellmax (6 u32), full flag; begincell copy of
// 2*nctt ints when `be` is not NULL. Returns nctt.
unsigned orc_sim_celldiv(void* h, unsigned* c6, int* be) {
  Sim* s = static_cast<Sim*>(h);
  if (c6) for (int k = 0; k < 3; k++) { c6[k] = s->cmin[k]; c6[3 + k] = s->cmax[k]; }
  if (be) std::memcpy(be, s->begincell.data(), 4 * s->begincell.size());
  return unsigned(s->begincell.size() / 2);
}
void orc_sim_pips(void* h, unsigned long long res[4]) {
  Sim* s = static_cast<Sim*>(h);
  const orc_div d = s->divdata();
  orc_pips(&s->c, &d, s->Np, s->Npb, s->NpbOk, s->dcell.data(), s->poscell.data(), res);
}
// Re-runs the interaction on the current state (forces of the current state).
void orc_sim_interact(void* h) { static_cast<Sim*>(h)->interaction(); }

// State of floating body cf (-1 when there is none).
int orc_sim_floating(void* h, unsigned cf, dsph_floating_state* out) {
  const Sim* s = static_cast<Sim*>(h);
  if (cf >= s->ft.size()) return -1;
  const auto& f = s->ft[cf];
  for (int k = 0; k < 3; k++) {
    out->center[k] = f.center[k];
    out->angles[k] = f.angles[k]; out->fvel[k] = f.fvel[k]; out->fomega[k] = f.fomega[k];
    out->facelin[k] = f.facelin[k]; out->faceang[k] = f.faceang[k];
  }
  return 0;
}
// Neighbour-list arrays of the current state: dcell, poscell (float4), and
// the constants used (orc_cte).
void orc_sim_arrays(void* h, unsigned* dcell, float* poscell, orc_cte* cte) {
  Sim* s = static_cast<Sim*>(h);
  if (dcell) std::memcpy(dcell, s->dcell.data(), 4 * size_t(s->Np));
  if (poscell) std::memcpy(poscell, s->poscell.data(), 16 * size_t(s->Np));
  if (cte) *cte = s->c;
}

}  // extern "C"
