// Oracle for the SPH step: a CPU restatement (C++17/OpenMP) of the
// reference's algorithm (DualSPHysics v5.4.355, paths relative to
// src/source/).
//
// TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load it, as the checker or as the timed CPU
// baseline; the product path (dualsphysics_amd) never calls it.
//
// Each function cites the reference code it restates. Pair arithmetic is
// templated on the real type: float reproduces the reference's FP32 pair
// arithmetic, double is the FP64 yardstick every parity result is measured
// against. Three distance/EOS/accumulation semantics are provided:
//   SEM_GPU : distances from cell-relative FP32 PosCell, EOS rho*(1/rho0),
//             fluid and boundary neighbours in one accumulator set
//             (KerInteractionForcesFluid, JSphGpu_ker.cu:786-990);
//   SEM_CPU : distances float(FP64 difference), EOS rho/rho0, fluid and
//             boundary neighbours in two passes, density diffusion through
//             the delta array (JSphCpu.cpp:733-1141,
//             1315-1372);
//   SEM_F64 : distances in FP64 from FP64 positions (yardstick).
//
// Pinning of each part (what checks it, in tests/): cell code, keys, sort,
// BeginEndCell and gathers against brute force and the reference CPU build's
// own outputs; kernels and EOS against closed forms; the interaction against
// O(N^2) brute-force FP64 sums and against the reference CPU build
// (oracle/_ref) on identical particles; integrators against the reference
// CPU build's particle states. See tests/test_oracle_*.py.
#include <omp.h>
#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

typedef uint16_t typecode;

namespace {

// ---- code and cell bits (DualSphDef.h:150-241, JDsDcellDef.h:28-45, DualSphDef.h:661-700)
constexpr unsigned kSpecial = 0xe000u, kOutIgnore = 0x4000u, kOutMov = 0x6000u,
                   kOutPos = 0x8000u, kOutRho = 0xa000u, kTypeMask = 0x1800u,
                   kTypeFloating = 0x1000u, kTypeFluid = 0x1800u;
inline unsigned cx_of(unsigned dcc, unsigned c) { return c >> ((dcc >> 10) & 31u); }
inline unsigned cy_of(unsigned dcc, unsigned c) { return (c << (dcc >> 25)) >> ((dcc >> 5) & 31u); }
inline unsigned cz_of(unsigned dcc, unsigned c) { return (c << (dcc & 31u)) >> (dcc & 31u); }
inline unsigned mkcell(unsigned dcc, unsigned x, unsigned y, unsigned z) {
  return (x << ((dcc >> 10) & 31u)) | (y << ((dcc >> 15) & 31u)) | z;
}
inline unsigned fbits(float f) { unsigned u; std::memcpy(&u, &f, 4); return u; }
inline float bitsf(unsigned u) { float f; std::memcpy(&f, &u, 4); return f; }

}  // namespace

extern "C" {

// Constants of one case (JSph members / StCteInteraction, JSphGpu_cte.h:27-62).
struct orc_cte {
  float massb, massf, kernelh, kernelsize2, poscellsize, awen, bwenh, cs0, eta2, ddtkh, ddtgz;
  float scell, kernelsize, dp, cteb, gamma, rhopzero, ovrhopzero, movlimit;
  unsigned cellcode;
  double domposmin[3], maprealposmin[3], maprealsize[3];
  double mapposmin[3];
  unsigned tboundary;  // 2 = mDBC; otherwise DBC
  unsigned periactive;        // periodic axes, bits 1 X, 2 Y, 4 Z
  double perinc[3][3];         // PeriXinc, PeriYinc, PeriZinc
  unsigned sim2d;              // 2-D simulation: mDBC systems over (1, x, z)
};

// Neighbour-search data (StDivDataCpu / StDivDataGpu).
struct orc_div {
  int scelldiv, ncx, ncy, ncz, ncw;
  unsigned cellfluid;
  int czx, czy, czz;
  const int* begincell;  // int2 {begin,end} per cell (canonical form)
};

// JDsDcell::CalcCellCode (JDsDcell.cpp:30-67).
unsigned orc_cell_code(unsigned ncx, unsigned ncy, unsigned ncz) {
  auto bits = [](unsigned v) { unsigned n = 2; for (; v >> n; n++) {} return n; };
  unsigned sx = bits(ncx), sy = bits(ncy), sz = bits(ncz);
  if (sx + sy + sz > 31) return 0;
  unsigned rest = 31 - (sx + sy + sz);
  while (rest) {
    if (rest) { sx++; rest--; }
    if (rest) { sy++; rest--; }
    if (rest) { sz++; rest--; }
  }
  return ((sx + 1) << 25) | (sy << 20) | (sz << 15) | ((sy + sz) << 10) | ((sx + 1 + sz) << 5) |
         (sx + 1 + sy);
}

// Cell of each particle from FP64 positions (JSph::LoadDcellParticles,
// JSph.cpp:1794-1816; JSphCpu::UpdatePos, JSphCpu.cpp:1507-1515).
void orc_dcell(unsigned n, const double* pos, const double domposmin[3], float scell,
               unsigned dcc, unsigned* dcell) {
  for (unsigned p = 0; p < n; p++) {
    const double* x = pos + 3 * size_t(p);
    dcell[p] = mkcell(dcc, unsigned((x[0] - domposmin[0]) / scell),
                      unsigned((x[1] - domposmin[1]) / scell), unsigned((x[2] - domposmin[2]) / scell));
  }
}

// Min/max cell of the non-excluded particles [pini,pini+n)
// (JCellDivCpu::LimitsCellBound/Fluid, JCellDivCpu.cpp:269-377).
void orc_limits(unsigned n, unsigned pini, unsigned dcc, const unsigned* dcell,
                const typecode* code, unsigned cmin[3], unsigned cmax[3]) {
  cmin[0] = cmin[1] = cmin[2] = UINT_MAX;
  cmax[0] = cmax[1] = cmax[2] = 0;
  for (unsigned p = pini; p < pini + n; p++) {
    if ((code[p] & kSpecial) >= kOutIgnore) continue;
    const unsigned c = dcell[p];
    const unsigned v[3] = {cx_of(dcc, c), cy_of(dcc, c), cz_of(dcc, c)};
    for (int k = 0; k < 3; k++) { cmin[k] = std::min(cmin[k], v[k]); cmax[k] = std::max(cmax[k], v[k]); }
  }
}

// Cell keys (JCellDivCpuSingle::PreSortFull / PreSortFluid,
// JCellDivCpuSingle.cpp:137-205). full!=0: all [0,np); else fluid [pini,np).
void orc_presort(int full, unsigned np, unsigned pini, unsigned dcc, const unsigned* dcell,
                 const typecode* code, const unsigned cellmin[3], const unsigned nc[3],
                 unsigned* cellpart) {
  const unsigned nsheet = nc[0] * nc[1], nct = nsheet * nc[2];
  const unsigned bbi = nct, bf = nct + 1, bbo = bf + nct, bfo = bbo + 1, bboi = bfo + 1,
                 bfoi = bboi + 1;
  for (unsigned p = (full ? 0 : pini); p < np; p++) {
    const unsigned c = dcell[p];
    const unsigned x = cx_of(dcc, c) - cellmin[0], y = cy_of(dcc, c) - cellmin[1],
                   z = cz_of(dcc, c) - cellmin[2];
    const unsigned key = x + y * nc[0] + z * nsheet;
    const unsigned type = code[p] & kTypeMask, out = code[p] & kSpecial;
    unsigned box;
    if (full && type < kTypeFloating)
      box = out < kOutIgnore ? ((x < nc[0] && y < nc[1] && z < nc[2]) ? key : bbi)
                             : (out == kOutIgnore ? bboi : bbo);
    else
      box = out <= kOutIgnore ? (out < kOutIgnore ? bf + key : bfoi)
                              : (type == kTypeFloating ? bbo : bfo);
    cellpart[p] = box;
  }
}

// Stable counting sort by key over [pini,np) (JCellDivCpuSingle::MakeSortFull /
// MakeSortFluid, JCellDivCpuSingle.cpp:207-262): sortpart[i] is the index of the
// particle that goes to slot i; keys sorted in place; begin/end of every key in
// [key0, nkeys) written as int2 {begin,end} ((0,0) when empty, as the GPU).
void orc_sort_cells(unsigned np, unsigned pini, unsigned nkeys, unsigned key0, unsigned* cellpart,
                    unsigned* sortpart, int* beginend) {
  std::vector<unsigned> cnt(nkeys + 1, 0), beg(nkeys + 1, 0);
  for (unsigned p = pini; p < np; p++) cnt[cellpart[p]]++;
  beg[key0] = pini;
  for (unsigned k = key0; k < nkeys; k++) beg[k + 1] = beg[k] + cnt[k];
  std::vector<unsigned> fill(beg.begin(), beg.end());
  std::vector<unsigned> sortedkeys(np);
  for (unsigned p = 0; p < pini; p++) sortpart[p] = p;
  for (unsigned p = pini; p < np; p++) {
    const unsigned k = cellpart[p];
    sortpart[fill[k]] = p;
    sortedkeys[fill[k]] = k;
    fill[k]++;
  }
  for (unsigned p = pini; p < np; p++) cellpart[p] = sortedkeys[p];
  if (beginend)
    for (unsigned k = key0; k < nkeys; k++) {
      beginend[2 * size_t(k)] = cnt[k] ? int(beg[k]) : 0;
      beginend[2 * size_t(k) + 1] = cnt[k] ? int(beg[k + 1]) : 0;
    }
}

// PosCell of FP64 positions (KerUpdatePosCell, JSphGpuSimple_ker.cu:41-59).
void orc_poscell(unsigned n, const double* pos, const double posmin[3], float pcs, float* poscell) {
  for (unsigned p = 0; p < n; p++) {
    const double* x = pos + 3 * size_t(p);
    const double dx = x[0] - posmin[0], dy = x[1] - posmin[1], dz = x[2] - posmin[2];
    const unsigned cx = unsigned(dx / pcs), cy = unsigned(dy / pcs), cz = unsigned(dz / pcs);
    float* o = poscell + 4 * size_t(p);
    o[0] = float(dx - double(pcs) * cx);
    o[1] = float(dy - double(pcs) * cy);
    o[2] = float(dz - double(pcs) * cz);
    o[3] = bitsf((cx << 19) | (cy << 9) | cz);
  }
}

}  // extern "C"

namespace {

enum { SEM_GPU = 0, SEM_CPU = 1, SEM_F64 = 2 };

// Power rounded once to R: for R=float the correctly rounded FP32 power (the
// FP64 power rounded to float), the common target of every FP32 path since
// the reference's own powf is platform dependent.
template <typename R>
inline R crpow(R x, R y) {
  return R(std::pow(double(x), double(y)));
}

// Wendland C2 (FunSphKernel.h:198-244) in precision R.
template <typename R>
inline R wend_fac(R rr2, R h, R bwenh) {
  const R rad = std::sqrt(rr2);
  const R qq = rad / h;
  const R w1 = R(1) - R(0.5) * qq;
  return bwenh * w1 * w1 * w1;
}
template <typename R>
inline R wend_wab(R rr2, R h, R awen) {
  const R rad = std::sqrt(rr2);
  const R qq = rad / h;
  const R wqq = qq + qq + R(1);
  const R w1 = R(1) - R(0.5) * qq;
  const R w2 = w1 * w1;
  return awen * wqq * w2 * w2;
}

struct Opts {
  int sem, tvisco, tddt, velgrad;
  float viscob, viscof;
  // mDBC2 (SlipMode no-slip): mode of each boundary particle (0 DBC, 1 mDBC2,
  // 2 mDBC2 switched off) and its tangential velocity; NULL without mDBC2.
  const unsigned char* bmode = nullptr;
  const float* tangenvel = nullptr;
  // Shifting (JSphCpu.cpp:943-950): mode 1 no boundaries, 2 no fixed
  // boundaries, 3 full; codes of the neighbours; 0 = off.
  int shiftmode = 0;
  const typecode* code = nullptr;
  // Floating bodies: mass of the particles of each body (FtObjs[].massp,
  // indexed by the code's type value) and the particle codes; NULL = none.
  const float* ftmassp = nullptr;
  // mDBC2 no penetration (MDBC2_NoPen): boundary normals (float3 per
  // particle) and the per-particle output (float4); NULL = off.
  const float* boundnor = nullptr;
  float* nopenshift = nullptr;
  // Motion velocity of the boundary particles (MotionVel, float3 per
  // particle) for the no-penetration sums; NULL = fixed boundaries (0).
  const float* motionvel = nullptr;
  // Particles whose forces are computed: fluid [fbeg, fend) and boundary
  // [bbeg, bend), clipped to [npb, np) and [0, npbok); the default is all.
  long fbeg = 0, fend = LONG_MAX, bbeg = 0, bend = LONG_MAX;
};

// Floating particle test and its body (CODE_IsFloating, CODE_GetTypeValue,
// DualSphDef.h:164-200).
inline bool is_floating(const Opts& o, unsigned p) {
  return o.ftmassp && (o.code[p] & kTypeMask) == kTypeFloating;
}
inline float ft_massp(const Opts& o, unsigned p) { return o.ftmassp[o.code[p] & 0x7ffu]; }

// Relative position p1-p2 in the chosen semantics.
template <typename R>
inline void rel(int sem, const orc_cte& c, const float* ps, const double* pos, unsigned p1,
                unsigned p2, R& dx, R& dy, R& dz) {
  if (sem == SEM_GPU) {
    const float* a = ps + 4 * size_t(p1);
    const float* b = ps + 4 * size_t(p2);
    const unsigned ua = fbits(a[3]), ub = fbits(b[3]);
    const int ax = int(ua >> 19), ay = int((ua & 0x0007fe00u) >> 9), az = int(ua & 0x1ffu);
    const int bx = int(ub >> 19), by = int((ub & 0x0007fe00u) >> 9), bz = int(ub & 0x1ffu);
    dx = R(a[0] - b[0] + c.poscellsize * float(ax - bx));
    dy = R(a[1] - b[1] + c.poscellsize * float(ay - by));
    dz = R(a[2] - b[2] + c.poscellsize * float(az - bz));
  } else if (sem == SEM_CPU) {
    dx = R(float(pos[3 * size_t(p1)] - pos[3 * size_t(p2)]));
    dy = R(float(pos[3 * size_t(p1) + 1] - pos[3 * size_t(p2) + 1]));
    dz = R(float(pos[3 * size_t(p1) + 2] - pos[3 * size_t(p2) + 2]));
  } else {
    dx = R(pos[3 * size_t(p1)] - pos[3 * size_t(p2)]);
    dy = R(pos[3 * size_t(p1) + 1] - pos[3 * size_t(p2) + 1]);
    dz = R(pos[3 * size_t(p1) + 2] - pos[3 * size_t(p2) + 2]);
  }
}

struct Row { unsigned a, b; };

// Cell rows around cell (cx,cy,cz) of the current domain (nsearch::Init +
// ParticleRange, JCellSearch_inline.h:33-85; GPU form JCellSearch_iker.h:34-166).
inline int rows_at(const orc_div& d, int cx, int cy, int cz, bool fluidcells, Row* out) {
  const int s = d.scelldiv;
  const int x0 = cx - std::min(cx, s), x1 = cx + std::min(d.ncx - cx - 1, s) + 1;
  const int y0 = cy - std::min(cy, s), y1 = cy + std::min(d.ncy - cy - 1, s) + 1;
  const int z0 = cz - std::min(cz, s), z1 = cz + std::min(d.ncz - cz - 1, s) + 1;
  const int base = fluidcells ? int(d.cellfluid) : 0;
  int nr = 0;
  for (int z = z0; z < z1; z++)
    for (int y = y0; y < y1; y++) {
      unsigned a = 0, b = 0;
      const int rb = base + z * d.ncw + y * d.ncx;
      for (int x = x0; x < x1; x++) {
        const int* be = d.begincell + 2 * size_t(rb + x);
        if (be[1]) {
          if (!b) a = unsigned(be[0]);
          b = unsigned(be[1]);
        }
      }
      out[nr++] = Row{a, b};
    }
  return nr;
}

// Fluid particle p1 against the rows of one kind (JSphCpu.cpp:733-1141 /
// KerInteractionForcesFluidBox JSphGpu_ker.cu:539-777).
template <typename R>
struct FAcc {
  R ax = 0, ay = 0, az = 0, ar = 0, visc = 0, delta = 0;
  R s[6] = {0, 0, 0, 0, 0, 0};
  R g[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  bool delta_off = false;
  float sh[4] = {0, 0, 0, 0};  // shifting sums in FP32 (x = FLT_MAX: cancelled)
  float npc[3] = {0, 0, 0}, nps[3] = {0, 0, 0};  // no-penetration counts and sums (FP32)
};

template <typename R>
void fluid_rows(const Opts& o, const orc_cte& c, bool boundp2, const Row* rows, int nr,
                unsigned p1, const float* ps, const double* pos, const float* velrho,
                const R* press, const float* tau, FAcc<R>& a) {
  const R massp2_kind = R(boundp2 ? c.massb : c.massf);
  const R visco = R(boundp2 ? o.viscob : o.viscof);
  const bool mdbc2 = boundp2 && o.bmode;
  const bool ftp1 = is_floating(o, p1);
  const float* v1 = velrho + 4 * size_t(p1);
  const R v1x = R(v1[0]), v1y = R(v1[1]), v1z = R(v1[2]), rho1 = R(v1[3]);
  const R press1 = press[p1];
  const R ks2 = R(c.kernelsize2), eta2 = R(c.eta2), h = R(c.kernelh), bwenh = R(c.bwenh);
  const R cbar = R(c.cs0), ddtkh = R(c.ddtkh), ddtgz = R(c.ddtgz), rho0 = R(c.rhopzero);
  const R invgamma = R(1) / R(c.gamma);
  for (int r = 0; r < nr; r++)
    for (unsigned p2 = rows[r].a; p2 < rows[r].b; p2++) {
      R dx, dy, dz;
      rel<R>(o.sem, c, ps, pos, p1, p2, dx, dy, dz);
      const R rr2 = dx * dx + dy * dy + dz * dz;
      if (!(rr2 <= ks2 && rr2 >= R(1e-18f))) continue;
      const R fac = wend_fac<R>(rr2, h, bwenh);
      const R frx = fac * dx, fry = fac * dy, frz = fac * dz;
      const float* v2 = velrho + 4 * size_t(p2);
      const R rho2 = R(v2[3]);
      // mDBC2: a boundary particle whose ghost node is dry has no mass
      // (JSphCpu.cpp:851-854; JSphGpu_ker.cu:585-588). A floating neighbour
      // has its body's particle mass and stops the Molteni-Colagrossi term
      // when it is not heavier than 1.2 fluid particles (DELTA_HEAVYFLOATING,
      // JSphCpu.cpp:838-846; JSphGpu_ker.cu:570-580).
      const bool ftp2 = !boundp2 && is_floating(o, p2);
      R massp2 = (mdbc2 && o.bmode[p2] == 2) ? R(0) : massp2_kind;
      if (ftp2) {
        massp2 = R(ft_massp(o, p2));
        if (o.tddt == 1 && float(massp2) <= c.massf * 1.2f) a.delta_off = true;
      }
      {
        const R prs = (press1 + press[p2]) / (rho1 * rho2);
        const R p_vpm = -prs * massp2;
        a.ax += p_vpm * frx; a.ay += p_vpm * fry; a.az += p_vpm * frz;
      }
      R dvx = v1x - R(v2[0]), dvy = v1y - R(v2[1]), dvz = v1z - R(v2[2]);
      a.ar += massp2 * (dvx * frx + dvy * fry + dvz * frz) * (rho1 / rho2);
      const R dot3 = dx * frx + dy * fry + dz * frz;
      if (o.tddt == 1 && !a.delta_off) {
        const R visc_densi = ddtkh * cbar * (rho1 / rho2 - R(1)) / (rr2 + eta2);
        // A boundary neighbour cancels the term with DBC only
        // (JSphCpu.cpp:930; JSphGpu_ker.cu:665).
        if (boundp2 && c.tboundary != 2) a.delta_off = true;
        else a.delta += visc_densi * dot3 * massp2;
      }
      if ((o.tddt == 2 || (o.tddt == 3 && !boundp2)) && !a.delta_off && !ftp2) {
        const R rh = R(1) + ddtgz * dz;
        const R drho = rho0 * crpow<R>(rh, invgamma) - rho0;
        const R visc_densi = ddtkh * cbar * ((rho2 - rho1) - drho) / (rr2 + eta2);
        const R delta = visc_densi * dot3 * massp2 / rho2;
        if (boundp2) a.delta_off = true;
        else a.delta -= delta;
      }
      if (o.shiftmode && a.sh[0] != FLT_MAX) {
        const float massrho = float(massp2) / v2[3];
        const bool noshift = boundp2 && (o.shiftmode == 1 ||
                                         (o.shiftmode == 2 && (o.code[p2] & kTypeMask) == 0));
        const float fx = float(frx), fy = float(fry), fz = float(frz);
        a.sh[0] = noshift ? FLT_MAX : a.sh[0] + massrho * fx;
        a.sh[1] += massrho * fy;
        a.sh[2] += massrho * fz;
        a.sh[3] -= massrho * float(dot3);
      }
      if (boundp2 && o.nopenshift && !ftp2) {
        // No penetration (JSphCpu.cpp:952-974 with ComputeNoPenVel, :716-726;
        // JSphGpu_ker.cu:686-707): a first-layer boundary neighbour
        // (|normal| < 1.75 dp) within 1.25 dp with the particle less than
        // 0.75 |normal| away along the normal; per axis, a particle moving
        // towards the wall (velocity relative to the wall's motion velocity)
        // counts it and gets -(3 - 4 max(|dr/n|, 1/4)) v n^2. FP32, as both
        // builds.
        const float fdx = float(dx), fdy = float(dy), fdz = float(dz);
        const float rrmag = std::sqrt(fdx * fdx + fdy * fdy + fdz * fdz);
        if (rrmag < 1.25f * c.dp) {
          const float* bn = o.boundnor + 3 * size_t(p2);
          const float norm = std::sqrt(bn[0] * bn[0] + bn[1] * bn[1] + bn[2] * bn[2]);
          const float n[3] = {bn[0] / norm, bn[1] / norm, bn[2] / norm};
          const float dr[3] = {fdx, fdy, fdz};
          const float* mv = o.motionvel ? o.motionvel + 3 * size_t(p2) : nullptr;
          const float v[3] = {mv ? float(v1[0]) - mv[0] : float(v1[0]),
                              mv ? float(v1[1]) - mv[1] : float(v1[1]),
                              mv ? float(v1[2]) - mv[2] : float(v1[2])};
          const float normdist = n[0] * fdx + n[1] * fdy + n[2] * fdz;
          if (normdist < 0.75f * norm && norm < 1.75f * c.dp) {
            for (int k = 0; k < 3; k++) {
              if (!(dr[k] * n[k] < 0.75f && std::fabs(n[k]) > 0.001f * c.dp)) continue;
              const float vfc = v[k] * n[k];
              if (vfc < 0.f) {
                const float ratio = std::max(std::fabs(dr[k] / n[k]), 0.25f);
                const float factor = -4.f * ratio + 3.f;
                a.npc[k] += 1.f;
                a.nps[k] -= factor * v[k] * n[k] * n[k];
              }
            }
          }
        }
      }
      if (mdbc2) {
        // Viscous terms against the no-slip tangential velocity
        // (JSphCpu.cpp:988-995; JSphGpu_ker.cu:725-730).
        const float* tv = o.tangenvel + 3 * size_t(p2);
        dvx = v1x - R(tv[0]); dvy = v1y - R(tv[1]); dvz = v1z - R(tv[2]);
      }
      const R dot = dx * dvx + dy * dvy + dz * dvz;
      const R dot_rr2 = dot / (rr2 + eta2);
      a.visc = std::max(dot_rr2, a.visc);
      if (o.tvisco == 1) {
        if (dot < 0) {
          const R amubar = h * dot_rr2;
          const R robar = (rho1 + rho2) * R(0.5);
          const R pi_visc = (-visco * cbar * amubar / robar) * massp2;
          a.ax -= pi_visc * frx; a.ay -= pi_visc * fry; a.az -= pi_visc * frz;
        }
      } else {
        const R robar2 = rho1 + rho2;
        const R temp = R(4) * visco / ((rr2 + eta2) * robar2);
        const R vtemp = massp2 * temp * dot3;
        a.ax += vtemp * dvx; a.ay += vtemp * dvy; a.az += vtemp * dvz;
      }
      if (o.tvisco == 2) {
        const float* t1 = tau + 6 * size_t(p1);
        R t[6] = {R(t1[0]), R(t1[1]), R(t1[2]), R(t1[3]), R(t1[4]), R(t1[5])};
        // tau of the neighbour when it is fluid; the strain of p1 when p1 is
        // fluid (JSphCpu.cpp:1019-1033).
        if (!boundp2 && !ftp2) {
          const float* t2 = tau + 6 * size_t(p2);
          for (int k = 0; k < 6; k++) t[k] += R(t2[k]);
        }
        a.ax += massp2 * (t[0] * frx + t[1] * fry + t[2] * frz);
        a.ay += massp2 * (t[1] * frx + t[3] * fry + t[4] * frz);
        a.az += massp2 * (t[2] * frx + t[4] * fry + t[5] * frz);
        if (!ftp1) {
          const R volp2 = -massp2 / rho2;
          R dv = dvx * volp2; a.s[0] += dv * frx; a.s[1] += dv * fry; a.s[2] += dv * frz;
          dv = dvy * volp2;   a.s[1] += dv * frx; a.s[3] += dv * fry; a.s[4] += dv * frz;
          dv = dvz * volp2;   a.s[2] += dv * frx; a.s[4] += dv * fry; a.s[5] += dv * frz;
        }
      }
      if (o.velgrad) {
        const R volp2 = -massp2 / rho2;
        const R ux = dvx * volp2, uy = dvy * volp2, uz = dvz * volp2;
        a.g[0] += ux * frx; a.g[1] += ux * fry; a.g[2] += ux * frz;
        a.g[3] += uy * frx; a.g[4] += uy * fry; a.g[5] += uy * frz;
        a.g[6] += uz * frx; a.g[7] += uz * fry; a.g[8] += uz * frz;
      }
    }
}

// Tait EOS (FunSphEos.h:38-66 / FunSphEos_iker.h:37-59). The CPU build's
// ComputePressMonaghan calls the C library's pow(double, double) on the FP32
// ratio rho/rho0, so its pressure is B*(x^gamma - 1) evaluated in FP64 and
// rounded once to FP32 (this reproduces every pressure of the reference CPU
// build bit for bit); the GPU build evaluates the same in FP32.
template <typename R>
R eos(int sem, const orc_cte& c, float rho) {
  if (sem == SEM_GPU)
    return R(c.cteb) * (crpow<R>(R(rho) * R(c.ovrhopzero), R(c.gamma)) - R(1));
  if (sem == SEM_CPU && sizeof(R) == sizeof(float)) {
    const float x = rho / c.rhopzero;
    return R(float(double(c.cteb) * (std::pow(double(x), double(c.gamma)) - 1.0)));
  }
  return R(c.cteb) * (crpow<R>(R(rho) / R(c.rhopzero), R(c.gamma)) - R(1));
}

// Mean no-penetration correction per axis over the boundary neighbours that
// asked for one; w = 10 marks a particle to correct, 0 none (the array is
// cleared before each interaction: JSphCpu.cpp:519, 1103-1123).
template <typename R>
void store_nopen(const FAcc<R>& a, float* ns) {
  ns[0] = ns[1] = ns[2] = ns[3] = 0.f;
  for (int k = 0; k < 3; k++)
    if (a.npc[k] > 0.f) {
      ns[k] = a.nps[k] / a.npc[k];
      ns[3] = 10.f;
    }
}

template <typename R>
void interaction(const Opts& o, const orc_cte& c, const orc_div& d, unsigned np, unsigned npb,
                 unsigned npbok, const unsigned* dcell, const float* ps, const double* pos,
                 const float* velrho, const float* tau, float* ace, float* ar, float* viscdt,
                 float* strain, float* velgrad, float* shiftposfs = nullptr) {
  std::vector<R> press(np);
#pragma omp parallel for schedule(static)
  for (long p = 0; p < long(np); p++) press[p] = eos<R>(o.sem, c, velrho[4 * size_t(p) + 3]);
  // Fluid particles.
  const long f0 = std::max(long(npb), o.fbeg), f1 = std::min(long(np), o.fend);
  const long b0 = std::max(0L, o.bbeg), b1 = std::min(long(npbok), o.bend);
#pragma omp parallel for schedule(guided)
  for (long pp = f0; pp < f1; pp++) {
    const unsigned p1 = unsigned(pp);
    Row rf[25], rb[25];
    const unsigned cel = dcell[p1];
    const int cx = int(cx_of(c.cellcode, cel)) - d.czx, cy = int(cy_of(c.cellcode, cel)) - d.czy,
              cz = int(cz_of(c.cellcode, cel)) - d.czz;
    const int nf = rows_at(d, cx, cy, cz, true, rf);
    const int nb = rows_at(d, cx, cy, cz, false, rb);
    R oax, oay, oaz, oar, ovisc;
    R os[6], og[9];
    if (o.sem == SEM_CPU) {
      // Two passes, each stored separately (JSphCpu.cpp:1331-1349). The
      // density-diffusion term goes to the per-particle delta array of the CPU
      // build (DDTArray, JSph.cpp:928; JSphCpu.cpp:1086): a boundary neighbour
      // in either pass drops the whole term, as on the GPU; otherwise the two
      // partial terms are summed and added to ar after both passes
      // (AddDelta, JSphCpuSingle.cpp:726).
      FAcc<R> a1, a2;
      // No density diffusion for floating particles (JSphCpu.cpp:774-778).
      a1.delta_off = a2.delta_off = is_floating(o, p1) && o.tddt != 0;
      if (shiftposfs) for (int k = 0; k < 4; k++) a1.sh[k] = shiftposfs[4 * size_t(p1) + k];
      fluid_rows<R>(o, c, false, rf, nf, p1, ps, pos, velrho, press.data(), tau, a1);
      // The CPU build stores the shifting sums after the fluid pass and
      // continues from them in the boundary pass (JSphCpu.cpp:769, 1101).
      for (int k = 0; k < 4; k++) a2.sh[k] = a1.sh[k];
      fluid_rows<R>(o, c, true, rb, nb, p1, ps, pos, velrho, press.data(), tau, a2);
      if (shiftposfs) for (int k = 0; k < 4; k++) shiftposfs[4 * size_t(p1) + k] = a2.sh[k];
      if (o.nopenshift) store_nopen(a2, o.nopenshift + 4 * size_t(p1));
      oax = R(float(a1.ax)) + a2.ax; oay = R(float(a1.ay)) + a2.ay; oaz = R(float(a1.az)) + a2.az;
      oar = R(float(R(float(a1.ar)) + a2.ar));
      if (o.tddt && !a1.delta_off && !a2.delta_off)
        oar = oar + R(float(R(float(a1.delta)) + a2.delta));
      ovisc = std::max(a1.visc, a2.visc);
      for (int k = 0; k < 6; k++) os[k] = a1.s[k] + a2.s[k];
      for (int k = 0; k < 9; k++) og[k] = a1.g[k] + a2.g[k];
    } else {
      FAcc<R> a;
      a.delta_off = is_floating(o, p1) && o.tddt != 0;  // JSphGpu_ker.cu:838-843
      if (shiftposfs) for (int k = 0; k < 4; k++) a.sh[k] = shiftposfs[4 * size_t(p1) + k];
      fluid_rows<R>(o, c, false, rf, nf, p1, ps, pos, velrho, press.data(), tau, a);
      fluid_rows<R>(o, c, true, rb, nb, p1, ps, pos, velrho, press.data(), tau, a);
      if (shiftposfs) for (int k = 0; k < 4; k++) shiftposfs[4 * size_t(p1) + k] = a.sh[k];
      if (o.nopenshift) store_nopen(a, o.nopenshift + 4 * size_t(p1));
      oax = a.ax; oay = a.ay; oaz = a.az;
      oar = a.ar + (o.tddt && !a.delta_off ? a.delta : R(0));
      ovisc = a.visc;
      for (int k = 0; k < 6; k++) os[k] = a.s[k];
      for (int k = 0; k < 9; k++) og[k] = a.g[k];
    }
    ace[3 * size_t(p1)] += float(oax);
    ace[3 * size_t(p1) + 1] += float(oay);
    ace[3 * size_t(p1) + 2] += float(oaz);
    ar[p1] += float(oar);
    if (float(ovisc) > viscdt[p1]) viscdt[p1] = float(ovisc);
    if (o.tvisco == 2)
      for (int k = 0; k < 6; k++) strain[6 * size_t(p1) + k] += float(os[k]);
    if (o.velgrad && velgrad)
      for (int k = 0; k < 9; k++) velgrad[9 * size_t(p1) + k] = float(og[k]);
  }
  // Boundary particles (JSphCpu.cpp:633-703 / KerInteractionForcesBound).
  const R ks2 = R(c.kernelsize2), eta2 = R(c.eta2), h = R(c.kernelh), bwenh = R(c.bwenh);
  const R massf = R(c.massf);
#pragma omp parallel for schedule(guided)
  for (long pp = b0; pp < b1; pp++) {
    const unsigned p1 = unsigned(pp);
    Row rf[25];
    const unsigned cel = dcell[p1];
    const int cx = int(cx_of(c.cellcode, cel)) - d.czx, cy = int(cy_of(c.cellcode, cel)) - d.czy,
              cz = int(cz_of(c.cellcode, cel)) - d.czz;
    const int nf = rows_at(d, cx, cy, cz, true, rf);
    const float* v1 = velrho + 4 * size_t(p1);
    R arp = 0, visc = 0;
    for (int r = 0; r < nf; r++)
      for (unsigned p2 = rf[r].a; p2 < rf[r].b; p2++) {
        R dx, dy, dz;
        rel<R>(o.sem, c, ps, pos, p1, p2, dx, dy, dz);
        const R rr2 = dx * dx + dy * dy + dz * dz;
        if (!(rr2 <= ks2 && rr2 >= R(1e-18f))) continue;
        const R fac = wend_fac<R>(rr2, h, bwenh);
        const R frx = fac * dx, fry = fac * dy, frz = fac * dz;
        const float* v2 = velrho + 4 * size_t(p2);
        const R dvx = R(v1[0]) - R(v2[0]), dvy = R(v1[1]) - R(v2[1]), dvz = R(v1[2]) - R(v2[2]);
        // A floating neighbour has its body's particle mass (JSphCpu.cpp:673-676).
        const R massp2 = is_floating(o, p2) ? R(ft_massp(o, p2)) : massf;
        arp += massp2 * (dvx * frx + dvy * fry + dvz * frz) * (R(v1[3]) / R(v2[3]));
        const R dot = dx * dvx + dy * dvy + dz * dvz;
        visc = std::max(dot / (rr2 + eta2), visc);
      }
    ar[p1] += float(arp);
    if (float(visc) > viscdt[p1]) viscdt[p1] = float(visc);
  }
}

}  // namespace

extern "C" {

void orc_interaction_mdbc2(int, int, int, int, float, float, const orc_cte*, const orc_div*,
                           unsigned, unsigned, unsigned, const unsigned*, const float*,
                           const double*, const float*, const float*, float*, float*, float*,
                           float*, float*, const unsigned char*, const float*);

void orc_interaction_ext(int, int, int, int, float, float, const orc_cte*, const orc_div*,
                         unsigned, unsigned, unsigned, const unsigned*, const float*,
                         const double*, const float*, const float*, float*, float*, float*,
                         float*, float*, const unsigned char*, const float*, int,
                         const typecode*, float*);

// Force interaction (see header). prec: 0 float, 1 double. sem: SEM_*.
// ace/ar/viscdt (and strain) accumulate as in the reference.
void orc_interaction(int prec, int sem, int tvisco, int tddt, float viscob, float viscof,
                     const orc_cte* c, const orc_div* d, unsigned np, unsigned npb, unsigned npbok,
                     const unsigned* dcell, const float* poscell, const double* pos,
                     const float* velrho, const float* tau, float* ace, float* ar,
                     float* viscdt, float* strain, float* velgrad) {
  orc_interaction_mdbc2(prec, sem, tvisco, tddt, viscob, viscof, c, d, np, npb, npbok, dcell,
                        poscell, pos, velrho, tau, ace, ar, viscdt, strain, velgrad, nullptr,
                        nullptr);
}

// Force interaction with mDBC2 boundary modes and tangential velocities
// (boundmode/tangenvel of StInterParmsc; NULL = no mDBC2).
void orc_interaction_mdbc2(int prec, int sem, int tvisco, int tddt, float viscob, float viscof,
                           const orc_cte* c, const orc_div* d, unsigned np, unsigned npb,
                           unsigned npbok, const unsigned* dcell, const float* poscell,
                           const double* pos, const float* velrho, const float* tau, float* ace,
                           float* ar, float* viscdt, float* strain, float* velgrad,
                           const unsigned char* bmode, const float* tangenvel) {
  orc_interaction_ext(prec, sem, tvisco, tddt, viscob, viscof, c, d, np, npb, npbok, dcell,
                      poscell, pos, velrho, tau, ace, ar, viscdt, strain, velgrad, bmode,
                      tangenvel, 0, nullptr, nullptr);
}

// Force interaction with every option of this path: mDBC2 modes (bmode,
// tangenvel or NULL) and shifting (shiftmode 1..3, the particle codes and
// the float4 sums per particle, fluid part cleared by the caller; 0 = off).
void orc_interaction_ft(int, int, int, int, float, float, const orc_cte*, const orc_div*,
                        unsigned, unsigned, unsigned, const unsigned*, const float*,
                        const double*, const float*, const float*, float*, float*, float*,
                        float*, float*, const unsigned char*, const float*, int,
                        const typecode*, float*, const float*);

void orc_interaction_ext(int prec, int sem, int tvisco, int tddt, float viscob, float viscof,
                         const orc_cte* c, const orc_div* d, unsigned np, unsigned npb,
                         unsigned npbok, const unsigned* dcell, const float* poscell,
                         const double* pos, const float* velrho, const float* tau, float* ace,
                         float* ar, float* viscdt, float* strain, float* velgrad,
                         const unsigned char* bmode, const float* tangenvel, int shiftmode,
                         const typecode* code, float* shiftposfs) {
  orc_interaction_ft(prec, sem, tvisco, tddt, viscob, viscof, c, d, np, npb, npbok, dcell,
                     poscell, pos, velrho, tau, ace, ar, viscdt, strain, velgrad, bmode,
                     tangenvel, shiftmode, code, shiftposfs, nullptr);
}

// The same with floating bodies: ftmassp = mass of the particles of each body
// (indexed by the code's type value), with the particle codes; NULL = none.
void orc_interaction_np(int, int, int, int, float, float, const orc_cte*, const orc_div*,
                        unsigned, unsigned, unsigned, const unsigned*, const float*,
                        const double*, const float*, const float*, float*, float*, float*,
                        float*, float*, const unsigned char*, const float*, int,
                        const typecode*, float*, const float*, const float*, float*);

void orc_interaction_ft(int prec, int sem, int tvisco, int tddt, float viscob, float viscof,
                        const orc_cte* c, const orc_div* d, unsigned np, unsigned npb,
                        unsigned npbok, const unsigned* dcell, const float* poscell,
                        const double* pos, const float* velrho, const float* tau, float* ace,
                        float* ar, float* viscdt, float* strain, float* velgrad,
                        const unsigned char* bmode, const float* tangenvel, int shiftmode,
                        const typecode* code, float* shiftposfs, const float* ftmassp) {
  orc_interaction_np(prec, sem, tvisco, tddt, viscob, viscof, c, d, np, npb, npbok, dcell,
                     poscell, pos, velrho, tau, ace, ar, viscdt, strain, velgrad, bmode,
                     tangenvel, shiftmode, code, shiftposfs, ftmassp, nullptr, nullptr);
}

void orc_interaction_range(int, int, int, int, float, float, const orc_cte*, const orc_div*,
                           unsigned, unsigned, unsigned, const unsigned*, const float*,
                           const double*, const float*, const float*, float*, float*, float*,
                           float*, float*, const unsigned char*, const float*, int,
                           const typecode*, float*, const float*, const float*, float*, long,
                           long, long, long);

// The same with mDBC2 no penetration: boundnor = float3 boundary normals per
// particle and nopenshift = float4 output per particle (written for the fluid;
// needs bmode), or NULL.
void orc_interaction_np(int prec, int sem, int tvisco, int tddt, float viscob, float viscof,
                        const orc_cte* c, const orc_div* d, unsigned np, unsigned npb,
                        unsigned npbok, const unsigned* dcell, const float* poscell,
                        const double* pos, const float* velrho, const float* tau, float* ace,
                        float* ar, float* viscdt, float* strain, float* velgrad,
                        const unsigned char* bmode, const float* tangenvel, int shiftmode,
                        const typecode* code, float* shiftposfs, const float* ftmassp,
                        const float* boundnor, float* nopenshift) {
  orc_interaction_range(prec, sem, tvisco, tddt, viscob, viscof, c, d, np, npb, npbok, dcell,
                        poscell, pos, velrho, tau, ace, ar, viscdt, strain, velgrad, bmode,
                        tangenvel, shiftmode, code, shiftposfs, ftmassp, boundnor, nopenshift,
                        0, LONG_MAX, 0, LONG_MAX);
}

// The same for a subset of the particles: fluid p1 in [fbeg, fend) and
// boundary p1 in [bbeg, bend) (the rest of the outputs is left alone), so
// that a large state can be checked on a sample of its particles.
void orc_interaction_range_mv(int prec, int sem, int tvisco, int tddt, float viscob,
                              float viscof, const orc_cte* c, const orc_div* d, unsigned np,
                              unsigned npb, unsigned npbok, const unsigned* dcell,
                              const float* poscell, const double* pos, const float* velrho,
                              const float* tau, float* ace, float* ar, float* viscdt,
                              float* strain, float* velgrad, const unsigned char* bmode,
                              const float* tangenvel, int shiftmode, const typecode* code,
                              float* shiftposfs, const float* ftmassp, const float* boundnor,
                              float* nopenshift, const float* motionvel, long fbeg, long fend,
                              long bbeg, long bend) {
  Opts o{sem, tvisco, tddt, velgrad != nullptr, viscob, viscof, bmode, tangenvel,
         shiftposfs ? shiftmode : 0, code, ftmassp && code ? ftmassp : nullptr,
         nopenshift ? boundnor : nullptr, boundnor ? nopenshift : nullptr,
         boundnor && nopenshift ? motionvel : nullptr, fbeg, fend, bbeg, bend};
  if (prec == 0)
    interaction<float>(o, *c, *d, np, npb, npbok, dcell, poscell, pos, velrho, tau, ace, ar,
                       viscdt, strain, velgrad, o.shiftmode ? shiftposfs : nullptr);
  else
    interaction<double>(o, *c, *d, np, npb, npbok, dcell, poscell, pos, velrho, tau, ace, ar,
                        viscdt, strain, velgrad, o.shiftmode ? shiftposfs : nullptr);
}

void orc_interaction_range(int prec, int sem, int tvisco, int tddt, float viscob, float viscof,
                           const orc_cte* c, const orc_div* d, unsigned np, unsigned npb,
                           unsigned npbok, const unsigned* dcell, const float* poscell,
                           const double* pos, const float* velrho, const float* tau,
                           float* ace, float* ar, float* viscdt, float* strain, float* velgrad,
                           const unsigned char* bmode, const float* tangenvel, int shiftmode,
                           const typecode* code, float* shiftposfs, const float* ftmassp,
                           const float* boundnor, float* nopenshift, long fbeg, long fend,
                           long bbeg, long bend) {
  orc_interaction_range_mv(prec, sem, tvisco, tddt, viscob, viscof, c, d, np, npb, npbok, dcell,
                           poscell, pos, velrho, tau, ace, ar, viscdt, strain, velgrad, bmode,
                           tangenvel, shiftmode, code, shiftposfs, ftmassp, boundnor, nopenshift,
                           nullptr, fbeg, fend, bbeg, bend);
}

// mDBC2 with moving boundaries (JSph*::RunMotion, CopyMotionVelAce,
// JSphCpu_mdbc.cpp:472-490, JSphGpu_mdbc_iker.cu:587-603): for the boundary
// particles [0, npb) with ids in [idbegin, idbegin+count), motion
// acceleration float((v - motionvel)/dt) in FP64 and motion velocity v, after
// the motion of the step.
void orc_motion_vel_ace(unsigned npb, unsigned idbegin, unsigned count, double dt,
                        const unsigned* idp, const float* velrho, float* motionvel,
                        float* motionace) {
  for (unsigned p = 0; p < npb; p++) {
    if (idp[p] - idbegin >= count) continue;
    const float* v = velrho + 4 * size_t(p);
    float* mv = motionvel + 3 * size_t(p);
    float* ma = motionace + 3 * size_t(p);
    for (int k = 0; k < 3; k++) {
      ma[k] = float((double(v[k]) - mv[k]) / dt);
      mv[k] = v[k];
    }
  }
}

}  // extern "C"

namespace {

// Determinant and inverse of a 4x4 matrix by cofactors (FP64).
inline double inverse4(const double a[4][4], double inv[4][4]) {
  double m[16], r[16];
  for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) m[4 * i + j] = a[i][j];
  r[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] + m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
  r[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] - m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
  r[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] + m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
  r[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] - m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
  r[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] - m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
  r[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] + m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
  r[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] - m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
  r[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] + m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
  r[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] + m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
  r[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] - m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
  r[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] + m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
  r[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] - m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
  r[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] - m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
  r[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] + m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
  r[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] - m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
  r[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] + m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
  const double det = m[0] * r[0] + m[1] * r[4] + m[2] * r[8] + m[3] * r[12];
  if (det != 0)
    for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) inv[i][j] = r[4 * i + j] / det;
  return det;
}

// Determinant and inverse of a 3x3 matrix (FP64): the 2-D systems over
// (1, x, z) of the reference's 2-D branches, with its term order (rule of
// Sarrus for the determinant, 2x2 cofactors over the determinant for the
// inverse; FunctionsMath.h:93-96 and 161-176, FunctionsMath_iker.h:111-114
// and 150-165).
inline double inverse3(const double a[3][3], double inv[3][3]) {
  const double det = a[0][0] * a[1][1] * a[2][2] + a[0][1] * a[1][2] * a[2][0] +
                     a[0][2] * a[1][0] * a[2][1] - a[2][0] * a[1][1] * a[0][2] -
                     a[2][1] * a[1][2] * a[0][0] - a[2][2] * a[1][0] * a[0][1];
  if (det != 0) {
    inv[0][0] = (a[1][1] * a[2][2] - a[1][2] * a[2][1]) / det;
    inv[0][1] = -(a[0][1] * a[2][2] - a[0][2] * a[2][1]) / det;
    inv[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) / det;
    inv[1][0] = -(a[1][0] * a[2][2] - a[1][2] * a[2][0]) / det;
    inv[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) / det;
    inv[1][2] = -(a[0][0] * a[1][2] - a[0][2] * a[1][0]) / det;
    inv[2][0] = (a[1][0] * a[2][1] - a[1][1] * a[2][0]) / det;
    inv[2][1] = -(a[0][0] * a[2][1] - a[0][1] * a[2][0]) / det;
    inv[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) / det;
  }
  return det;
}

// Solve of the mDBC moment system: in 3-D the 4x4 matrix over (1, x, y, z);
// in 2-D (all dy = 0, so the y row and column vanish) the 3x3 system over
// (1, x, z) the reference's 2-D branches build. a: the 4x4 sums; returns the
// determinant and writes the inverse into the 4x4 positions of the system's
// indices (the y row and column stay 0 in 2-D).
inline double mdbc_inverse(bool sim2d, const double a[4][4], double inv[4][4]) {
  if (!sim2d) return inverse4(a, inv);
  const int ix[3] = {0, 1, 3};
  double a3[3][3], i3[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) a3[i][j] = a[ix[i]][ix[j]];
  const double det = inverse3(a3, i3);
  for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) inv[i][j] = 0;
  if (det != 0)
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) inv[ix[i]][ix[j]] = i3[i][j];
  return det;
}

// mDBC density of boundary particles, SlipMode SLIP_Vel0: first-order
// (4x4) corrected Shepard density and gradient at the ghost node
// g = x_b + boundnor (normals already doubled), extrapolated back to the
// particle. CPU build: JSphCpu_mdbc.cpp:36-146 (FP64 positions, matrix
// accumulated in FP64); GPU build: KerInteractionMdbcCorrection_Fast,
// JSphGpu_mdbc_iker.cu:34-150 (distances from the ghost node's PosCell, FP32
// matrix). R = FP32 or FP64 (yardstick).
template <typename R, typename M>
void mdbc_correction(int sem, const orc_cte& c, const orc_div& d, unsigned n, const double* pos,
                     const float* ps, const typecode* code, const float* boundnor,
                     float* velrho) {
  const R ks2 = R(c.kernelsize2), h = R(c.kernelh), awen = R(c.awen), bwenh = R(c.bwenh);
  const R massf = R(c.massf);
#pragma omp parallel for schedule(guided)
  for (long pp = 0; pp < long(n); pp++) {
    const unsigned p1 = unsigned(pp);
    const float* bn = boundnor + 3 * size_t(p1);
    if (bn[0] == 0 && bn[1] == 0 && bn[2] == 0) continue;
    const double g[3] = {pos[3 * size_t(p1)] + bn[0], pos[3 * size_t(p1) + 1] + bn[1],
                         pos[3 * size_t(p1) + 2] + bn[2]};
    // PosCell of the ghost node (KerComputePosCell).
    float gps[4];
    {
      const double dx = g[0] - c.mapposmin[0], dy = g[1] - c.mapposmin[1], dz = g[2] - c.mapposmin[2];
      const double pcs = c.poscellsize;
      const unsigned cx = unsigned(dx / pcs), cy = unsigned(dy / pcs), cz = unsigned(dz / pcs);
      gps[0] = float(dx - pcs * cx); gps[1] = float(dy - pcs * cy); gps[2] = float(dz - pcs * cz);
      gps[3] = bitsf((cx << 19) | (cy << 9) | cz);
    }
    const int cx = int((g[0] - c.domposmin[0]) / c.scell) - d.czx;
    const int cy = int((g[1] - c.domposmin[1]) / c.scell) - d.czy;
    const int cz = int((g[2] - c.domposmin[2]) / c.scell) - d.czz;
    Row rows[25];
    const int nr = rows_at(d, cx, cy, cz, true, rows);
    R rhop1 = 0, gr[3] = {0, 0, 0};
    M a[4][4];
    for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) a[i][j] = 0;
    const unsigned ug = fbits(gps[3]);
    const int gx = int(ug >> 19), gy = int((ug & 0x0007fe00u) >> 9), gz = int(ug & 0x1ffu);
    for (int r = 0; r < nr; r++)
      for (unsigned p2 = rows[r].a; p2 < rows[r].b; p2++) {
        R dr[3];
        if (sem == SEM_GPU) {
          const float* b = ps + 4 * size_t(p2);
          const unsigned ub = fbits(b[3]);
          dr[0] = R(gps[0] - b[0] + c.poscellsize * float(gx - int(ub >> 19)));
          dr[1] = R(gps[1] - b[1] + c.poscellsize * float(gy - int((ub & 0x0007fe00u) >> 9)));
          dr[2] = R(gps[2] - b[2] + c.poscellsize * float(gz - int(ub & 0x1ffu)));
        } else if (sem == SEM_CPU) {
          for (int k = 0; k < 3; k++) dr[k] = R(float(g[k] - pos[3 * size_t(p2) + k]));
        } else {
          for (int k = 0; k < 3; k++) dr[k] = R(g[k] - pos[3 * size_t(p2) + k]);
        }
        const R rr2 = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
        if (!(rr2 <= ks2 && (code[p2] & kTypeMask) == kTypeFluid)) continue;
        const R wab = wend_wab<R>(rr2, h, awen);
        const R fac = wend_fac<R>(rr2, h, bwenh);
        const R fr[3] = {fac * dr[0], fac * dr[1], fac * dr[2]};
        const R volp2 = massf / R(velrho[4 * size_t(p2) + 3]);
        rhop1 += massf * wab;
        for (int k = 0; k < 3; k++) gr[k] += massf * fr[k];
        const R v[4] = {wab * volp2, fr[0] * volp2, fr[1] * volp2, fr[2] * volp2};
        for (int i = 0; i < 4; i++) {
          a[i][0] += v[i];
          for (int k = 0; k < 3; k++) a[i][k + 1] += dr[k] * v[i];
        }
      }
    float rhofinal = FLT_MAX;
    double ad[4][4], inv[4][4];
    for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) ad[i][j] = double(a[i][j]);
    const double det = mdbc_inverse(c.sim2d != 0, ad, inv);
    if (std::fabs(det) >= 1e-3) {
      // The GPU build rounds the inverse to FP32 (tmatrix4f); the CPU build keeps FP64.
      M im[4][4];
      for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) im[i][j] = M(inv[i][j]);
      const R b[4] = {rhop1, gr[0], gr[1], gr[2]};
      R o[4];
      for (int i = 0; i < 4; i++)
        o[i] = R(im[i][0] * b[0] + im[i][1] * b[1] + im[i][2] * b[2] + im[i][3] * b[3]);
      rhofinal = float(o[0] - o[1] * R(-bn[0]) - o[2] * R(-bn[1]) - o[3] * R(-bn[2]));
    } else if (a[0][0] > 0) {
      rhofinal = float(rhop1 / R(a[0][0]));
    }
    velrho[4 * size_t(p1) + 3] = (rhofinal != FLT_MAX ? rhofinal : c.rhopzero);
  }
}


// mDBC2 (SlipMode no-slip, MDBC2_Std, velocities kept: MDBC2_KEEPVEL,
// DualSphDef.h:471) of boundary particles [0,n). At the ghost node
// g = x_b + boundnor: Shepard sums, the 4x4 first-order correction matrix
// (columns 1..3 with -dr), a submerged test, the corrected density when the
// matrix is invertible and its scaled infinity-norm condition number
// dp^2 |A| |A^-1| <= 50 (else Shepard; Shepard clamped at rho0 when the
// kernel sum is < 0.1), pressure cloned to the particle with the hydrostatic
// gradient, and the tangential velocity of 2 v_motion - v_ghost (no-slip) or
// of v_ghost (free slip, JSphCpu_mdbc.cpp:412-420; JSphGpu_mdbc_iker.cu:464-473).
// CPU build: JSphCpu_mdbc.cpp:246-455 (dr from FP64 positions, matrix and
// velocity sums in FP64, FP64 Cs0 and Dp); GPU build:
// KerInteractionMdbc2Correction_Fast, JSphGpu_mdbc_iker.cu:289-487 (dr from
// PosCell, FP32 sums, FP64 determinant and inverse rounded to FP32, FP32
// Cs0). R = scalar sums, M = matrix and velocity sums. Writes velrho.w,
// bmode (1 = mDBC2, 2 = dry ghost: mass off) and tangenvel.
template <typename R, typename M>
void mdbc2_correction(int sem, const orc_cte& c, const orc_div& d, unsigned n, const double* pos,
                      const float* ps, const typecode* code, const float* boundnor,
                      const float* gravity, const float* motionvel, const float* motionace,
                      double cs0d, double dpd, float* velrho, unsigned char* bmode,
                      float* tangenvel, bool freeslip) {
  const R ks2 = R(c.kernelsize2), h = R(c.kernelh), awen = R(c.awen), bwenh = R(c.bwenh);
  const R massf = R(c.massf);
  const bool f64 = (sem == SEM_F64);
  // cs0^2 as each build forms it.
  const double cs2 = sem == SEM_GPU ? double(c.cs0 * c.cs0) : cs0d * cs0d;
  const float rho0 = c.rhopzero;
  auto tangential = [](const float* bn, const float v[3], float* out) {
    const float sn = std::sqrt(bn[0] * bn[0] + bn[1] * bn[1] + bn[2] * bn[2]);
    const float nx = bn[0] / sn, ny = bn[1] / sn, nz = bn[2] / sn;
    const float vdn = v[0] * nx + v[1] * ny + v[2] * nz;
    out[0] = v[0] - vdn * nx; out[1] = v[1] - vdn * ny; out[2] = v[2] - vdn * nz;
  };
#pragma omp parallel for schedule(guided)
  for (long pp = 0; pp < long(n); pp++) {
    const unsigned p1 = unsigned(pp);
    const float* bn = boundnor + 3 * size_t(p1);
    if (bn[0] == 0 && bn[1] == 0 && bn[2] == 0) continue;
    const float zero3[3] = {0, 0, 0};
    const float* mv = motionvel ? motionvel + 3 * size_t(p1) : zero3;
    const float* ma = motionace ? motionace + 3 * size_t(p1) : zero3;
    const double g[3] = {pos[3 * size_t(p1)] + bn[0], pos[3 * size_t(p1) + 1] + bn[1],
                         pos[3 * size_t(p1) + 2] + bn[2]};
    float gps[4];
    {
      const double dx = g[0] - c.mapposmin[0], dy = g[1] - c.mapposmin[1], dz = g[2] - c.mapposmin[2];
      const double pcs = c.poscellsize;
      const unsigned cx = unsigned(dx / pcs), cy = unsigned(dy / pcs), cz = unsigned(dz / pcs);
      gps[0] = float(dx - pcs * cx); gps[1] = float(dy - pcs * cy); gps[2] = float(dz - pcs * cz);
      gps[3] = bitsf((cx << 19) | (cy << 9) | cz);
    }
    const int cx = int((g[0] - c.domposmin[0]) / c.scell) - d.czx;
    const int cy = int((g[1] - c.domposmin[1]) / c.scell) - d.czy;
    const int cz = int((g[2] - c.domposmin[2]) / c.scell) - d.czz;
    Row rows[25];
    const int nr = rows_at(d, cx, cy, cz, true, rows);
    R rhop1 = 0, gr[3] = {0, 0, 0}, sumwab = 0, submerged = 0;
    M velp[3] = {0, 0, 0};
    M a[4][4];
    for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) a[i][j] = 0;
    const unsigned ug = fbits(gps[3]);
    const int gx = int(ug >> 19), gy = int((ug & 0x0007fe00u) >> 9), gz = int(ug & 0x1ffu);
    for (int r = 0; r < nr; r++)
      for (unsigned p2 = rows[r].a; p2 < rows[r].b; p2++) {
        R dr[3];
        if (sem == SEM_GPU) {
          const float* b = ps + 4 * size_t(p2);
          const unsigned ub = fbits(b[3]);
          dr[0] = R(gps[0] - b[0] + c.poscellsize * float(gx - int(ub >> 19)));
          dr[1] = R(gps[1] - b[1] + c.poscellsize * float(gy - int((ub & 0x0007fe00u) >> 9)));
          dr[2] = R(gps[2] - b[2] + c.poscellsize * float(gz - int(ub & 0x1ffu)));
        } else if (sem == SEM_CPU) {
          for (int k = 0; k < 3; k++) dr[k] = R(float(g[k] - pos[3 * size_t(p2) + k]));
        } else {
          for (int k = 0; k < 3; k++) dr[k] = R(g[k] - pos[3 * size_t(p2) + k]);
        }
        const R rr2 = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
        if (!(rr2 <= ks2 && (code[p2] & kTypeMask) == kTypeFluid)) continue;
        const R wab = wend_wab<R>(rr2, h, awen);
        const R fac = wend_fac<R>(rr2, h, bwenh);
        const R fr[3] = {fac * dr[0], fac * dr[1], fac * dr[2]};
        const float* v2 = velrho + 4 * size_t(p2);
        const R volp2 = massf / R(v2[3]);
        submerged -= volp2 * (dr[0] * fr[0] + dr[1] * fr[1] + dr[2] * fr[2]);
        rhop1 += massf * wab;
        for (int k = 0; k < 3; k++) gr[k] += massf * fr[k];
        const R vwab = wab * volp2;
        sumwab += vwab;
        for (int k = 0; k < 3; k++) velp[k] += M(vwab * R(v2[k]));
        const R v[4] = {vwab, fr[0] * volp2, fr[1] * volp2, fr[2] * volp2};
        for (int i = 0; i < 4; i++) {
          a[i][0] += M(v[i]);
          for (int k = 0; k < 3; k++) a[i][k + 1] -= M(dr[k] * v[i]);
        }
      }
    float* vr1 = velrho + 4 * size_t(p1);
    float* tv = tangenvel + 3 * size_t(p1);
    if (!(submerged > R(0))) {
      bmode[p1] = 2;
      vr1[3] = rho0;
      tangential(bn, mv, tv);
      continue;
    }
    bmode[p1] = 1;
    float rhofinal;
    if (sumwab < R(0.1f)) {
      rhofinal = std::max(rho0, float(M(rhop1) / a[0][0]));
    } else {
      double ad[4][4], inv[4][4];
      for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) ad[i][j] = double(a[i][j]);
      const double det = mdbc_inverse(c.sim2d != 0, ad, inv);
      if (std::fabs(det) >= 0.001) {
        M im[4][4];
        for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) im[i][j] = M(inv[i][j]);
        float na = 0, ni = 0;
        for (int i = 0; i < 4; i++) {
          const float ra = float(std::fabs(a[i][0]) + std::fabs(a[i][1]) + std::fabs(a[i][2]) + std::fabs(a[i][3]));
          const float ri = float(std::fabs(im[i][0]) + std::fabs(im[i][1]) + std::fabs(im[i][2]) + std::fabs(im[i][3]));
          na = std::max(na, ra);
          ni = std::max(ni, ri);
        }
        const float dp2 = sem == SEM_GPU ? c.dp * c.dp : float(dpd * dpd);
        const float condinf = dp2 * na * ni;
        if (condinf <= 50)
          rhofinal = float(im[0][0] * M(rhop1) + im[0][1] * M(gr[0]) + im[0][2] * M(gr[1]) +
                           im[0][3] * M(gr[2]));
        else
          rhofinal = float(M(rhop1) / a[0][0]);
      } else {
        rhofinal = float(M(rhop1) / a[0][0]);
      }
    }
    // Pressure cloning (Mdbc2PressClone, JSphCpu_mdbc.cpp:185-216 /
    // KerMdbc2PressClone, JSphGpu_mdbc_iker.cu:210-236).
    {
      const float sn = std::sqrt(bn[0] * bn[0] + bn[1] * bn[1] + bn[2] * bn[2]);
      const float nx = bn[0] / sn, ny = bn[1] / sn, nz = bn[2] / sn;
      const float normpos = (-bn[0]) * nx + (-bn[1]) * ny + (-bn[2]) * nz;
      const float fx = gravity[0] - ma[0], fy = gravity[1] - ma[1], fz = gravity[2] - ma[2];
      const float normforce = rho0 * (fx * nx + fy * ny + fz * nz);
      if (f64) {
        const double pghost = cs2 * (double(rhofinal) - double(rho0));
        const double pf = pghost + double(normforce) * double(normpos);
        rhofinal = float(double(rho0) + pf / cs2);
      } else {
        const float pghost = float(cs2 * double(rhofinal - rho0));
        const float pf = pghost + normforce * normpos;
        rhofinal = rho0 + float(double(pf) / cs2);
      }
    }
    vr1[3] = rhofinal;
    float vg[3], v2[3];
    for (int k = 0; k < 3; k++) vg[k] = float(velp[k] / a[0][0]);
    for (int k = 0; k < 3; k++) v2[k] = freeslip ? vg[k] : mv[k] + mv[k] - vg[k];
    tangential(bn, v2, tv);
  }
}

}  // namespace

extern "C" {

// SPS stress tau/rho^2 (JSphCpu.cpp:1279-1309).
void orc_sps_tau(unsigned np, unsigned npb, float smag, float blin, const float* velrho,
                 const float* s2, float* tau) {
  for (unsigned p = npb; p < np; p++) {
    const float* s = s2 + 6 * size_t(p);
    const float xx = s[0], xy = s[1], xz = s[2], yy = s[3], yz = s[4], zz = s[5];
    const float pow1 = xx * xx + yy * yy + zz * zz;
    const float prr = xy * xy + xz * xz + yz * yz + pow1 + pow1;
    const float visc_sps = smag * std::sqrt(prr);
    const float div_u = xx + yy + zz;
    const float sps_k = visc_sps * div_u;
    const float sps_blin = blin * prr;
    const float sumsps = -(sps_k + sps_blin);
    const float two = visc_sps + visc_sps;
    const float one_rho = 1.0f / velrho[4 * size_t(p) + 3];
    float* t = tau + 6 * size_t(p);
    t[0] = one_rho * (two * xx + sumsps);
    t[1] = one_rho * (visc_sps * xy);
    t[2] = one_rho * (visc_sps * xz);
    t[3] = one_rho * (two * yy + sumsps);
    t[4] = one_rho * (visc_sps * yz);
    t[5] = one_rho * (two * zz + sumsps);
  }
}

// Symplectic predictor (JSphCpu.cpp:1687-1844; KerComputeStepSymplecticPre,
// JSphGpuSimple_ker.cu:265-353) for DBC, floating particles included.
void orc_sym_pre(unsigned np, unsigned npb, const float* vrpre, const float* ar, const float* ace,
                 double dtm, float rho0, float rhomin, float rhomax, const float g[3],
                 typecode* code, double* mov, float* velrho) {
  for (unsigned p = 0; p < np; p++) {
    const float* pre = vrpre + 4 * size_t(p);
    float* nw = velrho + 4 * size_t(p);
    if (p < npb) {
      float w = float(double(pre[3]) + dtm * ar[p]);
      w = (w < rho0 ? rho0 : w);
      nw[0] = pre[0]; nw[1] = pre[1]; nw[2] = pre[2]; nw[3] = w;
      continue;
    }
    const float rhonew = float(double(pre[3]) + dtm * ar[p]);
    if ((code[p] & kTypeMask) == kTypeFloating) {
      // Floating particles keep their velocity and position (moved by
      // RunFloating); density clamped at rho0 (JSphCpu.cpp:1795-1799).
      nw[0] = pre[0]; nw[1] = pre[1]; nw[2] = pre[2];
      nw[3] = (rhonew < rho0 ? rho0 : rhonew);
      continue;
    }
    mov[3 * size_t(p)] = double(pre[0]) * dtm;
    mov[3 * size_t(p) + 1] = double(pre[1]) * dtm;
    mov[3 * size_t(p) + 2] = double(pre[2]) * dtm;
    for (int k = 0; k < 3; k++)
      nw[k] = float(double(pre[k]) + (double(ace[3 * size_t(p) + k]) + g[k]) * dtm);
    nw[3] = rhonew;
    if ((rhonew < rhomin || rhonew > rhomax) && (code[p] & kSpecial) == 0)
      code[p] = typecode((code[p] & ~kSpecial) | kOutRho);
  }
}

// Symplectic corrector (JSphCpu.cpp:1850-2029; KerComputeStepSymplecticCor).
void orc_sym_cor(unsigned np, unsigned npb, const float* vrpre, const float* ar, const float* ace,
                 double dtm, double dt, float rho0, float rhomin, float rhomax, const float g[3],
                 typecode* code, double* mov, float* velrho) {
  for (unsigned p = 0; p < np; p++) {
    const float* pre = vrpre + 4 * size_t(p);
    float* nw = velrho + 4 * size_t(p);
    const double eps = (-double(ar[p]) / double(nw[3])) * dt;
    const float rhonew = float(double(pre[3]) * (2. - eps) / (2. + eps));
    if (p < npb) {
      nw[0] = nw[1] = nw[2] = 0;
      nw[3] = (rhonew < rho0 ? rho0 : rhonew);
      continue;
    }
    if ((code[p] & kTypeMask) == kTypeFloating) {  // JSphCpu.cpp:1979-1983
      nw[0] = pre[0]; nw[1] = pre[1]; nw[2] = pre[2];
      nw[3] = (rhonew < rho0 ? rho0 : rhonew);
      continue;
    }
    float v[3];
    for (int k = 0; k < 3; k++)
      v[k] = float(double(pre[k]) + (double(ace[3 * size_t(p) + k]) + g[k]) * dt);
    for (int k = 0; k < 3; k++) mov[3 * size_t(p) + k] = (double(pre[k]) + double(v[k])) * dtm;
    nw[0] = v[0]; nw[1] = v[1]; nw[2] = v[2]; nw[3] = rhonew;
    if ((rhonew < rhomin || rhonew > rhomax) && (code[p] & kSpecial) == 0)
      code[p] = typecode((code[p] & ~kSpecial) | kOutRho);
  }
}

// Verlet (JSphCpu.cpp:1523-1680; KerComputeStepVerlet JSphGpuSimple_ker.cu:132-225).
void orc_verlet(unsigned np, unsigned npb, const float* vr1, const float* vr2, const float* ar,
                const float* ace, double dt, double dt2, float rho0, float rhomin, float rhomax,
                const float g[3], typecode* code, double* mov, float* vrnew) {
  const double dt205 = 0.5 * dt * dt;
  for (unsigned p = 0; p < np; p++) {
    float* nw = vrnew + 4 * size_t(p);
    if (p < npb) {
      float w = float(double(vr2[4 * size_t(p) + 3]) + dt2 * ar[p]);
      nw[0] = nw[1] = nw[2] = 0;
      nw[3] = (w < rho0 ? rho0 : w);
      continue;
    }
    const float rhonew = float(double(vr2[4 * size_t(p) + 3]) + dt2 * ar[p]);
    if ((code[p] & kTypeMask) == kTypeFloating) {
      // Floating particle (JSphCpu.cpp:1592-1596): current velocity, density
      // clamped at rho0, no displacement (RunFloating moves it).
      for (int k = 0; k < 3; k++) nw[k] = vr1[4 * size_t(p) + k];
      nw[3] = (rhonew < rho0 ? rho0 : rhonew);
      continue;
    }
    double acg[3];
    for (int k = 0; k < 3; k++) acg[k] = double(ace[3 * size_t(p) + k]) + g[k];
    float v[3];
    for (int k = 0; k < 3; k++) {
      mov[3 * size_t(p) + k] = double(vr1[4 * size_t(p) + k]) * dt + acg[k] * dt205;
      v[k] = float(double(vr2[4 * size_t(p) + k]) + acg[k] * dt2);
    }
    nw[0] = v[0]; nw[1] = v[1]; nw[2] = v[2]; nw[3] = rhonew;
    if ((rhonew < rhomin || rhonew > rhomax) && (code[p] & kSpecial) == 0)
      code[p] = typecode((code[p] & ~kSpecial) | kOutRho);
  }
}

// Position update of fluid [npb,np): x = xpre + mov in FP64, OUTMOV/OUTPOS
// checks and new cell (JSphCpu::UpdatePos, JSphCpu.cpp:1464-1516).
void orc_step_pos(unsigned np, unsigned npb, const orc_cte* c, const double* pospre,
                  const double* mov, double* pos, unsigned* dcell, typecode* code) {
  for (unsigned p = npb; p < np; p++) {
    const bool outrho = (code[p] & kSpecial) == kOutRho;
    // Periodic copies keep their position (KerComputeStepPos2 'normal').
    if (c->periactive && !outrho && (code[p] & kSpecial) != 0) continue;
    const double* m = mov + 3 * size_t(p);
    const bool outmov = (std::fabs(float(m[0])) > c->movlimit || std::fabs(float(m[1])) > c->movlimit ||
                         std::fabs(float(m[2])) > c->movlimit);
    double x[3];
    for (int k = 0; k < 3; k++) x[k] = pospre[3 * size_t(p) + k] + m[k];
    double dd[3];
    bool out = false;
    for (int k = 0; k < 3; k++) {
      dd[k] = x[k] - c->maprealposmin[k];
      if (dd[k] != dd[k] || dd[k] < 0 || dd[k] >= c->maprealsize[k]) out = true;
    }
    // UpdatePeriodicPos (JSph.cpp:2076-2101): re-enter through the opposite
    // face of a periodic axis; the cell is then taken from DomPosMin.
    if (c->periactive && out) {
      for (int a = 0; a < 3; a++) {
        if (!(c->periactive & (1u << a))) continue;
        if (dd[a] < 0) for (int k = 0; k < 3; k++) dd[k] -= c->perinc[a][k];
        if (dd[a] >= c->maprealsize[a]) for (int k = 0; k < 3; k++) dd[k] += c->perinc[a][k];
      }
      out = false;
      for (int a = 0; a < 3; a++)
        if (!(c->periactive & (1u << a)) && (dd[a] < 0 || dd[a] >= c->maprealsize[a])) out = true;
      for (int k = 0; k < 3; k++) x[k] = dd[k] + c->maprealposmin[k];
    }
    if (c->periactive)
      for (int k = 0; k < 3; k++) dd[k] = x[k] - c->domposmin[k];
    for (int k = 0; k < 3; k++) pos[3 * size_t(p) + k] = x[k];
    if (outrho || outmov || out) {
      unsigned rc = code[p] & ~kSpecial;
      rc |= out ? kOutPos : (outrho ? kOutRho : kOutMov);
      code[p] = typecode(rc);
      dcell[p] = 0xffffffffu;
    } else {
      dcell[p] = mkcell(c->cellcode, unsigned(dd[0] / c->scell), unsigned(dd[1] / c->scell),
                        unsigned(dd[2] / c->scell));
    }
  }
}

// DtVariable (JSphGpu.cpp:1103-1147; JSphCpu.cpp:2035-2078).
double orc_dt_variable(double velmax, double acemax, float viscdtmax, float kernelh, double cs0,
                       double cfl, double dtmin) {
  const double dt1 = (acemax ? std::sqrt(double(kernelh) / acemax) : DBL_MAX);
  const double dt2 = double(kernelh) / (std::max(cs0, velmax * 10.) + double(kernelh) * viscdtmax);
  double dt = cfl * std::min(dt1, dt2);
  if (dt < dtmin) dt = dtmin;
  return dt;
}

// Real and checked neighbour counts (JDsPips.cpp ComputeCpu; JDsPips_ker.cu:54-124):
// res = {real fluid, real bound, checked fluid, checked bound}.
void orc_pips(const orc_cte* c, const orc_div* d, unsigned np, unsigned npb, unsigned npbok,
              const unsigned* dcell, const float* poscell, unsigned long long res[4]) {
  unsigned long long rf = 0, rb = 0, cf = 0, cb = 0;
#pragma omp parallel for schedule(guided) reduction(+ : rf, rb, cf, cb)
  for (long pp = 0; pp < long(np); pp++) {
    const unsigned p1 = unsigned(pp);
    const bool fluid = p1 >= npb;
    if (!fluid && p1 >= npbok) continue;
    Row rows[50];
    const unsigned cel = dcell[p1];
    const int cx = int(cx_of(c->cellcode, cel)) - d->czx, cy = int(cy_of(c->cellcode, cel)) - d->czy,
              cz = int(cz_of(c->cellcode, cel)) - d->czz;
    int nr = rows_at(*d, cx, cy, cz, true, rows);
    if (fluid) nr += rows_at(*d, cx, cy, cz, false, rows + nr);
    unsigned long long nreal = 0, nchk = 0;
    for (int r = 0; r < nr; r++) {
      for (unsigned p2 = rows[r].a; p2 < rows[r].b; p2++) {
        float dx, dy, dz;
        rel<float>(SEM_GPU, *c, poscell, nullptr, p1, p2, dx, dy, dz);
        const float rr2 = dx * dx + dy * dy + dz * dz;
        if (rr2 <= c->kernelsize2 && rr2 >= 1e-18f) nreal++;
      }
      nchk += rows[r].b - rows[r].a;
    }
    if (fluid) { rf += nreal; cf += nchk; } else { rb += nreal; cb += nchk; }
  }
  res[0] = rf; res[1] = rb; res[2] = cf; res[3] = cb;
}

// Brute-force O(N^2) neighbour sums in FP64 from FP64 positions: the
// interaction without any cell list (checks the cell-list traversal).
void orc_bruteforce_fp64(const orc_cte* c, int tvisco, int tddt, float viscob, float viscof,
                         unsigned np, unsigned npb, const double* pos, const float* velrho,
                         double* ace, double* ar, double* viscdt) {
  const double h = c->kernelh, ks2 = c->kernelsize2, eta2 = c->eta2, cbar = c->cs0;
  std::vector<double> press(np);
  for (unsigned p = 0; p < np; p++)
    press[p] = double(c->cteb) * (std::pow(double(velrho[4 * size_t(p) + 3]) / double(c->rhopzero),
                                           double(c->gamma)) - 1.0);
#pragma omp parallel for schedule(dynamic, 16)
  for (long pp = 0; pp < long(np); pp++) {
    const unsigned p1 = unsigned(pp);
    const bool fl1 = p1 >= npb;
    double a[3] = {0, 0, 0}, arp = 0, visc = 0, delta = 0;
    bool delta_off = false;
    const float* v1 = velrho + 4 * size_t(p1);
    // Same neighbour order class as the reference: fluid first, then boundary.
    for (int pass = 0; pass < 2; pass++) {
      const bool boundp2 = pass == 1;
      if (!fl1 && boundp2) break;
      const unsigned q0 = boundp2 ? 0 : npb, q1 = boundp2 ? npb : np;
      for (unsigned p2 = q0; p2 < q1; p2++) {
        const double dx = pos[3 * size_t(p1)] - pos[3 * size_t(p2)];
        const double dy = pos[3 * size_t(p1) + 1] - pos[3 * size_t(p2) + 1];
        const double dz = pos[3 * size_t(p1) + 2] - pos[3 * size_t(p2) + 2];
        const double rr2 = dx * dx + dy * dy + dz * dz;
        if (!(rr2 <= ks2 && rr2 >= 1e-18)) continue;
        const double rad = std::sqrt(rr2), qq = rad / h, w1 = 1 - 0.5 * qq;
        const double fac = double(c->bwenh) * w1 * w1 * w1;
        const double fr[3] = {fac * dx, fac * dy, fac * dz};
        const float* v2 = velrho + 4 * size_t(p2);
        const double rho1 = v1[3], rho2 = v2[3];
        const double dv[3] = {double(v1[0]) - v2[0], double(v1[1]) - v2[1], double(v1[2]) - v2[2]};
        const double massp2 = boundp2 ? c->massb : c->massf;
        const double dvfr = dv[0] * fr[0] + dv[1] * fr[1] + dv[2] * fr[2];
        const double dot = dx * dv[0] + dy * dv[1] + dz * dv[2];
        visc = std::max(dot / (rr2 + eta2), visc);
        if (!fl1) {
          arp += double(c->massf) * dvfr * (rho1 / rho2);
          continue;
        }
        const double prs = (press[p1] + press[p2]) / (rho1 * rho2);
        for (int k = 0; k < 3; k++) a[k] += -prs * massp2 * fr[k];
        arp += massp2 * dvfr * (rho1 / rho2);
        const double dot3 = dx * fr[0] + dy * fr[1] + dz * fr[2];
        if ((tddt == 2 || (tddt == 3 && !boundp2)) && !delta_off) {
          const double rh = 1 + double(c->ddtgz) * dz;
          const double drho = double(c->rhopzero) * std::pow(rh, 1.0 / double(c->gamma)) - c->rhopzero;
          const double vd = double(c->ddtkh) * cbar * ((rho2 - rho1) - drho) / (rr2 + eta2);
          if (boundp2) delta_off = true;
          else delta -= vd * dot3 * massp2 / rho2;
        }
        const double visco = boundp2 ? viscob : viscof;
        if (tvisco == 1) {
          if (dot < 0) {
            const double pi = (-visco * cbar * (h * dot / (rr2 + eta2)) / ((rho1 + rho2) * 0.5)) * massp2;
            for (int k = 0; k < 3; k++) a[k] -= pi * fr[k];
          }
        } else {
          const double vt = massp2 * (4 * visco / ((rr2 + eta2) * (rho1 + rho2))) * dot3;
          for (int k = 0; k < 3; k++) a[k] += vt * dv[k];
        }
      }
    }
    for (int k = 0; k < 3; k++) ace[3 * size_t(p1) + k] = a[k];
    ar[p1] = arp + (tddt && !delta_off ? delta : 0);
    viscdt[p1] = visc;
  }
}

// Kernel and EOS of one value in precision prec (0 = FP32, 1 = FP64), the
// same helpers the interaction uses (FunSphKernel.h:198-244, FunSphEos.h:38-66).
void orc_kernel_eval(int prec, float rr2, float h, float awen, float bwenh, double* wab,
                     double* fac) {
  if (prec == 0) {
    *wab = wend_wab<float>(rr2, h, awen);
    *fac = wend_fac<float>(rr2, h, bwenh);
  } else {
    *wab = wend_wab<double>(rr2, h, awen);
    *fac = wend_fac<double>(rr2, h, bwenh);
  }
}
// mDBC correction of boundary particles [0,n) (see mdbc_correction).
// prec: 0 FP32, 1 FP64; sem: SEM_GPU (FP32 matrix) or SEM_CPU (FP64 matrix).
void orc_mdbc(int prec, int sem, const orc_cte* c, const orc_div* d, unsigned n,
              const double* pos, const float* poscell, const typecode* code,
              const float* boundnor2, float* velrho) {
  if (prec == 1) mdbc_correction<double, double>(SEM_F64, *c, *d, n, pos, poscell, code, boundnor2, velrho);
  else if (sem == SEM_CPU) mdbc_correction<float, double>(sem, *c, *d, n, pos, poscell, code, boundnor2, velrho);
  else mdbc_correction<float, float>(sem, *c, *d, n, pos, poscell, code, boundnor2, velrho);
}
// mDBC2 correction of boundary particles [0,n) (see mdbc2_correction).
// prec: 0 FP32, 1 FP64; sem: SEM_GPU or SEM_CPU. motionvel/motionace may be
// NULL (fixed boundaries). cs0 and dp are the FP64 case constants.
void orc_mdbc2_slip(int, int, int, const orc_cte*, const orc_div*, unsigned, const double*,
                    const float*, const typecode*, const float*, const float*, const float*,
                    const float*, double, double, float*, unsigned char*, float*);

void orc_mdbc2(int prec, int sem, const orc_cte* c, const orc_div* d, unsigned n,
               const double* pos, const float* poscell, const typecode* code,
               const float* boundnor, const float* gravity, const float* motionvel,
               const float* motionace, double cs0, double dp, float* velrho,
               unsigned char* bmode, float* tangenvel) {
  orc_mdbc2_slip(prec, sem, 2, c, d, n, pos, poscell, code, boundnor, gravity, motionvel,
                 motionace, cs0, dp, velrho, bmode, tangenvel);
}

// The same with the slip mode: 2 no-slip, 3 free slip.
void orc_mdbc2_slip(int prec, int sem, int slipmode, const orc_cte* c, const orc_div* d,
                    unsigned n, const double* pos, const float* poscell, const typecode* code,
                    const float* boundnor, const float* gravity, const float* motionvel,
                    const float* motionace, double cs0, double dp, float* velrho,
                    unsigned char* bmode, float* tangenvel) {
  const bool fs = slipmode == 3;
  if (prec == 1)
    mdbc2_correction<double, double>(SEM_F64, *c, *d, n, pos, poscell, code, boundnor, gravity,
                                     motionvel, motionace, cs0, dp, velrho, bmode, tangenvel, fs);
  else if (sem == SEM_CPU)
    mdbc2_correction<float, double>(sem, *c, *d, n, pos, poscell, code, boundnor, gravity,
                                    motionvel, motionace, cs0, dp, velrho, bmode, tangenvel, fs);
  else
    mdbc2_correction<float, float>(sem, *c, *d, n, pos, poscell, code, boundnor, gravity,
                                   motionvel, motionace, cs0, dp, velrho, bmode, tangenvel, fs);
}

// No-penetration step correction (JSphCpu.cpp:1555-1570, 1920-1935;
// JSphGpuSimple_ker.cu:177-191, 447-461) on the outputs of the Symplectic
// corrector or the Verlet step: a flagged fluid particle (w > 5) takes, on
// each axis with a non-zero correction, velocity base + correction and
// displacement velocity * dt. velbase: velrhopre (corrector) or the current
// velocity (Verlet); mov: double3 per particle.
void orc_nopen_step(unsigned np, unsigned npb, const float* nopenshift, const float* velbase,
                    double dt, float* velrho, double* mov) {
  for (unsigned p = npb; p < np; p++) {
    const float* ns = nopenshift + 4 * size_t(p);
    if (!(ns[3] > 5.f)) continue;
    for (int k = 0; k < 3; k++)
      if (ns[k] != 0) {
        velrho[4 * size_t(p) + k] = velbase[4 * size_t(p) + k] + ns[k];
        mov[3 * size_t(p) + k] = double(velrho[4 * size_t(p) + k]) * dt;
      }
  }
}
// Shifting displacement of fluid [pini, pini+n) (JSphShifting::RunCpu,
// JSphShifting.cpp:390-420; GPU KerRunShifting JSphShifting_ker.cu:123-151):
// umagn = dt*coef*h*|v| in FP64, free-surface factor (w - tfs)/(3 - tfs)
// (2 - tfs in 2-D, JSphShifting.cpp:394) above tfs (0 below; tfs = 0: off),
// shift = sums*umagn clamped to +-float(dp*0.1); cancelled particles
// (x = FLT_MAX) get 0.
void orc_run_shifting(unsigned n, unsigned pini, double dt, float shiftcoef, float kernelh,
                      float tfs, double dp, const float* velrho, float* shiftposfs,
                      unsigned sim2d) {
  const double coefumagn = dt * shiftcoef * kernelh;
  const double coeftfs = (sim2d ? 2.0 : 3.0) - tfs;
  const float maxdist = float(dp * 0.1);
  for (unsigned p = pini; p < pini + n; p++) {
    float* rs = shiftposfs + 4 * size_t(p);
    if (rs[0] != FLT_MAX) {
      const float* v = velrho + 4 * size_t(p);
      const double vx = v[0], vy = v[1], vz = v[2];
      double umagn = coefumagn * std::sqrt(vx * vx + vy * vy + vz * vz);
      if (tfs) {
        if (rs[3] < tfs) umagn = 0;
        else umagn *= (double(rs[3]) - tfs) / coeftfs;
      }
      for (int k = 0; k < 3; k++) {
        const float sd = float(double(rs[k]) * umagn);
        rs[k] = (std::fabs(sd) < maxdist ? sd : (sd >= 0 ? maxdist : -maxdist));
      }
    } else {
      rs[0] = rs[1] = rs[2] = rs[3] = 0;
    }
  }
}
// Adds the shifting displacement to the fluid displacement mov (FP64).
void orc_add_shift(unsigned np, unsigned npb, const float* shiftposfs, double* mov) {
  for (unsigned p = npb; p < np; p++)
    for (int k = 0; k < 3; k++) mov[3 * size_t(p) + k] += double(shiftposfs[4 * size_t(p) + k]);
}
// Integrator rule of mDBC2 boundaries (mode >= 1 keeps the state it starts
// from: KerComputeStepSymplecticPre/Cor JSphGpuSimple_ker.cu:276,406-412,
// KerComputeStepVerlet :142; JSphCpu::ComputeVelrhoBound JSphCpu.cpp:1607-1616):
// velrho[p] = src[p] for p < npb with bmode[p] >= 1, after the DBC update.
void orc_mdbc2_keep(unsigned npb, const unsigned char* bmode, const float* src, float* velrho) {
  for (unsigned p = 0; p < npb; p++)
    if (bmode[p] >= 1)
      for (int k = 0; k < 4; k++) velrho[4 * size_t(p) + k] = src[4 * size_t(p) + k];
}
double orc_eos(int prec, int sem, const orc_cte* c, float rho) {
  return prec == 0 ? double(eos<float>(sem, *c, rho)) : eos<double>(sem, *c, rho);
}

}  // extern "C"
