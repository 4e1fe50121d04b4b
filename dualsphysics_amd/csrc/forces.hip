// Force interaction of the SPH step on gfx950 (FP32 pair arithmetic).
//
// Physics follow the reference GPU kernels KerInteractionForcesFluid /
// KerInteractionForcesFluidBox / KerInteractionForcesBound (JSphGpu_ker.cu:
// 450-990): for every fluid particle the neighbour cell rows are visited
// z-outer, y-inner, fluid cells first and then boundary cells, the particles of
// a row in ascending sorted index, all contributions in one accumulator set;
// distances from the cell-relative FP32 PosCell.
//
// Two fluid kernels produce the same sums in the same order:
//  - k_forces_fluid_rows: one lane per particle looping over its rows (the
//    reference's structure). A wave executes the pair body whenever any lane
//    has a neighbour, so most lanes idle (about a quarter of them do work on a
//    jittered lattice).
//  - k_forces_fluid_tile (default): a wave walks the rows shared by its lanes
//    once, stages each row in LDS with coalesced loads, lets every lane run a
//    cheap conservative distance test against the staged candidates and queue
//    its hits in LDS, then evaluates the exact pair body for the queued pairs
//    in wave-uniform batches (every lane busy), so the expensive body runs
//    only for real pairs. Queue order is the scan order, so the summation
//    order is unchanged.
#include <cfloat>
#include <cstring>
#include <map>
#include <mutex>
#include <type_traits>
#include "ensemble.h"
#include "sph_math.h"

using namespace dsph;

namespace {

constexpr unsigned kBsForces = 128;
constexpr unsigned kBs = 256;

// Per-interaction preparation: pressure (EOS once per particle; each pair sees
// the same value the reference computes per pair), the reciprocal density
// (v_rcp_f32, as a fast-math quotient by rho computes it), and float positions
// in the map frame for the conservative neighbour test.
__device__ inline void prep_one(unsigned p, const Cte& c, const float4* __restrict__ poscell,
    const float4* __restrict__ velrho, float2* __restrict__ press, float4* __restrict__ posf) {
  const float rho = velrho[p].w;
  press[p] = make_float2(eos_press(rho, c), fast_rcp(rho));
  if (posf) {
    const float4 ps = poscell[p];
    const double pcs = c.pcs;
    posf[p] = make_float4(float(double(pscel_fx(ps.w)) * pcs + double(ps.x)),
                          float(double(pscel_fy(ps.w)) * pcs + double(ps.y)),
                          float(double(pscel_fz(ps.w)) * pcs + double(ps.z)), 0.f);
  }
}

__global__ void __launch_bounds__(kBs) k_prep(unsigned n, const float4* __restrict__ poscell,
    const float4* __restrict__ velrho, float2* __restrict__ press, float4* __restrict__ posf,
    Cte c) {
  const unsigned p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  prep_one(p, c, poscell, velrho, press, posf);
}

// XCD-aware block order. The dispatcher sends block i to XCD i % 8; each
// XCD has its own L2. Renumbering so that XCD x runs one contiguous range of
// logical blocks keeps neighbouring particles (which read the same neighbour
// data) on the same L2.
constexpr unsigned kNumXcd = 8;
__device__ inline unsigned xcd_block(unsigned i, unsigned nblocks) {
  const unsigned per = nblocks / kNumXcd, rem = nblocks % kNumXcd;
  const unsigned x = i % kNumXcd, k = i / kNumXcd;
  return x * per + min(x, rem) + k;
}

struct Acc {
  float ax, ay, az, ar, visc, delta;
  float s[6];   // two_strain {xx,xy,xz,yy,yz,zz}
  float g[9];   // full velocity gradient (learned closure features)
  float4 sh;    // shifting: sum m/rho gradW (x = FLT_MAX: cancelled), w = -sum m/rho r.gradW
  float npc[3], nps[3];  // no-penetration: boundary neighbours counted and corrections, per axis
};

__device__ inline void acc_init(Acc& a) {
  a.ax = a.ay = a.az = a.ar = a.visc = a.delta = 0;
  a.sh = make_float4(0, 0, 0, 0);
#pragma unroll
  for (int k = 0; k < 3; k++) a.npc[k] = a.nps[k] = 0;
#pragma unroll
  for (int k = 0; k < 6; k++) a.s[k] = 0;
#pragma unroll
  for (int k = 0; k < 9; k++) a.g[k] = 0;
}

struct P1 {
  float4 ps, vr;
  float press, rrho;
  float tau[6];
  bool ftp1;   // floating particle (FT kernels only)
};

// Exact pair body for fluid p1 and neighbour p2 (KerInteractionForcesFluidBox,
// JSphGpu_ker.cu:539-777, in-scope options). boundp2 selects the boundary
// variant (mass, viscosity, no SPS tau of p2, delta-SPH rules).
// pair_eval takes the neighbour's PosCell, velrho and {pressure, 1/rho} already loaded
// and applies the exact neighbour test rr2 <= (2h)^2, rr2 >= ALMOSTZERO.
// MDBC2: boundary neighbours in mDBC2 mode (boundmode, tangential velocity):
// no mass when their ghost node is dry, viscous terms against the tangential
// velocity (JSphGpu_ker.cu:585-588, 725-730; velocities kept for continuity,
// MDBC2_KEEPVEL).
// SHIFT: standard shifting (JSphGpu_ker.cu:677-684): accumulates m2/rho2 gradW
// and -m2/rho2 r.gradW; a boundary neighbour cancels it (x = FLT_MAX) for
// shiftmode 1 (no boundaries) and, for shiftmode 2, when it is fixed.
// FT: floating bodies. A floating neighbour (code type floating, in the
// fluid range) has its body's particle mass ftomassp[type value], stops the
// Molteni-Colagrossi term when it is not heavier than 1.2 fluid particles,
// takes no part in the Fourtakas term and gives no SPS tau; a floating p1
// has no density diffusion and no strain (JSphGpu_ker.cu:570-583, 668,
// 755-766, 838-843).
// NOPEN: mDBC2 no-penetration sums against boundary neighbours, with the
// velocity relative to their motion velocity (motionvel, float3 per
// particle; NULL = fixed boundaries) (MDBC2_NoPen, JSphGpu_ker.cu:686-707
// with ComputeNoPenVel, :428-440).
template <int TVISCO, int TDDT, bool VELGRAD, bool MDBC2 = false, bool SHIFT = false,
          bool FT = false, bool NOPEN = false>
__device__ __forceinline__ void pair_eval(bool boundp2, unsigned p2, const float4& ps2,
    const float4& vr2, float2 pr2, const P1& q, const float* __restrict__ tau, float viscob,
    float viscof, const Cte& c, Acc& a, const uint8_t* __restrict__ bmode = nullptr,
    const float* __restrict__ tangenvel = nullptr, int shiftmode = 0,
    const typecode* __restrict__ code = nullptr, const float* __restrict__ ftomassp = nullptr,
    const float* __restrict__ boundnor = nullptr, const float* __restrict__ motionvel = nullptr) {
  const float3 dr = pscel_dr(q.ps, ps2, c.pcs);
  const float rr2 = dr.x * dr.x + dr.y * dr.y + dr.z * dr.z;
  if (!(rr2 <= c.ks2 && rr2 >= kAlmostZero)) return;
  float massp2 = boundp2 ? c.massb : c.massf;
  if (MDBC2 && boundp2 && bmode[p2] == DSPH_BMODE_MDBC2OFF) massp2 = 0.f;
  bool ftp2 = false;
  if (FT && !boundp2) {
    const unsigned cd = code[p2];
    ftp2 = code_type(cd) == kCodeTypeFloating;
    if (ftp2) {
      const float ftm = ftomassp[cd & kCodeMaskValue];
      if (TDDT == DSPH_DDT_DDT && ftm <= massp2 * 1.2f) a.delta = FLT_MAX;
      massp2 = ftm;
    }
  }
  const float visco = boundp2 ? viscob : viscof;
  const float fac = wendland_fac(rr2, c);
  const float frx = fac * dr.x, fry = fac * dr.y, frz = fac * dr.z;
  const float4& vr1 = q.vr;
  {
    const float prs = (q.press + pr2.x) * (q.rrho * pr2.y);
    const float p_vpm = -prs * massp2;
    a.ax += p_vpm * frx; a.ay += p_vpm * fry; a.az += p_vpm * frz;
  }
  float dvx = vr1.x - vr2.x, dvy = vr1.y - vr2.y, dvz = vr1.z - vr2.z;
  const float rrho2 = pr2.y;
  a.ar += massp2 * (dvx * frx + dvy * fry + dvz * frz) * (vr1.w * rrho2);
  const float cbar = c.cs0;
  const float dot3 = dr.x * frx + dr.y * fry + dr.z * frz;
  const float rr2eta = fast_rcp(rr2 + c.eta2);
  if (TDDT == DSPH_DDT_DDT && a.delta != FLT_MAX) {
    const float rhop1over2 = vr1.w * rrho2;
    const float visc_densi = c.ddtkh * cbar * (rhop1over2 - 1.f) * rr2eta;
    const float delta = visc_densi * dot3 * massp2;
    // A boundary neighbour cancels the term with DBC only (JSphGpu_ker.cu:665).
    a.delta = ((boundp2 && c.tboundary == DSPH_BC_DBC) ? FLT_MAX : a.delta + delta);
  }
  if ((TDDT == DSPH_DDT_DDT2 && a.delta != FLT_MAX && !(FT && ftp2)) ||
      (TDDT == DSPH_DDT_DDT2FULL && !boundp2 && !(FT && (ftp2 || a.delta == FLT_MAX)))) {
    const float rh = 1.f + c.ddtgz * dr.z;
    const float drho = c.rhopzero * pow_invgamma(rh, c) - c.rhopzero;
    const float visc_densi = c.ddtkh * cbar * ((vr2.w - vr1.w) - drho) * rr2eta;
    const float delta = visc_densi * dot3 * massp2 * rrho2;
    a.delta = (boundp2 ? FLT_MAX : a.delta - delta);
  }
  if (SHIFT && a.sh.x != FLT_MAX) {
    const float massrho = massp2 * rrho2;
    const bool noshift = boundp2 && (shiftmode == 1 ||
                                     (shiftmode == 2 && code_type(code[p2]) == kCodeTypeFixed));
    a.sh.x = noshift ? FLT_MAX : a.sh.x + massrho * frx;
    a.sh.y += massrho * fry;
    a.sh.z += massrho * frz;
    a.sh.w -= massrho * dot3;
  }
  if (NOPEN && boundp2) {
    // A boundary neighbour of the first layer (|normal| < 1.75 dp) within
    // 1.25 dp, with the particle less than 0.75 |normal| away along the
    // normal: per axis, a particle moving towards the wall (v relative to
    // the wall's motion velocity) gets -(3 - 4 max(|dr/n|, 1/4)) v n^2.
    if (sqrtf(rr2) < 1.25f * c.dp) {
      const float* bn = boundnor + 3 * size_t(p2);
      const float norm = sqrtf(bn[0] * bn[0] + bn[1] * bn[1] + bn[2] * bn[2]);
      const float nrm[3] = {fast_div(bn[0], norm), fast_div(bn[1], norm), fast_div(bn[2], norm)};
      const float d[3] = {dr.x, dr.y, dr.z};
      const float* mv = motionvel ? motionvel + 3 * size_t(p2) : nullptr;
      const float v[3] = {mv ? q.vr.x - mv[0] : q.vr.x, mv ? q.vr.y - mv[1] : q.vr.y,
                          mv ? q.vr.z - mv[2] : q.vr.z};
      if (nrm[0] * dr.x + nrm[1] * dr.y + nrm[2] * dr.z < 0.75f * norm && norm < 1.75f * c.dp) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
          if (d[k] * nrm[k] < 0.75f && fabsf(nrm[k]) > 0.001f * c.dp && v[k] * nrm[k] < 0.f) {
            const float ratio = fmaxf(fabsf(fast_div(d[k], nrm[k])), 0.25f);
            const float factor = -4.f * ratio + 3.f;
            a.npc[k] += 1.f;
            a.nps[k] -= factor * v[k] * nrm[k] * nrm[k];
          }
        }
      }
    }
  }
  if (MDBC2 && boundp2) {
    const float* tv = tangenvel + 3 * size_t(p2);
    dvx = vr1.x - tv[0]; dvy = vr1.y - tv[1]; dvz = vr1.z - tv[2];
  }
  const float dot = dr.x * dvx + dr.y * dvy + dr.z * dvz;
  const float dot_rr2 = dot * rr2eta;
  a.visc = fmaxf(dot_rr2, a.visc);
  if (TVISCO == DSPH_VISCO_ARTIFICIAL) {
    if (dot < 0) {
      const float amubar = c.h * dot_rr2;
      const float robar = (vr1.w + vr2.w) * 0.5f;
      const float pi_visc = fast_div(-visco * cbar * amubar, robar) * massp2;
      a.ax -= pi_visc * frx; a.ay -= pi_visc * fry; a.az -= pi_visc * frz;
    }
  } else {
    const float robar2 = (vr1.w + vr2.w);
    const float temp = fast_div(4.f * visco, (rr2 + c.eta2) * robar2);
    const float vtemp = massp2 * temp * dot3;
    a.ax += vtemp * dvx; a.ay += vtemp * dvy; a.az += vtemp * dvz;
  }
  if (TVISCO == DSPH_VISCO_LAMINARSPS) {
    float sxx = q.tau[0], sxy = q.tau[1], sxz = q.tau[2], syy = q.tau[3], syz = q.tau[4],
          szz = q.tau[5];
    if (!boundp2 && !(FT && ftp2)) {
      const float* t2 = tau + 6 * size_t(p2);
      sxx += t2[0]; sxy += t2[1]; sxz += t2[2]; syy += t2[3]; syz += t2[4]; szz += t2[5];
    }
    a.ax += massp2 * (sxx * frx + sxy * fry + sxz * frz);
    a.ay += massp2 * (sxy * frx + syy * fry + syz * frz);
    a.az += massp2 * (sxz * frx + syz * fry + szz * frz);
    if (!(FT && q.ftp1)) {
      const float volp2 = -massp2 * rrho2;
      float dv = dvx * volp2; a.s[0] += dv * frx; a.s[1] += dv * fry; a.s[2] += dv * frz;
      dv = dvy * volp2;       a.s[1] += dv * frx; a.s[3] += dv * fry; a.s[4] += dv * frz;
      dv = dvz * volp2;       a.s[2] += dv * frx; a.s[4] += dv * fry; a.s[5] += dv * frz;
    }
  }
  if (VELGRAD) {
    // Full gradient G_ij = sum_b -m_b/rho_b (v_a-v_b)_i dW/dx_j.
    const float volp2 = -massp2 * rrho2;
    const float ux = dvx * volp2, uy = dvy * volp2, uz = dvz * volp2;
    a.g[0] += ux * frx; a.g[1] += ux * fry; a.g[2] += ux * frz;
    a.g[3] += uy * frx; a.g[4] += uy * fry; a.g[5] += uy * frz;
    a.g[6] += uz * frx; a.g[7] += uz * fry; a.g[8] += uz * frz;
  }
}

// Reference-structured pair: loads PosCell first and the rest only for a
// real neighbour.
template <int TVISCO, int TDDT, bool VELGRAD, bool MDBC2 = false, bool SHIFT = false,
          bool FT = false, bool NOPEN = false>
__device__ __forceinline__ void pair_body(bool boundp2, unsigned p2, const P1& q,
    const float4* __restrict__ poscell, const float4* __restrict__ velrho,
    const float2* __restrict__ press, const float* __restrict__ tau, float viscob, float viscof,
    const Cte& c, Acc& a, const uint8_t* __restrict__ bmode = nullptr,
    const float* __restrict__ tangenvel = nullptr, int shiftmode = 0,
    const typecode* __restrict__ code = nullptr, const float* __restrict__ ftomassp = nullptr,
    const float* __restrict__ boundnor = nullptr, const float* __restrict__ motionvel = nullptr) {
  const float4 ps2 = poscell[p2];
  const float3 dr = pscel_dr(q.ps, ps2, c.pcs);
  const float rr2 = dr.x * dr.x + dr.y * dr.y + dr.z * dr.z;
  if (!(rr2 <= c.ks2 && rr2 >= kAlmostZero)) return;
  pair_eval<TVISCO, TDDT, VELGRAD, MDBC2, SHIFT, FT, NOPEN>(
      boundp2, p2, ps2, velrho[p2], press[p2], q, tau, viscob, viscof, c, a, bmode, tangenvel,
      shiftmode, code, ftomassp, boundnor, motionvel);
}

template <int TVISCO>
__device__ inline void load_p1(unsigned p1, const float4* __restrict__ poscell,
                               const float4* __restrict__ velrho, const float2* __restrict__ press,
                               const float* __restrict__ tau, P1& q) {
  q.ps = poscell[p1];
  q.vr = velrho[p1];
  const float2 pr1 = press[p1];
  q.press = pr1.x;
  q.rrho = pr1.y;
  q.ftp1 = false;
#pragma unroll
  for (int k = 0; k < 6; k++)
    q.tau[k] = (TVISCO == DSPH_VISCO_LAMINARSPS) ? tau[6 * size_t(p1) + k] : 0.f;
}

template <int TVISCO, int TDDT, bool VELGRAD>
__device__ inline void store_results(unsigned p1, const Acc& a, float* __restrict__ viscdt,
                                     float* __restrict__ ar, float* __restrict__ ace,
                                     float* __restrict__ strain, float* __restrict__ velgrad) {
  if (a.ar != 0 || a.ax != 0 || a.ay != 0 || a.az != 0 || a.visc != 0) {
    float arp1 = a.ar;
    if (TDDT != DSPH_DDT_NONE && a.delta != FLT_MAX) arp1 += a.delta;
    ar[p1] += arp1;
    float* ap = ace + 3 * size_t(p1);
    ap[0] += a.ax; ap[1] += a.ay; ap[2] += a.az;
    if (a.visc > viscdt[p1]) viscdt[p1] = a.visc;
    if (TVISCO == DSPH_VISCO_LAMINARSPS) {
      float* sp = strain + 6 * size_t(p1);
#pragma unroll
      for (int k = 0; k < 6; k++) sp[k] += a.s[k];
    }
  }
  if (VELGRAD) {
    float* gp = velgrad + 9 * size_t(p1);
#pragma unroll
    for (int k = 0; k < 9; k++) gp[k] = a.g[k];
  }
}

__device__ inline NgLimits p1_limits(unsigned p1, const unsigned* __restrict__ dcell,
                                     const DivData& dvd) {
  const unsigned cel = dcell[p1];
  const int cx = int(dcel_cellx(dvd.domcellcode, cel)) - dvd.czx;
  const int cy = int(dcel_celly(dvd.domcellcode, cel)) - dvd.czy;
  const int cz = int(dcel_cellz(dvd.domcellcode, cel)) - dvd.czz;
  return ng_limits(cx, cy, cz, dvd);
}

// ------------------------------------------------------------ reference-order rows kernel
template <int TVISCO, int TDDT, bool VELGRAD, bool MDBC2 = false, bool SHIFT = false,
          bool FT = false, bool NOPEN = false>
__device__ __forceinline__ void fluid_rows_body(unsigned p1, float viscob, float viscof,
    const DivData& dvd, const Cte& c, const unsigned* __restrict__ dcell,
    const float4* __restrict__ poscell, const float4* __restrict__ velrho,
    const float2* __restrict__ press, const float* __restrict__ tau, float* __restrict__ viscdt,
    float* __restrict__ ar, float* __restrict__ ace, float* __restrict__ strain,
    float* __restrict__ velgrad, const uint8_t* __restrict__ bmode = nullptr,
    const float* __restrict__ tangenvel = nullptr, int shiftmode = 0,
    const typecode* __restrict__ code = nullptr, float4* __restrict__ shiftposfs = nullptr,
    const float* __restrict__ ftomassp = nullptr, const float* __restrict__ boundnor = nullptr,
    float4* __restrict__ nopenshift = nullptr, const float* __restrict__ motionvel = nullptr) {
  Acc a;
  acc_init(a);
  if (SHIFT) a.sh = shiftposfs[p1];
  P1 q;
  load_p1<TVISCO>(p1, poscell, velrho, press, tau, q);
  if (FT) {
    q.ftp1 = code_type(code[p1]) == kCodeTypeFloating;
    if (q.ftp1 && TDDT != DSPH_DDT_NONE) a.delta = FLT_MAX;
  }
  const NgLimits l = p1_limits(p1, dcell, dvd);
  for (int kind = 0; kind < 2; kind++) {
    const int base = kind == 0 ? int(dvd.cellfluid) : 0;
    for (int z = l.z0; z < l.z1; z++)
      for (int y = l.y0; y < l.y1; y++) {
        unsigned pini, pfin;
        row_range(dvd.begincell, base + z * dvd.ncw + y * dvd.ncx, l.x0, l.x1, pini, pfin);
        for (unsigned p2 = pini; p2 < pfin; p2++)
          pair_body<TVISCO, TDDT, VELGRAD, MDBC2, SHIFT, FT, NOPEN>(
              kind == 1, p2, q, poscell, velrho, press, tau, viscob, viscof, c, a, bmode,
              tangenvel, shiftmode, code, ftomassp, boundnor, motionvel);
      }
  }
  store_results<TVISCO, TDDT, VELGRAD>(p1, a, viscdt, ar, ace, strain, velgrad);
  if (SHIFT) shiftposfs[p1] = a.sh;
  if (NOPEN) {
    // Mean correction per axis over the boundary neighbours that asked for
    // one; w = 10 marks a particle to correct (JSphGpu_ker.cu:958-976).
    float4 ns = make_float4(0, 0, 0, 0);
    if (a.npc[0] > 0.f) { ns.x = fast_div(a.nps[0], a.npc[0]); ns.w = 10.f; }
    if (a.npc[1] > 0.f) { ns.y = fast_div(a.nps[1], a.npc[1]); ns.w = 10.f; }
    if (a.npc[2] > 0.f) { ns.z = fast_div(a.nps[2], a.npc[2]); ns.w = 10.f; }
    nopenshift[p1] = ns;
  }
}

template <int TVISCO, int TDDT, bool VELGRAD, bool MDBC2, bool SHIFT = false, bool FT = false,
          bool NOPEN = false>
__global__ void __launch_bounds__(kBsForces) k_forces_fluid_rows(
    unsigned n, unsigned pinit, float viscob, float viscof, DivData dvd, Cte c,
    const unsigned* __restrict__ dcell, const float4* __restrict__ poscell,
    const float4* __restrict__ velrho, const float2* __restrict__ press,
    const float* __restrict__ tau, float* __restrict__ viscdt, float* __restrict__ ar,
    float* __restrict__ ace, float* __restrict__ strain, float* __restrict__ velgrad,
    const uint8_t* __restrict__ bmode, const float* __restrict__ tangenvel, int shiftmode = 0,
    const typecode* __restrict__ code = nullptr, float4* __restrict__ shiftposfs = nullptr,
    const float* __restrict__ ftomassp = nullptr, const float* __restrict__ boundnor = nullptr,
    float4* __restrict__ nopenshift = nullptr, const float* __restrict__ motionvel = nullptr) {
  const unsigned p = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (p >= n) return;
  fluid_rows_body<TVISCO, TDDT, VELGRAD, MDBC2, SHIFT, FT, NOPEN>(
      p + pinit, viscob, viscof, dvd, c, dcell, poscell, velrho, press, tau, viscdt, ar, ace,
      strain, velgrad, bmode, tangenvel, shiftmode, code, shiftposfs, ftomassp, boundnor,
      nopenshift, motionvel);
}

// Boundary particle against one fluid neighbour that passed the neighbour
// test (KerInteractionForcesBound, JSphGpu_ker.cu:450-493): density rate and
// the viscous dt term. Shared by the one-lane-per-particle and the tile
// boundary kernels (same expressions, same bits).
__device__ __forceinline__ void bound_pair(const float4& vr1, const float3& dr, float rr2,
                                           const float4& vr2, float massp2, const Cte& c,
                                           float& arp1, float& visc) {
  const float fac = wendland_fac(rr2, c);
  const float frx = fac * dr.x, fry = fac * dr.y, frz = fac * dr.z;
  const float dvx = vr1.x - vr2.x, dvy = vr1.y - vr2.y, dvz = vr1.z - vr2.z;
  arp1 += massp2 * (dvx * frx + dvy * fry + dvz * frz) * fast_div(vr1.w, vr2.w);
  const float dot = dr.x * dvx + dr.y * dvy + dr.z * dvz;
  const float dot_rr2 = fast_div(dot, rr2 + c.eta2);
  visc = fmaxf(dot_rr2, visc);
}

// ------------------------------------------------------------ cell-row tile kernel
constexpr unsigned kTile = 64;   // candidates staged in LDS per tile (one bit each)
constexpr unsigned kMinBatch = 16;
constexpr unsigned kScanBlock = 8;  // staged candidates tested per loop trip
#ifndef DSPH_RING_TILES
#define DSPH_RING_TILES 12
#endif
// Tiles of hit masks kept per wave (8 bytes per lane and tile).
constexpr unsigned kRingTiles = DSPH_RING_TILES;
#ifndef DSPH_TILE_WAVES
#define DSPH_TILE_WAVES 1
#endif
// Waves per workgroup of the single-case tile kernel: the waves of one
// workgroup run on one CU and take consecutive groups of 64 particles, so
// they share the L1 copy of their common neighbour rows.
constexpr unsigned kTileWaves = DSPH_TILE_WAVES;

// Orders the LDS accesses of one wave whose lanes exchange data through LDS.
// The LDS executes the instructions of a wave in issue order, so fences at
// wavefront scope (which keep the compiler from moving accesses across them)
// are all that is needed; there is no workgroup barrier, as the waves of a
// workgroup run loops of different lengths.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline unsigned wave_min_u32(unsigned v) {
  for (int off = 32; off > 0; off >>= 1) v = min(v, (unsigned)__shfl_xor((int)v, off));
  return v;
}
__device__ inline unsigned wave_max_u32(unsigned v) {
  for (int off = 32; off > 0; off >>= 1) v = max(v, (unsigned)__shfl_xor((int)v, off));
  return v;
}
__device__ inline int wave_max_i32(int v) {
  for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off));
  return v;
}
__device__ inline float wave_max_f32(float v) {
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off));
  return v;
}
__device__ inline int uniform_i(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline float uniform_f(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
// Shifts the result of d <= thr into bit 0 of bits (bits = 2*bits + (thr >= d)):
// a compare into VCC and an add with carry, two instructions per candidate.
__device__ __forceinline__ void shift_in_le(unsigned& bits, float d, float thr) {
  asm("v_cmp_ge_f32_e32 vcc, %1, %2\n\tv_addc_co_u32_e32 %0, vcc, %0, %0, vcc"
      : "+v"(bits)
      : "v"(thr), "v"(d)
      : "vcc");
}

// Conservative distance test of one lane against the first nblk blocks of 8
// staged candidates (wave-uniform nblk <= 8): bit k of the result is set when
// candidate k passes d = |q2|^2 - 2 q1.q2 <= thr. Bits of candidates beyond
// the staged ones are undefined (the caller masks them with its window).
__device__ __forceinline__ unsigned long long scan_tile(const float4* tile, unsigned nblk,
                                                        float m2x, float m2y, float m2z,
                                                        float thr) {
  unsigned w[2];
#pragma unroll
  for (unsigned h = 0; h < 2; h++) {
    const unsigned nb = min(nblk - min(nblk, 4u * h), 4u);
    unsigned b = 0;
#pragma unroll 1
    for (unsigned i = 0; i < nb; i++) {
#pragma unroll
      for (unsigned k = 0; k < kScanBlock; k++) {
        const float4 f = tile[32 * h + kScanBlock * i + k];
        float d = fmaf(m2x, f.x, f.w);
        d = fmaf(m2y, f.y, d);
        d = fmaf(m2z, f.z, d);
        shift_in_le(b, d, thr);
      }
    }
    // Candidate 32h + j is at bit 8nb-1-j: left-align it to bit 31-j, reverse.
    w[h] = nb ? __builtin_bitreverse32(b << (32u - kScanBlock * nb)) : 0u;
  }
  return ((unsigned long long)w[1] << 32) | w[0];
}

// Bits [k0, k1) of a 64-bit mask, 0 <= k0 <= k1 <= 64.
__device__ inline unsigned long long bit_range(unsigned k0, unsigned k1) {
  const unsigned long long hi = k1 >= 64 ? ~0ull : ((1ull << k1) - 1ull);
  const unsigned long long lo = k0 >= 64 ? ~0ull : ((1ull << k0) - 1ull);
  return hi & ~lo;
}

// One wave = 64 consecutive fluid particles (sorted order). Lanes in the same
// cell row (cy, cz) form a group (normally the whole wave); the group walks
// the union of its lanes' neighbour rows once, fluid cells then boundary
// cells, z-outer / y-inner (JSphGpu_ker.cu:867-896). Each row is a contiguous
// range of sorted particles, staged in LDS 64 at a time with coalesced loads
// (positions relative to the group origin and their squared norm). Every lane
// tests the staged candidates (LDS broadcast) with a conservative distance
// bound, keeps the hits inside its own index window of the row (cells
// cx-s..cx+s) as a 64-bit mask per tile, and the masks of the last kRing
// tiles stay in LDS. The exact pair body then runs for the queued pairs in
// wave-uniform batches (every lane busy), each lane popping its hits in tile
// and bit order, i.e. in the reference's summation order: the results equal
// k_forces_fluid_rows bit for bit. Loads of the next pair are issued before
// the arithmetic of the current one.
// LDS of one wave of the tile kernel.
template <unsigned kRing>
struct TileLds {
  float4 tile[kTile];
  unsigned long long masks[kRing][64];
  unsigned tbase[kRing];
};

// BP1: the boundary kernel's form (KerInteractionForcesBound): p1 is a
// boundary particle, only fluid cells are scanned, bound_pair is the pair
// body and ar/viscdt the outputs (no neighbour pressure is needed).
template <int TVISCO, int TDDT, bool VELGRAD, unsigned kRing, bool MDBC2 = false,
          bool SHIFT = false, bool EOS7 = false, bool BP1 = false>
__device__ __forceinline__ void fluid_tile_body(unsigned p, bool valid, unsigned pinit,
    float viscob, float viscof, const DivData& dvd, const Cte& c, float ks2scan,
    const unsigned* __restrict__ dcell, const float4* __restrict__ poscell,
    const float4* __restrict__ posf, const float4* __restrict__ velrho,
    const float2* __restrict__ press, const float* __restrict__ tau, float* __restrict__ viscdt,
    float* __restrict__ ar, float* __restrict__ ace, float* __restrict__ strain,
    float* __restrict__ velgrad, TileLds<kRing>& lds, const uint8_t* __restrict__ bmode = nullptr,
    const float* __restrict__ tangenvel = nullptr, int shiftmode = 0,
    const typecode* __restrict__ code = nullptr, float4* __restrict__ shiftposfs = nullptr) {
  float4* tile = lds.tile;
  auto& masks = lds.masks;
  unsigned* tbase = lds.tbase;
  const unsigned lane = threadIdx.x & 63u;
  const unsigned p1 = (valid ? p : 0) + pinit;
  Acc a;
  acc_init(a);
  P1 q;
  load_p1<TVISCO>(p1, poscell, velrho, press, tau, q);
  if (SHIFT) a.sh = shiftposfs[p1];
  const float4 pf1 = posf[p1];
  int cx, cy, cz;
  {
    const unsigned cel = dcell[p1];
    cx = int(dcel_cellx(dvd.domcellcode, cel)) - dvd.czx;
    cy = int(dcel_celly(dvd.domcellcode, cel)) - dvd.czy;
    cz = int(dcel_cellz(dvd.domcellcode, cel)) - dvd.czz;
  }
  const int s = dvd.scelldiv;
  const int lx0 = cx - min(cx, s), lx1 = cx + min(dvd.ncx - cx - 1, s) + 1;

  // Queue state. wt: tiles scanned (uniform). Per lane: rt = tile being
  // consumed, cm = its remaining hits, tb = its base index | kind << 31,
  // len = queued hits. Invariant: len > 0 => cm != 0; len == 0 => rt == wt.
  unsigned wt = 0, rt = 0, tb = 0, len = 0;
  unsigned long long cm = 0;
  auto pop = [&](unsigned& j) {
    const unsigned bit = unsigned(__builtin_ctzll(cm));  // cm != 0 while len > 0
    j = tb + bit;  // kind bit 31 carried along
    cm &= cm - 1ull;
    len--;
    if (!cm) {
      if (len) {
        do {
          rt++;
          cm = masks[rt % kRing][lane];
        } while (!cm);
        tb = tbase[rt % kRing];
      } else {
        rt = wt;
      }
    }
  };
  // k wave-uniform iterations; each lane pops up to k of its queued pairs.
  // Pops the next queued pair and issues its loads (32-bit offsets from the
  // array bases).
  // EOS7 (gamma = 7): the neighbour's {press, 1/rho} are evaluated from its
  // density with the functions k_prep applies (so the same bits; the power is
  // four FP64 products) instead of being gathered: two gathers per pair
  // instead of three. Otherwise they are loaded.
  auto issue = [&](unsigned& j, float4& ps2, float4& vr2, float2& pr2) {
    pop(j);
    const unsigned jj = j & 0x7fffffffu;
    ps2 = *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(poscell) + (jj << 4));
    vr2 = *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(velrho) + (jj << 4));
    if (!EOS7 && !BP1)
      pr2 = *reinterpret_cast<const float2*>(reinterpret_cast<const char*>(press) + (jj << 3));
  };
  auto eval = [&](unsigned j, const float4& ps2, const float4& vr2, const float2& pr2) {
    if constexpr (BP1) {
      const float3 dr = pscel_dr(q.ps, ps2, c.pcs);
      const float rr2 = dr.x * dr.x + dr.y * dr.y + dr.z * dr.z;
      if (rr2 <= c.ks2 && rr2 >= kAlmostZero) bound_pair(q.vr, dr, rr2, vr2, c.massf, c, a.ar, a.visc);
      return;
    }
    const float2 pr =
        EOS7 ? make_float2(c.cteb * (pow7_f64(vr2.w * c.ovrhopzero) - 1.0f), fast_rcp(vr2.w))
             : pr2;
    pair_eval<TVISCO, TDDT, VELGRAD, MDBC2, SHIFT>((j >> 31) != 0, j & 0x7fffffffu, ps2, vr2,
                                                   pr, q, tau, viscob, viscof, c, a, bmode,
                                                   tangenvel, shiftmode, code);
  };
  // k wave-uniform iterations; each lane pops up to k of its queued pairs.
  // Two register sets alternate so the loads of the next pair are in flight
  // during the arithmetic of the current one. With kFull every lane has at
  // least k queued pairs, so no lane test is needed.
  auto process = [&](unsigned k, auto full) {
    constexpr bool kFull = decltype(full)::value;
    unsigned ja = 0, jb = 0;
    float4 psa = make_float4(0, 0, 0, 0), vra = psa, psb = psa, vrb = psa;
    float2 pra = make_float2(0, 0), prb = pra;
    bool ha = k != 0 && (kFull || len != 0);
    if (ha) issue(ja, psa, vra, pra);
    for (unsigned i = 0; i < k; i += 2) {
      const bool hb = i + 1 < k && (kFull || len != 0);
      if (hb) issue(jb, psb, vrb, prb);
      if (ha) eval(ja, psa, vra, pra);
      ha = i + 2 < k && (kFull || len != 0);
      if (ha) issue(ja, psa, vra, pra);
      if (hb) eval(jb, psb, vrb, prb);
    }
  };
  using Full = std::true_type;
  using Some = std::false_type;

  bool pending = valid;
  for (;;) {
    const unsigned long long pend = __ballot(pending);
    if (!pend) break;
    const int leader = __ffsll((long long)pend) - 1;
    const int gy = uniform_i(__shfl(cy, leader)), gz = uniform_i(__shfl(cz, leader));
    const bool member = pending && cy == gy && cz == gz;
    pending = pending && !member;
    const int gxmin = uniform_i(__shfl(cx, leader));
    const int gxmax = uniform_i(wave_max_i32(member ? cx : -1));
    const int X0 = gxmin - min(gxmin, s), X1 = gxmax + min(dvd.ncx - gxmax - 1, s) + 1;
    const int Y0 = gy - min(gy, s), Y1 = gy + min(dvd.ncy - gy - 1, s) + 1;
    const int Z0 = gz - min(gz, s), Z1 = gz + min(dvd.ncz - gz - 1, s) + 1;
    // Group origin and the per-lane threshold of the expanded distance test
    // |q1|^2 + |q2|^2 - 2 q1.q2 <= ks2scan (q relative to the origin). For a
    // real neighbour |q|,|q2| <= Q = max|q1| + 2h, and the rounding of the
    // expanded form stays below 8 * 2^-24 * (2Q)^2.
    const float ox = uniform_f(__shfl(pf1.x, leader)), oy = uniform_f(__shfl(pf1.y, leader)),
                oz = uniform_f(__shfl(pf1.z, leader));
    const float q1x = pf1.x - ox, q1y = pf1.y - oy, q1z = pf1.z - oz;
    const float s1 = q1x * q1x + q1y * q1y + q1z * q1z;
    const float qm = sqrtf(uniform_f(wave_max_f32(member ? s1 : 0.f))) + 1.01f * sqrtf(ks2scan);
    const float thr = ks2scan - s1 + 2.0e-6f * qm * qm;
    const float m2x = -2.f * q1x, m2y = -2.f * q1y, m2z = -2.f * q1z;
    for (int kind = 0; kind < (BP1 ? 1 : 2); kind++) {
      const int base = kind == 0 ? int(dvd.cellfluid) : 0;
      const unsigned kbit = unsigned(kind) << 31;
      for (int z = Z0; z < Z1; z++)
        for (int y = Y0; y < Y1; y++) {
          const int rowbase = base + z * dvd.ncw + y * dvd.ncx;
          unsigned gini, gfin;
          row_range(dvd.begincell, rowbase, X0, X1, gini, gfin);
          gini = unsigned(uniform_i(int(gini)));
          gfin = unsigned(uniform_i(int(gfin)));
          if (gfin <= gini) continue;
          unsigned lo = 0, hi = 0;
          if (member) row_range(dvd.begincell, rowbase, lx0, lx1, lo, hi);
          for (unsigned t = gini; t < gfin; t += kTile) {
            // The ring slot of tile wt must have been consumed by every lane.
            while (wt >= kRing && wave_min_u32(rt) <= wt - kRing) process(kMinBatch, Some());
            wave_sync();
            if (t + lane < gfin) {
              const float4 f = posf[t + lane];
              const float x = f.x - ox, y2 = f.y - oy, z2 = f.z - oz;
              tile[lane] = make_float4(x, y2, z2, x * x + y2 * y2 + z2 * z2);
            }
            if (lane == 0) tbase[wt % kRing] = t | kbit;
            wave_sync();
            // Hits of this lane among the staged candidates (blocks of 8 up
            // to the last staged one; the window drops the rest).
            unsigned long long hm = 0;
            if (member) {
              const unsigned k0 = lo > t ? min(lo - t, kTile) : 0u;
              const unsigned k1 = hi > t ? min(hi - t, kTile) : 0u;
              const unsigned long long win = k1 > k0 ? bit_range(k0, k1) : 0ull;
              if (win) {
                const unsigned nblk = (min(kTile, gfin - t) + kScanBlock - 1) / kScanBlock;
                hm = scan_tile(tile, nblk, m2x, m2y, m2z, thr) & win;
              }
            }
            masks[wt % kRing][lane] = hm;
            if (len == 0) {
              // Caught up: consume the new tile directly.
              cm = hm;
              tb = t | kbit;
              rt = hm ? wt : wt + 1;
            }
            len += unsigned(__popcll(hm));
            wt++;
            const unsigned mn = wave_min_u32(member ? len : 0xffffffffu);
            if (mn >= kMinBatch) {
              // All 64 lanes have mn pairs queued (the usual case: one group,
              // all lanes valid): batches without per-lane tests.
              if (wave_min_u32(len) >= mn) process(mn, Full());
              else process(mn, Some());
            }
          }
        }
    }
  }
  process(wave_max_u32(len), Some());
  if (BP1) {
    if (valid && (a.ar != 0 || a.visc != 0)) {
      ar[p1] += a.ar;
      if (a.visc > viscdt[p1]) viscdt[p1] = a.visc;
    }
  } else if (valid) {
    store_results<TVISCO, TDDT, VELGRAD>(p1, a, viscdt, ar, ace, strain, velgrad);
    if (SHIFT) shiftposfs[p1] = a.sh;
  }
}

// Waves per SIMD the register allocation of the tile kernel must allow. The
// kernel's speed is set by the latency of its neighbour gathers, so by the
// waves that cover it: 5 (96 registers, a few bytes spilled outside the pair
// loop) for the artificial and laminar accumulator sets, measured 10% faster
// than the 4 waves of the unconstrained allocation; 4 (128 registers) for
// the larger sets, which would otherwise get 3.
template <int TVISCO, bool VELGRAD, bool MDBC2, bool SHIFT>
constexpr int tile_min_waves() {
  return (TVISCO != DSPH_VISCO_LAMINARSPS && !VELGRAD && !MDBC2 && !SHIFT) ? 5 : 4;
}

template <int TVISCO, int TDDT, bool VELGRAD, unsigned kRing, bool MDBC2 = false,
          bool SHIFT = false, bool EOS7 = false>
__global__ void __launch_bounds__(64 * kTileWaves)
__attribute__((amdgpu_waves_per_eu(tile_min_waves<TVISCO, VELGRAD, MDBC2, SHIFT>(), 8)))
k_forces_fluid_tile(
    unsigned n, unsigned pinit, float viscob, float viscof, DivData dvd, Cte c, float ks2scan,
    const unsigned* __restrict__ dcell, const float4* __restrict__ poscell,
    const float4* __restrict__ posf, const float4* __restrict__ velrho,
    const float2* __restrict__ press, const float* __restrict__ tau, float* __restrict__ viscdt,
    float* __restrict__ ar, float* __restrict__ ace, float* __restrict__ strain,
    float* __restrict__ velgrad, const uint8_t* __restrict__ bmode = nullptr,
    const float* __restrict__ tangenvel = nullptr, int shiftmode = 0,
    const typecode* __restrict__ code = nullptr, float4* __restrict__ shiftposfs = nullptr) {
  __shared__ TileLds<kRing> lds[kTileWaves];
  const unsigned w = threadIdx.x >> 6;
  const unsigned p =
      (xcd_block(blockIdx.x, gridDim.x) * kTileWaves + w) * 64 + (threadIdx.x & 63u);
  fluid_tile_body<TVISCO, TDDT, VELGRAD, kRing, MDBC2, SHIFT, EOS7>(
      p, p < n, pinit, viscob, viscof, dvd, c, ks2scan, dcell, poscell, posf, velrho, press, tau,
      viscdt, ar, ace, strain, velgrad, lds[w], bmode, tangenvel, shiftmode, code, shiftposfs);
}

// Boundary particle p1 against fluid neighbours (KerInteractionForcesBound,
// JSphGpu_ker.cu:450-533).
// FT: a floating neighbour has its body's particle mass (JSphGpu_ker.cu:470-483).
template <bool FT = false>
__device__ __forceinline__ void bound_body(unsigned p1, const DivData& dvd, const Cte& c,
    const unsigned* __restrict__ dcell, const float4* __restrict__ poscell,
    const float4* __restrict__ velrho, float* __restrict__ viscdt, float* __restrict__ ar,
    const typecode* __restrict__ code = nullptr, const float* __restrict__ ftomassp = nullptr) {
  float visc = 0, arp1 = 0;
  const float4 ps1 = poscell[p1];
  const float4 vr1 = velrho[p1];
  const NgLimits l = p1_limits(p1, dcell, dvd);
  for (int z = l.z0; z < l.z1; z++)
    for (int y = l.y0; y < l.y1; y++) {
      unsigned pini, pfin;
      row_range(dvd.begincell, int(dvd.cellfluid) + z * dvd.ncw + y * dvd.ncx, l.x0, l.x1,
                pini, pfin);
      for (unsigned p2 = pini; p2 < pfin; p2++) {
        const float4 ps2 = poscell[p2];
        const float3 dr = pscel_dr(ps1, ps2, c.pcs);
        const float rr2 = dr.x * dr.x + dr.y * dr.y + dr.z * dr.z;
        if (rr2 <= c.ks2 && rr2 >= kAlmostZero) {
          const float4 vr2 = velrho[p2];
          float massp2 = c.massf;
          if (FT) {
            const unsigned cd = code[p2];
            if (code_type(cd) == kCodeTypeFloating) massp2 = ftomassp[cd & kCodeMaskValue];
          }
          bound_pair(vr1, dr, rr2, vr2, massp2, c, arp1, visc);
        }
      }
    }
  if (arp1 != 0 || visc != 0) {
    ar[p1] += arp1;
    if (visc > viscdt[p1]) viscdt[p1] = visc;
  }
}

template <bool FT = false>
__global__ void __launch_bounds__(kBsForces) k_forces_bound(unsigned n, DivData dvd, Cte c,
    const unsigned* __restrict__ dcell, const float4* __restrict__ poscell,
    const float4* __restrict__ velrho, float* __restrict__ viscdt, float* __restrict__ ar,
    const typecode* __restrict__ code = nullptr, const float* __restrict__ ftomassp = nullptr) {
  const unsigned p1 = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (p1 >= n) return;
  bound_body<FT>(p1, dvd, c, dcell, poscell, velrho, viscdt, ar, code, ftomassp);
}

// Boundary interaction in the tile form (same pairs, same order, same bits as
// k_forces_bound): a wave of 64 consecutive boundary particles stages the
// fluid rows around them in LDS and each lane tests the staged candidates.
__global__ void __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(5, 8)))
k_forces_bound_tile(unsigned n, DivData dvd, Cte c, float ks2scan,
    const unsigned* __restrict__ dcell, const float4* __restrict__ poscell,
    const float4* __restrict__ posf, const float4* __restrict__ velrho,
    const float2* __restrict__ press, float* __restrict__ viscdt, float* __restrict__ ar) {
  __shared__ TileLds<kRingTiles> lds;
  const unsigned p = xcd_block(blockIdx.x, gridDim.x) * 64 + threadIdx.x;
  fluid_tile_body<DSPH_VISCO_ARTIFICIAL, DSPH_DDT_NONE, false, kRingTiles, false, false, true,
                  true>(p, p < n, 0, 0.f, 0.f, dvd, c, ks2scan, dcell, poscell, posf, velrho,
                        press, nullptr, viscdt, ar, nullptr, nullptr, nullptr, lds);
}

// Radius^2 of the conservative scan test: (2h + slack)^2 where the slack
// bounds the rounding of map-frame float positions (each coordinate within
// half an ulp of the largest map extent, PosCell within half an ulp of pcs).
float scan_radius2(const dsph_ctx* ctx, const Cte&) { return dsph::ens_scan_radius2(ctx->cte); }

template <int TVISCO, int TDDT, bool VELGRAD>
void launch_fluid(dsph_ctx* ctx, const dsph_inter_parms* t, const DivData& dvd, const Cte& c,
                  const float2* press, const float4* posf) {
  KTimer kt(ctx, "k_forces_fluid");
  const unsigned npf = t->np - t->npb;
  if (t->d_nopenshift) {
    // mDBC2 with no penetration: the one-lane-per-particle kernel with the
    // no-penetration sums (shifting and floating bodies are refused with it
    // by dsph_interaction_forces).
    k_forces_fluid_rows<TVISCO, TDDT, VELGRAD, true, false, false, true>
        <<<grid_for(npf, kBsForces), kBsForces, 0, ctx->stream>>>(
            npf, t->npb, t->viscob, t->viscof, dvd, c, t->d_dcell, (const float4*)t->d_poscell,
            (const float4*)t->d_velrho, press, t->d_spstaurho2, t->d_viscdt, t->d_ar, t->d_ace,
            t->d_sps2strain, t->d_velgrad, t->d_boundmode, t->d_tangenvel, 0, nullptr, nullptr,
            nullptr, t->d_boundnor, (float4*)t->d_nopenshift, t->d_motionvel);
  } else if (t->d_ftomassp) {
    // Floating bodies: the one-lane-per-particle kernel with the floating
    // rules (shifting, mDBC2 and the velocity gradient are refused with them
    // by dsph_interaction_forces).
    if constexpr (!VELGRAD)
      k_forces_fluid_rows<TVISCO, TDDT, false, false, false, true>
          <<<grid_for(npf, kBsForces), kBsForces, 0, ctx->stream>>>(
              npf, t->npb, t->viscob, t->viscof, dvd, c, t->d_dcell, (const float4*)t->d_poscell,
              (const float4*)t->d_velrho, press, t->d_spstaurho2, t->d_viscdt, t->d_ar, t->d_ace,
              t->d_sps2strain, nullptr, nullptr, nullptr, 0, t->d_code, nullptr, t->d_ftomassp);
  } else if (t->d_shiftposfs && t->d_boundmode) {
    // Shifting with mDBC2 boundaries: the one-lane-per-particle kernel with
    // both rules (a boundary switched off has mass 0 in the shifting sums
    // too, JSphGpu_ker.cu:586-588, 733-740).
    k_forces_fluid_rows<TVISCO, TDDT, false, true, true><<<grid_for(npf, kBsForces), kBsForces,
                                                          0, ctx->stream>>>(
        npf, t->npb, t->viscob, t->viscof, dvd, c, t->d_dcell, (const float4*)t->d_poscell,
        (const float4*)t->d_velrho, press, t->d_spstaurho2, t->d_viscdt, t->d_ar, t->d_ace,
        t->d_sps2strain, nullptr, t->d_boundmode, t->d_tangenvel, t->shiftmode, t->d_code,
        (float4*)t->d_shiftposfs);
  } else if (t->d_shiftposfs && ctx->forces_kernel == 0) {
    // Shifting in the tile kernel (no penetration and the velocity gradient
    // are refused with shifting by dsph_interaction_forces).
    k_forces_fluid_tile<TVISCO, TDDT, false, kRingTiles, false, true>
        <<<grid_for(npf, 64 * kTileWaves), 64 * kTileWaves, 0, ctx->stream>>>(
        npf, t->npb, t->viscob, t->viscof, dvd, c, scan_radius2(ctx, c), t->d_dcell,
        (const float4*)t->d_poscell, posf, (const float4*)t->d_velrho, press, t->d_spstaurho2,
        t->d_viscdt, t->d_ar, t->d_ace, t->d_sps2strain, nullptr, nullptr, nullptr, t->shiftmode,
        t->d_code, (float4*)t->d_shiftposfs);
  } else if (t->d_shiftposfs) {
    k_forces_fluid_rows<TVISCO, TDDT, false, false, true><<<grid_for(npf, kBsForces), kBsForces,
                                                           0, ctx->stream>>>(
        npf, t->npb, t->viscob, t->viscof, dvd, c, t->d_dcell, (const float4*)t->d_poscell,
        (const float4*)t->d_velrho, press, t->d_spstaurho2, t->d_viscdt, t->d_ar, t->d_ace,
        t->d_sps2strain, nullptr, nullptr, nullptr, t->shiftmode, t->d_code,
        (float4*)t->d_shiftposfs);
  } else if (t->d_boundmode && ctx->forces_kernel == 0 && !VELGRAD) {
    // mDBC2 boundaries in the tile kernel.
    k_forces_fluid_tile<TVISCO, TDDT, false, kRingTiles, true>
        <<<grid_for(npf, 64 * kTileWaves), 64 * kTileWaves, 0, ctx->stream>>>(
        npf, t->npb, t->viscob, t->viscof, dvd, c, scan_radius2(ctx, c), t->d_dcell,
        (const float4*)t->d_poscell, posf, (const float4*)t->d_velrho, press, t->d_spstaurho2,
        t->d_viscdt, t->d_ar, t->d_ace, t->d_sps2strain, nullptr, t->d_boundmode, t->d_tangenvel);
  } else if (t->d_boundmode) {
    // mDBC2 with the velocity gradient (learned closure) or the rows kernel
    // requested: the one-lane-per-particle kernel.
    k_forces_fluid_rows<TVISCO, TDDT, VELGRAD, true><<<grid_for(npf, kBsForces), kBsForces, 0,
                                                       ctx->stream>>>(
        npf, t->npb, t->viscob, t->viscof, dvd, c, t->d_dcell, (const float4*)t->d_poscell,
        (const float4*)t->d_velrho, press, t->d_spstaurho2, t->d_viscdt, t->d_ar, t->d_ace,
        t->d_sps2strain, t->d_velgrad, t->d_boundmode, t->d_tangenvel);
  } else if (ctx->forces_kernel == 0) {
    // gamma = 7 without the gradient output: the neighbour pressure is
    // evaluated in the kernel instead of being gathered (Laminar+SPS too: 127
    // registers, no spill, 9 % faster on the wave flume).
    constexpr bool kEos7Built = !VELGRAD;
    if (kEos7Built && c.gamma == 7.f)
      k_forces_fluid_tile<TVISCO, TDDT, false, kRingTiles, false, false, kEos7Built>
          <<<grid_for(npf, 64 * kTileWaves), 64 * kTileWaves, 0, ctx->stream>>>(
              npf, t->npb, t->viscob, t->viscof, dvd, c, scan_radius2(ctx, c), t->d_dcell,
              (const float4*)t->d_poscell, posf, (const float4*)t->d_velrho, press,
              t->d_spstaurho2, t->d_viscdt, t->d_ar, t->d_ace, t->d_sps2strain, nullptr);
    else
      k_forces_fluid_tile<TVISCO, TDDT, VELGRAD, kRingTiles>
          <<<grid_for(npf, 64 * kTileWaves), 64 * kTileWaves, 0, ctx->stream>>>(
              npf, t->npb, t->viscob, t->viscof, dvd, c, scan_radius2(ctx, c), t->d_dcell,
              (const float4*)t->d_poscell, posf, (const float4*)t->d_velrho, press,
              t->d_spstaurho2, t->d_viscdt, t->d_ar, t->d_ace, t->d_sps2strain, t->d_velgrad);
  } else {
    k_forces_fluid_rows<TVISCO, TDDT, VELGRAD, false><<<grid_for(npf, kBsForces), kBsForces, 0,
                                                        ctx->stream>>>(
        npf, t->npb, t->viscob, t->viscof, dvd, c, t->d_dcell, (const float4*)t->d_poscell,
        (const float4*)t->d_velrho, press, t->d_spstaurho2, t->d_viscdt, t->d_ar, t->d_ace,
        t->d_sps2strain, t->d_velgrad, nullptr, nullptr);
  }
}

template <int TVISCO, bool VELGRAD>
int dispatch_ddt(dsph_ctx* ctx, const dsph_inter_parms* t, const DivData& dvd, const Cte& c,
                 const float2* press, const float4* posf) {
  switch (t->tdensity) {
    case DSPH_DDT_NONE: launch_fluid<TVISCO, DSPH_DDT_NONE, VELGRAD>(ctx, t, dvd, c, press, posf); break;
    case DSPH_DDT_DDT: launch_fluid<TVISCO, DSPH_DDT_DDT, VELGRAD>(ctx, t, dvd, c, press, posf); break;
    case DSPH_DDT_DDT2: launch_fluid<TVISCO, DSPH_DDT_DDT2, VELGRAD>(ctx, t, dvd, c, press, posf); break;
    case DSPH_DDT_DDT2FULL:
      launch_fluid<TVISCO, DSPH_DDT_DDT2FULL, VELGRAD>(ctx, t, dvd, c, press, posf);
      break;
    default: return set_error(ctx, DSPH_E_ARG, "dsph_interaction_forces: invalid tdensity");
  }
  return DSPH_OK;
}

// ---------------------------------------------------------------- ensembles
// The same bodies over the cases of an ensemble; the block map gives the case
// and the first case-local index of each workgroup.
__global__ void k_fill_u32(unsigned* __restrict__ a, unsigned n, unsigned v) {
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

__global__ void __launch_bounds__(kBs) k_ens_prep(const EnsCase* __restrict__ cases,
    const EnsState* __restrict__ st, const EnsBlock* __restrict__ pmap,
    const float4* __restrict__ poscell, const float4* __restrict__ velrho,
    float2* __restrict__ press, float4* __restrict__ posf) {
  const EnsBlock e = pmap[blockIdx.x];
  if (!st[e.x].active) return;
  const unsigned l = e.y + threadIdx.x;
  if (l >= st[e.x].np) return;
  const EnsCase& ec = cases[e.x];
  prep_one(ec.p0 + l, ec.c, poscell, velrho, press, posf);
}

template <int TVISCO, int TDDT, unsigned kRing, bool EOS7>
__global__ void __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(tile_min_waves<TVISCO, false, false, false>(), 8)))
k_ens_fluid_tile(const EnsCase* __restrict__ cases,
    const EnsState* __restrict__ st, const EnsBlock* __restrict__ wmap,
    const unsigned* __restrict__ dcell, const float4* __restrict__ poscell,
    const float4* __restrict__ posf, const float4* __restrict__ velrho,
    const float2* __restrict__ press, float* __restrict__ viscdt, float* __restrict__ ar,
    float* __restrict__ ace) {
  __shared__ TileLds<kRing> lds;
  const EnsBlock e = wmap[xcd_block(blockIdx.x, gridDim.x)];
  if (!st[e.x].active) return;
  const EnsCase& ec = cases[e.x];
  const unsigned npf = st[e.x].np - ec.npb;
  if (e.y >= npf) return;
  const unsigned p = e.y + threadIdx.x;
  fluid_tile_body<TVISCO, TDDT, false, kRing, false, false, EOS7>(
      p, p < npf, ec.p0 + ec.npb, ec.viscob, ec.viscof,
                                              ec.dvd, ec.c, ec.ks2scan, dcell, poscell, posf,
                                              velrho, press, nullptr, viscdt, ar, ace, nullptr,
                                              nullptr, lds);
}

// Persistent form of k_ens_fluid_tile: a grid sized to the resident waves
// pulls work items (waves of 64 fluid particles, entries of wmap) in chunks of
// `chunk` consecutive items from counters, until the list is exhausted. The
// list is split into kNumXcd contiguous segments, one per XCD (block i runs
// on XCD i % 8), so neighbouring particles stay on one L2; a wave drains its
// own XCD's segment first, then helps the others. Every wave leaves once each
// segment's counter has passed its end (counters only grow), so the grid
// drains. ctr[0..kNumXcd) must be zero at launch. Items are independent:
// the schedule changes no result.
template <int TVISCO, int TDDT, unsigned kRing, bool EOS7>
__global__ void __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(tile_min_waves<TVISCO, false, false, false>(), 8)))
k_ens_fluid_tile_persist(const EnsCase* __restrict__ cases,
    const EnsState* __restrict__ st, const EnsBlock* __restrict__ wmap, unsigned nitems,
    unsigned chunk, unsigned* __restrict__ ctr,
    const unsigned* __restrict__ dcell, const float4* __restrict__ poscell,
    const float4* __restrict__ posf, const float4* __restrict__ velrho,
    const float2* __restrict__ press, float* __restrict__ viscdt, float* __restrict__ ar,
    float* __restrict__ ace) {
  __shared__ TileLds<kRing> lds;
  const unsigned per = nitems / kNumXcd, rem = nitems % kNumXcd;
  const unsigned x0 = blockIdx.x % kNumXcd;
  for (unsigned s = 0; s < kNumXcd; s++) {
    const unsigned x = (x0 + s) % kNumXcd;
    const unsigned seg0 = x * per + min(x, rem), segn = per + (x < rem ? 1u : 0u);
    for (;;) {
      unsigned i0 = 0;
      if (threadIdx.x == 0) i0 = atomicAdd(ctr + x, chunk);
      i0 = __builtin_amdgcn_readfirstlane(i0);
      if (i0 >= segn) break;
      const unsigned i1 = min(i0 + chunk, segn);
      for (unsigned i = i0; i < i1; i++) {
        const EnsBlock e = wmap[seg0 + i];
        if (!st[e.x].active) continue;
        const EnsCase& ec = cases[e.x];
        const unsigned npf = st[e.x].np - ec.npb;
        if (e.y >= npf) continue;
        const unsigned p = e.y + threadIdx.x;
        fluid_tile_body<TVISCO, TDDT, false, kRing, false, false, EOS7>(
            p, p < npf, ec.p0 + ec.npb, ec.viscob, ec.viscof, ec.dvd, ec.c, ec.ks2scan, dcell,
            poscell, posf, velrho, press, nullptr, viscdt, ar, ace, nullptr, nullptr, lds);
      }
    }
  }
}

template <int TVISCO, int TDDT>
__global__ void __launch_bounds__(64) k_ens_fluid_rows(const EnsCase* __restrict__ cases,
    const EnsState* __restrict__ st, const EnsBlock* __restrict__ wmap,
    const unsigned* __restrict__ dcell, const float4* __restrict__ poscell,
    const float4* __restrict__ velrho, const float2* __restrict__ press,
    float* __restrict__ viscdt, float* __restrict__ ar, float* __restrict__ ace) {
  const EnsBlock e = wmap[xcd_block(blockIdx.x, gridDim.x)];
  if (!st[e.x].active) return;
  const EnsCase& ec = cases[e.x];
  const unsigned l = e.y + threadIdx.x;
  if (l >= st[e.x].np - ec.npb) return;
  fluid_rows_body<TVISCO, TDDT, false>(ec.p0 + ec.npb + l, ec.viscob, ec.viscof, ec.dvd, ec.c,
                                       dcell, poscell, velrho, press, nullptr, viscdt, ar, ace,
                                       nullptr, nullptr);
}

// Any block size up to kEnsBsBoundMax; bmap has one entry per block of
// blockDim.x boundary particles.
__global__ void __launch_bounds__(kEnsBsBoundMax) k_ens_bound(const EnsCase* __restrict__ cases,
    const EnsState* __restrict__ st, const EnsBlock* __restrict__ bmap,
    const unsigned* __restrict__ dcell, const float4* __restrict__ poscell,
    const float4* __restrict__ velrho, float* __restrict__ viscdt, float* __restrict__ ar) {
  const EnsBlock e = bmap[xcd_block(blockIdx.x, gridDim.x)];
  if (!st[e.x].active) return;
  const unsigned l = e.y + threadIdx.x;
  if (l >= st[e.x].npbok) return;
  const EnsCase& ec = cases[e.x];
  bound_body<false>(ec.p0 + l, ec.dvd, ec.c, dcell, poscell, velrho, viscdt, ar);
}

// The boundary interaction of an ensemble in the tile form (k_forces_bound_tile
// per case): bmap has one entry per 64 boundary particles.
__global__ void __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(5, 8)))
k_ens_bound_tile(const EnsCase* __restrict__ cases, const EnsState* __restrict__ st,
    const EnsBlock* __restrict__ bmap, const unsigned* __restrict__ dcell,
    const float4* __restrict__ poscell, const float4* __restrict__ posf,
    const float4* __restrict__ velrho, const float2* __restrict__ press,
    float* __restrict__ viscdt, float* __restrict__ ar) {
  __shared__ TileLds<kRingTiles> lds;
  const EnsBlock e = bmap[xcd_block(blockIdx.x, gridDim.x)];
  if (!st[e.x].active) return;
  const unsigned npbok = st[e.x].npbok;
  if (e.y >= npbok) return;
  const EnsCase& ec = cases[e.x];
  const unsigned l = e.y + threadIdx.x;
  fluid_tile_body<DSPH_VISCO_ARTIFICIAL, DSPH_DDT_NONE, false, kRingTiles, false, false, true,
                  true>(l, l < npbok, ec.p0, 0.f, 0.f, ec.dvd, ec.c, ec.ks2scan, dcell, poscell,
                        posf, velrho, press, nullptr, viscdt, ar, nullptr, nullptr, nullptr, lds);
}

template <int TVISCO, int TDDT, bool EOS7>
void launch_ens_fluid(dsph_ctx* ctx, int variant, const EnsCase* cases, const EnsState* st,
                      const EnsBlock* wmap, unsigned nb, const EnsPersist* pers,
                      const unsigned* dcell,
                      const float4* poscell, const float4* posf, const float4* velrho,
                      const float2* press, float* viscdt, float* ar, float* ace) {
  if (pers) {
    const unsigned g = pers->grid;
    switch (variant) {
      case kEnsTileRing8:
        k_ens_fluid_tile_persist<TVISCO, TDDT, 8, EOS7><<<g, 64, 0, ctx->stream>>>(
            cases, st, wmap, nb, pers->chunk, pers->ctr, dcell, poscell, posf, velrho, press,
            viscdt, ar, ace);
        break;
      case kEnsTileRing12:
        k_ens_fluid_tile_persist<TVISCO, TDDT, 12, EOS7><<<g, 64, 0, ctx->stream>>>(
            cases, st, wmap, nb, pers->chunk, pers->ctr, dcell, poscell, posf, velrho, press,
            viscdt, ar, ace);
        break;
      default:
        k_ens_fluid_tile_persist<TVISCO, TDDT, 16, EOS7><<<g, 64, 0, ctx->stream>>>(
            cases, st, wmap, nb, pers->chunk, pers->ctr, dcell, poscell, posf, velrho, press,
            viscdt, ar, ace);
        break;
    }
    return;
  }
  switch (variant) {
    case kEnsTileRing8:
      k_ens_fluid_tile<TVISCO, TDDT, 8, EOS7><<<nb, 64, 0, ctx->stream>>>(
          cases, st, wmap, dcell, poscell, posf, velrho, press, viscdt, ar, ace);
      break;
    case kEnsTileRing12:
      k_ens_fluid_tile<TVISCO, TDDT, 12, EOS7><<<nb, 64, 0, ctx->stream>>>(
          cases, st, wmap, dcell, poscell, posf, velrho, press, viscdt, ar, ace);
      break;
    case kEnsTileRing16:
      k_ens_fluid_tile<TVISCO, TDDT, 16, EOS7><<<nb, 64, 0, ctx->stream>>>(
          cases, st, wmap, dcell, poscell, posf, velrho, press, viscdt, ar, ace);
      break;
    default:
      k_ens_fluid_rows<TVISCO, TDDT><<<nb, 64, 0, ctx->stream>>>(
          cases, st, wmap, dcell, poscell, velrho, press, viscdt, ar, ace);
      break;
  }
}

}  // namespace

namespace dsph {

int ens_prep(dsph_ctx* ctx, const EnsCase* cases, const EnsState* st, const EnsBlock* pmap,
             unsigned npblocks, const float4* poscell, const float4* velrho, float2* press,
             float4* posf) {
  if (!npblocks) return DSPH_OK;
  KTimer kt(ctx, "k_prep");
  k_ens_prep<<<npblocks, kBs, 0, ctx->stream>>>(cases, st, pmap, poscell, velrho, press, posf);
  DSPH_LAUNCHED(ctx, "k_ens_prep");
  return DSPH_OK;
}

int ens_forces_fluid(dsph_ctx* ctx, int variant, int tvisco, int tddt, bool eos7,
                     const EnsCase* cases, const EnsState* st, const EnsBlock* wmap,
                     unsigned nblocks, const EnsPersist* pers, const unsigned* dcell,
                     const float4* poscell,
                     const float4* posf, const float4* velrho, const float2* press,
                     float* viscdt, float* ar, float* ace) {
  if (!nblocks) return DSPH_OK;
  DSPH_REQUIRE(ctx, variant >= 0 && variant < kEnsNumVariants, "ensemble: invalid kernel variant");
  if (pers) {
    DSPH_REQUIRE(ctx, variant != kEnsRows64, "ensemble: the persistent form runs the tile kernel");
    DSPH_REQUIRE(ctx, pers->ctr && pers->grid && pers->chunk,
                 "ensemble: persistent launch without counters, grid or chunk");
    // The counters start at zero (fill kernel: safe inside a captured graph).
    k_fill_u32<<<1, kNumXcd, 0, ctx->stream>>>(pers->ctr, kNumXcd, 0u);
    DSPH_LAUNCHED(ctx, "k_fill_u32");
  }
  KTimer kt(ctx, "k_forces_fluid");
#define DSPH_ENS_FLUID(V, D)                                                              \
  (eos7 ? launch_ens_fluid<V, D, true>(ctx, variant, cases, st, wmap, nblocks, pers, dcell,   \
                                       poscell, posf, velrho, press, viscdt, ar, ace)         \
        : launch_ens_fluid<V, D, false>(ctx, variant, cases, st, wmap, nblocks, pers, dcell,  \
                                        poscell, posf, velrho, press, viscdt, ar, ace))
  if (tvisco == DSPH_VISCO_ARTIFICIAL && tddt == DSPH_DDT_DDT2FULL)
    DSPH_ENS_FLUID(DSPH_VISCO_ARTIFICIAL, DSPH_DDT_DDT2FULL);
  else if (tvisco == DSPH_VISCO_ARTIFICIAL && tddt == DSPH_DDT_NONE)
    DSPH_ENS_FLUID(DSPH_VISCO_ARTIFICIAL, DSPH_DDT_NONE);
  else if (tvisco == DSPH_VISCO_LAMINAR && tddt == DSPH_DDT_DDT2FULL)
    DSPH_ENS_FLUID(DSPH_VISCO_LAMINAR, DSPH_DDT_DDT2FULL);
  else if (tvisco == DSPH_VISCO_LAMINAR && tddt == DSPH_DDT_NONE)
    DSPH_ENS_FLUID(DSPH_VISCO_LAMINAR, DSPH_DDT_NONE);
  else
    return set_error(ctx, DSPH_E_ARG,
                     "ensembles are built for artificial or laminar viscosity with "
                     "DDT2Full or no density diffusion");
#undef DSPH_ENS_FLUID
  DSPH_LAUNCHED(ctx, "k_ens_fluid");
  return DSPH_OK;
}

int ens_forces_bound(dsph_ctx* ctx, const EnsCase* cases, const EnsState* st,
                     const EnsBlock* bmap, unsigned nblocks, unsigned bs, const unsigned* dcell,
                     const float4* poscell, const float4* velrho, float* viscdt, float* ar) {
  if (!nblocks) return DSPH_OK;
  DSPH_REQUIRE(ctx, bs >= 64 && bs <= kEnsBsBoundMax && bs % 64 == 0,
               "ensemble: boundary block size must be a multiple of 64 up to 256");
  KTimer kt(ctx, "k_forces_bound");
  k_ens_bound<<<nblocks, bs, 0, ctx->stream>>>(cases, st, bmap, dcell, poscell, velrho,
                                              viscdt, ar);
  DSPH_LAUNCHED(ctx, "k_ens_bound");
  return DSPH_OK;
}

int ens_forces_bound_tile(dsph_ctx* ctx, const EnsCase* cases, const EnsState* st,
                          const EnsBlock* bmap64, unsigned nblocks, const unsigned* dcell,
                          const float4* poscell, const float4* posf, const float4* velrho,
                          const float2* press, float* viscdt, float* ar) {
  if (!nblocks) return DSPH_OK;
  KTimer kt(ctx, "k_forces_bound");
  k_ens_bound_tile<<<nblocks, 64, 0, ctx->stream>>>(cases, st, bmap64, dcell, poscell, posf,
                                                    velrho, press, viscdt, ar);
  DSPH_LAUNCHED(ctx, "k_ens_bound_tile");
  return DSPH_OK;
}

// Radius^2 of the conservative scan test for a case (see scan_radius2).
float ens_scan_radius2(const dsph_cte& cte) {
  const Cte c = make_cte(cte);
  const double lmax = std::max(std::max(cte.maprealsizex, cte.maprealsizey), cte.maprealsizez) +
                      2.0 * c.pcs;
  const double ulp = lmax * 1.1920928955078125e-7;
  const double r = std::sqrt(double(c.ks2)) + 4.0 * ulp + 1e-6 * std::sqrt(double(c.ks2));
  return float(r * r);
}

}  // namespace dsph

extern "C" {

int dsph_ctx_set_option(dsph_ctx* ctx, int key, int value) {
  if (!ctx) return DSPH_E_ARG;
  switch (key) {
    case DSPH_OPT_FORCES_KERNEL:
      DSPH_REQUIRE(ctx, value == 0 || value == 1, "forces kernel must be 0 (tile) or 1 (rows)");
      ctx->forces_kernel = value;
      return DSPH_OK;
    case DSPH_OPT_BOUND_KERNEL:
      DSPH_REQUIRE(ctx, value >= 0 && value <= 2,
                   "boundary kernel must be 0 (auto), 1 (tile) or 2 (rows)");
      ctx->bound_kernel = value;
      ctx->bound_count = 0;
      if (value) ctx->bound_choice = value;
      return DSPH_OK;
    default: return set_error(ctx, DSPH_E_ARG, "unknown option");
  }
}

}  // extern "C"

namespace {

// Argument checks shared by the interaction entry points.
int check_inter_parms(dsph_ctx* ctx, const dsph_inter_parms* t) {
  if (!ctx->cte_set) return set_error(ctx, DSPH_E_STATE, "constants not set");
  DSPH_REQUIRE(ctx, t->tkernel == DSPH_KERNEL_WENDLAND, "only the Wendland kernel is built");
  DSPH_REQUIRE(ctx, t->np >= t->npb && t->npb >= t->npbok, "invalid particle counts");
  DSPH_REQUIRE(ctx, t->np < DSPH_NP_LIMIT,
               "dsph_interaction_forces: np >= DSPH_NP_LIMIT (2^28: 32-bit byte offsets)");
  DSPH_REQUIRE(ctx, t->d_dcell && t->d_poscell && t->d_velrho && t->d_viscdt && t->d_ar &&
                        t->d_ace && t->divdata.d_beginendcell,
               "dsph_interaction_forces: null array");
  DSPH_REQUIRE(ctx, !t->d_boundmode || t->d_tangenvel, "mDBC2 needs boundmode and tangenvel");
  DSPH_REQUIRE(ctx, !t->d_shiftposfs || (t->shiftmode >= 1 && t->shiftmode <= 3 && t->d_code),
               "shifting needs shiftmode 1..3 and the particle codes");
  DSPH_REQUIRE(ctx, !t->d_shiftposfs || (!t->d_nopenshift && !t->d_velgrad),
               "shifting with mDBC2 no penetration or the velocity gradient is not built");
  DSPH_REQUIRE(ctx, !t->d_ftomassp || t->d_code, "floating bodies need the particle codes");
  DSPH_REQUIRE(ctx, !t->d_ftomassp || (!t->d_shiftposfs && !t->d_boundmode && !t->d_velgrad),
               "floating bodies with shifting, mDBC2 or the velocity gradient are not built");
  DSPH_REQUIRE(ctx, !t->d_nopenshift || (t->d_boundmode && t->d_boundnor),
               "no penetration needs mDBC2 (boundmode, tangenvel) and the boundary normals");
  const bool sps = t->tvisco == DSPH_VISCO_LAMINARSPS;
  DSPH_REQUIRE(ctx, !sps || (t->d_spstaurho2 && t->d_sps2strain),
               "Laminar+SPS needs spstaurho2 and sps2strain");
  return DSPH_OK;
}

// Interactions between two trials of the boundary kernel's forms (auto).
constexpr unsigned long long kBoundRetrial = 4096;

// The interaction with the per-particle preparation in press/posf (np
// entries each): k_prep, the fluid kernel, the boundary kernel. allow_trial:
// the boundary kernel's auto choice may time its forms (events, one host
// synchronisation per trial); without it the current choice runs.
int interaction_forces(dsph_ctx* ctx, const dsph_inter_parms* t, float2* press, float4* posf,
                       bool allow_trial) {
  const Cte c = make_cte(ctx->cte);
  const DivData dvd = make_divdata(t->divdata);
  const unsigned npf = t->np - t->npb;
  int rc = DSPH_OK;
  if (npf) {
    {
      KTimer kt(ctx, "k_prep");
      k_prep<<<grid_for(t->np, kBs), kBs, 0, ctx->stream>>>(
          t->np, (const float4*)t->d_poscell, (const float4*)t->d_velrho, press,
          ctx->forces_kernel == 0 ? posf : nullptr, c);
    }
    const bool vg = t->d_velgrad != nullptr;
    switch (t->tvisco) {
      case DSPH_VISCO_ARTIFICIAL:
        rc = vg ? dispatch_ddt<DSPH_VISCO_ARTIFICIAL, true>(ctx, t, dvd, c, press, posf)
                : dispatch_ddt<DSPH_VISCO_ARTIFICIAL, false>(ctx, t, dvd, c, press, posf);
        break;
      case DSPH_VISCO_LAMINAR:
        rc = vg ? dispatch_ddt<DSPH_VISCO_LAMINAR, true>(ctx, t, dvd, c, press, posf)
                : dispatch_ddt<DSPH_VISCO_LAMINAR, false>(ctx, t, dvd, c, press, posf);
        break;
      case DSPH_VISCO_LAMINARSPS:
        rc = vg ? dispatch_ddt<DSPH_VISCO_LAMINARSPS, true>(ctx, t, dvd, c, press, posf)
                : dispatch_ddt<DSPH_VISCO_LAMINARSPS, false>(ctx, t, dvd, c, press, posf);
        break;
      default: return set_error(ctx, DSPH_E_ARG, "dsph_interaction_forces: invalid tvisco");
    }
    if (rc) return rc;
    DSPH_LAUNCHED(ctx, "k_forces_fluid");
  }
  if (t->npbok) {
    KTimer kt(ctx, "k_forces_bound");
    if (ctx->forces_kernel == 0 && !t->d_ftomassp && npf) {
      // Tile or one-lane form (same pairs, same order, same bits): which is
      // faster depends on how many walls are wet, so by default it is timed.
      hipEvent_t* ev = nullptr;
      int form = ctx->bound_kernel ? ctx->bound_kernel : ctx->bound_choice;
      hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
      if (!ctx->bound_kernel && allow_trial &&
          hipStreamIsCapturing(ctx->stream, &cs) == hipSuccess &&
          cs == hipStreamCaptureStatusNone) {
        // Not while a graph is being captured: the count only advances on
        // eager interactions, so the events of a trial are always recorded.
        const unsigned long long k = ctx->bound_count++ % kBoundRetrial;
        if (k < 5) {
          if (k < 4) {
            form = k < 2 ? 1 : 2;
            if (k == 1 || k == 3) {
              hipEvent_t* e = ctx->bound_ev + (k == 1 ? 0 : 2);
              for (int q = 0; q < 2; q++)
                if (!e[q]) DSPH_HIP(ctx, hipEventCreate(&e[q]));
              ev = e;
            }
          } else {
            float ms_tile = 0, ms_rows = 0;
            DSPH_HIP(ctx, hipEventSynchronize(ctx->bound_ev[3]));
            DSPH_HIP(ctx, hipEventElapsedTime(&ms_tile, ctx->bound_ev[0], ctx->bound_ev[1]));
            DSPH_HIP(ctx, hipEventElapsedTime(&ms_rows, ctx->bound_ev[2], ctx->bound_ev[3]));
            ctx->bound_choice = ms_rows < ms_tile ? 2 : 1;
            form = ctx->bound_choice;
          }
        }
      }
      if (ev) DSPH_HIP(ctx, hipEventRecord(ev[0], ctx->stream));
      if (form == 1)
        k_forces_bound_tile<<<grid_for(t->npbok, 64), 64, 0, ctx->stream>>>(
            t->npbok, dvd, c, scan_radius2(ctx, c), t->d_dcell, (const float4*)t->d_poscell,
            posf, (const float4*)t->d_velrho, press, t->d_viscdt, t->d_ar);
      else
        k_forces_bound<false><<<grid_for(t->npbok, kBsForces), kBsForces, 0, ctx->stream>>>(
            t->npbok, dvd, c, t->d_dcell, (const float4*)t->d_poscell,
            (const float4*)t->d_velrho, t->d_viscdt, t->d_ar);
      if (ev) DSPH_HIP(ctx, hipEventRecord(ev[1], ctx->stream));
    } else if (t->d_ftomassp)
      k_forces_bound<true><<<grid_for(t->npbok, kBsForces), kBsForces, 0, ctx->stream>>>(
          t->npbok, dvd, c, t->d_dcell, (const float4*)t->d_poscell,
          (const float4*)t->d_velrho, t->d_viscdt, t->d_ar, t->d_code, t->d_ftomassp);
    else
      k_forces_bound<false><<<grid_for(t->npbok, kBsForces), kBsForces, 0, ctx->stream>>>(
          t->npbok, dvd, c, t->d_dcell, (const float4*)t->d_poscell,
          (const float4*)t->d_velrho, t->d_viscdt, t->d_ar);
    DSPH_LAUNCHED(ctx, "k_forces_bound");
  }
  return DSPH_OK;
}

// Scratch of dsph_interaction_forces_noalloc: {press, 1/rho} (float2) for
// DSPH_NOALLOC_NP_MAX particles, then their map-frame positions (float4).
// Every GPU holds its own copy for the life of the process; host code uses
// the device address that hipGetSymbolAddress gives on each device, never
// the array's name.
constexpr size_t kNoallocPressBytes = size_t(DSPH_NOALLOC_NP_MAX) * sizeof(float2);
constexpr size_t kNoallocBytes = kNoallocPressBytes + size_t(DSPH_NOALLOC_NP_MAX) * sizeof(float4);
static __device__ __align__(256) unsigned char scratch[kNoallocBytes];

}  // namespace

namespace dsph {

// Device address of the no-alloc scratch on the current device (the
// context's device), looked up once per device.
int noalloc_scratch_init(dsph_ctx* ctx) {
  static std::mutex mu;
  static std::map<int, void*> addr;
  std::lock_guard<std::mutex> lock(mu);
  auto it = addr.find(ctx->device);
  if (it == addr.end()) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return set_error(ctx, DSPH_E_LAUNCH,
                       std::string("hipGetSymbolAddress(no-alloc scratch): ") +
                           hipGetErrorString(e));
    it = addr.emplace(ctx->device, ptr).first;
  }
  ctx->noalloc_scratch = it->second;
  return DSPH_OK;
}

}  // namespace dsph

extern "C" {

int dsph_interaction_forces(dsph_ctx* ctx, const dsph_inter_parms* t) {
  if (!ctx || !t) return DSPH_E_ARG;
  int rc = check_inter_parms(ctx, t);
  if (rc) return rc;
  if (t->np - t->npb) {
    rc = ensure_prep_scratch(ctx, t->np);
    if (rc) return rc;
  }
  return interaction_forces(ctx, t, reinterpret_cast<float2*>(ctx->prep_press),
                            reinterpret_cast<float4*>(ctx->prep_posf), true);
}

int dsph_interaction_forces_noalloc(dsph_ctx* ctx, const dsph_inter_parms* t) {
  if (!ctx || !t) return DSPH_E_ARG;
  DSPH_REQUIRE(ctx, t->np <= DSPH_NOALLOC_NP_MAX,
               "dsph_interaction_forces_noalloc: np " + std::to_string(t->np) +
                   " > DSPH_NOALLOC_NP_MAX (" + std::to_string(DSPH_NOALLOC_NP_MAX) +
                   "); use dsph_interaction_forces");
  int rc = check_inter_parms(ctx, t);
  if (rc) return rc;
  if (!ctx->noalloc_scratch)
    return set_error(ctx, DSPH_E_STATE, "no-alloc scratch address not set up for this device");
  unsigned char* base = static_cast<unsigned char*>(ctx->noalloc_scratch);
  return interaction_forces(ctx, t, reinterpret_cast<float2*>(base),
                            reinterpret_cast<float4*>(base + kNoallocPressBytes), false);
}

}  // extern "C"
