// Particle interaction kernels of the SPH step on gfx950 (FP32 pair
// arithmetic): fluid-fluid + fluid-boundary forces, boundary-fluid continuity,
// Smagorinsky SPS stress, mDBC ghost-node density, and the PIPS pair counter.
//
// Physics and summation order follow the reference GPU kernels
// (KerInteractionForcesFluid / KerInteractionForcesFluidBox /
// KerInteractionForcesBound, JSphGpu_ker.cu:450-990): for every particle the
// neighbour cell rows are visited z-outer, y-inner, fluid cells first and then
// boundary cells, particles of a row in ascending sorted index, and all
// contributions go into one set of accumulators. Distances come from the
// cell-relative FP32 PosCell. The pressure of every particle is evaluated once
// per interaction (same function and input as the reference's per-pair call,
// so each pair sees the identical value).
#include <cfloat>
#include <cstring>
#include "dsph_ctx.h"

using namespace dsph;

namespace {

struct Cte {
  float massb, massf, h, ks2, pcs, awen, bwenh, cs0, eta2, ddtkh, ddtgz, rhopzero,
      invgamma, cteb, gamma, ovrhopzero, dp;
  unsigned cellcode;
};

Cte make_cte(const dsph_cte& c) {
  Cte r;
  r.massb = c.massb; r.massf = c.massf; r.h = c.kernelh; r.ks2 = c.kernelsize2;
  r.pcs = c.poscellsize; r.awen = c.awen; r.bwenh = c.bwenh; r.cs0 = c.cs0; r.eta2 = c.eta2;
  r.ddtkh = c.ddtkh; r.ddtgz = c.ddtgz; r.rhopzero = c.rhopzero; r.invgamma = 1.f / c.gamma;
  r.cteb = c.cteb; r.gamma = c.gamma; r.ovrhopzero = c.ovrhopzero; r.dp = c.dp;
  r.cellcode = c.cellcode;
  return r;
}

// Tait EOS of FunSphEos_iker.h:37-59 (B*((rho/rho0)^gamma-1), rho/rho0 as
// rho*(1/rho0)).
__device__ inline float eos_press(float rho, const Cte& c) {
  return c.cteb * (powf(rho * c.ovrhopzero, c.gamma) - 1.0f);
}

// Wendland C2 gradient factor (FunSphKernel_iker.h:117-124).
__device__ inline float wendland_fac(float rr2, const Cte& c) {
  const float rad = sqrtf(rr2);
  const float qq = rad / c.h;
  const float wqq1 = 1.f - 0.5f * qq;
  return c.bwenh * wqq1 * wqq1 * wqq1;
}
// Wendland C2 kernel value and gradient factor (FunSphKernel_iker.h:106-140).
__device__ inline float wendland_wab_fac(float rr2, const Cte& c, float& fac) {
  const float rad = sqrtf(rr2);
  const float qq = rad / c.h;
  const float wqq1 = 1.f - 0.5f * qq;
  const float wqq2 = wqq1 * wqq1;
  fac = c.bwenh * wqq2 * wqq1;
  const float wqq = qq + qq + 1.f;
  return c.awen * wqq * wqq2 * wqq2;
}

constexpr unsigned kBsForces = 128;
constexpr unsigned kBs = 256;

__global__ void __launch_bounds__(kBs) k_press(unsigned n, const float4* __restrict__ velrho,
                                               float* __restrict__ press, Cte c) {
  const unsigned p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) press[p] = eos_press(velrho[p].w, c);
}

struct Acc {
  float3 ace;
  float ar, visc, delta;
  float s[6];   // two_strain {xx,xy,xz,yy,yz,zz}
  float g[9];   // full velocity gradient (learned closure features)
};

// One fluid particle p1 against the particles [pini,pfin) of a cell row
// (KerInteractionForcesFluidBox, JSphGpu_ker.cu:539-777, in-scope options).
template <int TVISCO, int TDDT, bool BOUNDP2, bool VELGRAD>
__device__ inline void fluid_row(unsigned pini, unsigned pfin, float visco, float massp2,
                                 const float4& ps1, const float4& vr1, float press1,
                                 const float* tau1, const float4* __restrict__ poscell,
                                 const float4* __restrict__ velrho,
                                 const float* __restrict__ press,
                                 const float* __restrict__ tau, const Cte& c, Acc& a) {
  for (unsigned p2 = pini; p2 < pfin; p2++) {
    const float4 ps2 = poscell[p2];
    const float3 dr = pscel_dr(ps1, ps2, c.pcs);
    const float rr2 = dr.x * dr.x + dr.y * dr.y + dr.z * dr.z;
    if (rr2 <= c.ks2 && rr2 >= kAlmostZero) {
      const float fac = wendland_fac(rr2, c);
      const float frx = fac * dr.x, fry = fac * dr.y, frz = fac * dr.z;
      const float4 vr2 = velrho[p2];
      // Momentum: pressure term.
      {
        const float press2 = press[p2];
        const float prs = (press1 + press2) / (vr1.w * vr2.w);
        const float p_vpm = -prs * massp2;
        a.ace.x += p_vpm * frx; a.ace.y += p_vpm * fry; a.ace.z += p_vpm * frz;
      }
      // Continuity.
      const float dvx = vr1.x - vr2.x, dvy = vr1.y - vr2.y, dvz = vr1.z - vr2.z;
      a.ar += massp2 * (dvx * frx + dvy * fry + dvz * frz) * (vr1.w / vr2.w);
      const float cbar = c.cs0;
      const float dot3 = (TDDT != DSPH_DDT_NONE) ? dr.x * frx + dr.y * fry + dr.z * frz : 0.f;
      // Density diffusion (Molteni & Colagrossi 2009), DBC boundaries.
      if (TDDT == DSPH_DDT_DDT && a.delta != FLT_MAX) {
        const float rhop1over2 = vr1.w / vr2.w;
        const float visc_densi = c.ddtkh * cbar * (rhop1over2 - 1.f) / (rr2 + c.eta2);
        const float delta = visc_densi * dot3 * massp2;
        a.delta = (BOUNDP2 ? FLT_MAX : a.delta + delta);
      }
      // Density diffusion (Fourtakas et al. 2019).
      if ((TDDT == DSPH_DDT_DDT2 || (TDDT == DSPH_DDT_DDT2FULL && !BOUNDP2)) &&
          a.delta != FLT_MAX) {
        const float rh = 1.f + c.ddtgz * dr.z;
        const float drho = c.rhopzero * powf(rh, c.invgamma) - c.rhopzero;
        const float visc_densi = c.ddtkh * cbar * ((vr2.w - vr1.w) - drho) / (rr2 + c.eta2);
        const float delta = visc_densi * dot3 * massp2 / vr2.w;
        a.delta = (BOUNDP2 ? FLT_MAX : a.delta - delta);
      }
      // Viscosity.
      const float dot = dr.x * dvx + dr.y * dvy + dr.z * dvz;
      const float dot_rr2 = dot / (rr2 + c.eta2);
      a.visc = fmaxf(dot_rr2, a.visc);
      if (TVISCO == DSPH_VISCO_ARTIFICIAL) {
        if (dot < 0) {
          const float amubar = c.h * dot_rr2;
          const float robar = (vr1.w + vr2.w) * 0.5f;
          const float pi_visc = (-visco * cbar * amubar / robar) * massp2;
          a.ace.x -= pi_visc * frx; a.ace.y -= pi_visc * fry; a.ace.z -= pi_visc * frz;
        }
      } else {
        const float robar2 = (vr1.w + vr2.w);
        const float temp = 4.f * visco / ((rr2 + c.eta2) * robar2);
        const float vtemp = massp2 * temp * (dr.x * frx + dr.y * fry + dr.z * frz);
        a.ace.x += vtemp * dvx; a.ace.y += vtemp * dvy; a.ace.z += vtemp * dvz;
      }
      if (TVISCO == DSPH_VISCO_LAMINARSPS) {
        float sxx = tau1[0], sxy = tau1[1], sxz = tau1[2], syy = tau1[3], syz = tau1[4],
              szz = tau1[5];
        if (!BOUNDP2) {
          const float* t2 = tau + 6 * size_t(p2);
          sxx += t2[0]; sxy += t2[1]; sxz += t2[2]; syy += t2[3]; syz += t2[4]; szz += t2[5];
        }
        a.ace.x += massp2 * (sxx * frx + sxy * fry + sxz * frz);
        a.ace.y += massp2 * (sxy * frx + syy * fry + syz * frz);
        a.ace.z += massp2 * (sxz * frx + syz * fry + szz * frz);
        const float volp2 = -massp2 / vr2.w;
        float dv = dvx * volp2; a.s[0] += dv * frx; a.s[1] += dv * fry; a.s[2] += dv * frz;
        dv = dvy * volp2;       a.s[1] += dv * frx; a.s[3] += dv * fry; a.s[4] += dv * frz;
        dv = dvz * volp2;       a.s[2] += dv * frx; a.s[4] += dv * fry; a.s[5] += dv * frz;
      }
      if (VELGRAD) {
        // Full gradient G_ij = sum_b -m_b/rho_b (v_a-v_b)_i dW/dx_j.
        const float volp2 = -massp2 / vr2.w;
        const float ux = dvx * volp2, uy = dvy * volp2, uz = dvz * volp2;
        a.g[0] += ux * frx; a.g[1] += ux * fry; a.g[2] += ux * frz;
        a.g[3] += uy * frx; a.g[4] += uy * fry; a.g[5] += uy * frz;
        a.g[6] += uz * frx; a.g[7] += uz * fry; a.g[8] += uz * frz;
      }
    }
  }
}

template <int TVISCO, int TDDT, bool VELGRAD>
__global__ void __launch_bounds__(kBsForces) k_forces_fluid(
    unsigned n, unsigned pinit, float viscob, float viscof, DivData dvd, Cte c,
    const unsigned* __restrict__ dcell, const float4* __restrict__ poscell,
    const float4* __restrict__ velrho, const float* __restrict__ press,
    const float* __restrict__ tau, float* __restrict__ viscdt, float* __restrict__ ar,
    float* __restrict__ ace, float* __restrict__ strain, float* __restrict__ velgrad) {
  const unsigned p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const unsigned p1 = p + pinit;
  Acc a;
  a.ace = make_float3(0, 0, 0); a.ar = 0; a.visc = 0; a.delta = 0;
#pragma unroll
  for (int k = 0; k < 6; k++) a.s[k] = 0;
#pragma unroll
  for (int k = 0; k < 9; k++) a.g[k] = 0;
  const float4 ps1 = poscell[p1];
  const float4 vr1 = velrho[p1];
  const float press1 = press[p1];
  float tau1[6] = {0, 0, 0, 0, 0, 0};
  if (TVISCO == DSPH_VISCO_LAMINARSPS) {
#pragma unroll
    for (int k = 0; k < 6; k++) tau1[k] = tau[6 * size_t(p1) + k];
  }
  const unsigned cel = dcell[p1];
  const int cx = int(dcel_cellx(dvd.domcellcode, cel)) - dvd.czx;
  const int cy = int(dcel_celly(dvd.domcellcode, cel)) - dvd.czy;
  const int cz = int(dcel_cellz(dvd.domcellcode, cel)) - dvd.czz;
  const NgLimits l = ng_limits(cx, cy, cz, dvd);
  // Fluid neighbours.
  for (int z = l.z0; z < l.z1; z++)
    for (int y = l.y0; y < l.y1; y++) {
      unsigned pini, pfin;
      row_range(dvd.begincell, int(dvd.cellfluid) + z * dvd.ncw + y * dvd.ncx, l.x0, l.x1,
                pini, pfin);
      if (pfin)
        fluid_row<TVISCO, TDDT, false, VELGRAD>(pini, pfin, viscof, c.massf, ps1, vr1, press1,
                                                tau1, poscell, velrho, press, tau, c, a);
    }
  // Boundary neighbours.
  for (int z = l.z0; z < l.z1; z++)
    for (int y = l.y0; y < l.y1; y++) {
      unsigned pini, pfin;
      row_range(dvd.begincell, z * dvd.ncw + y * dvd.ncx, l.x0, l.x1, pini, pfin);
      if (pfin)
        fluid_row<TVISCO, TDDT, true, VELGRAD>(pini, pfin, viscob, c.massb, ps1, vr1, press1,
                                               tau1, poscell, velrho, press, tau, c, a);
    }
  if (a.ar != 0 || a.ace.x != 0 || a.ace.y != 0 || a.ace.z != 0 || a.visc != 0) {
    float arp1 = a.ar;
    if (TDDT != DSPH_DDT_NONE && a.delta != FLT_MAX) arp1 += a.delta;
    ar[p1] += arp1;
    float* ap = ace + 3 * size_t(p1);
    ap[0] += a.ace.x; ap[1] += a.ace.y; ap[2] += a.ace.z;
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

// Boundary particle p1 against fluid neighbours (KerInteractionForcesBound,
// JSphGpu_ker.cu:450-533).
__global__ void __launch_bounds__(kBsForces) k_forces_bound(unsigned n, DivData dvd, Cte c,
    const unsigned* __restrict__ dcell, const float4* __restrict__ poscell,
    const float4* __restrict__ velrho, float* __restrict__ viscdt, float* __restrict__ ar) {
  const unsigned p1 = blockIdx.x * blockDim.x + threadIdx.x;
  if (p1 >= n) return;
  float visc = 0, arp1 = 0;
  const float4 ps1 = poscell[p1];
  const float4 vr1 = velrho[p1];
  const unsigned cel = dcell[p1];
  const int cx = int(dcel_cellx(dvd.domcellcode, cel)) - dvd.czx;
  const int cy = int(dcel_celly(dvd.domcellcode, cel)) - dvd.czy;
  const int cz = int(dcel_cellz(dvd.domcellcode, cel)) - dvd.czz;
  const NgLimits l = ng_limits(cx, cy, cz, dvd);
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
          const float fac = wendland_fac(rr2, c);
          const float frx = fac * dr.x, fry = fac * dr.y, frz = fac * dr.z;
          const float4 vr2 = velrho[p2];
          const float dvx = vr1.x - vr2.x, dvy = vr1.y - vr2.y, dvz = vr1.z - vr2.z;
          arp1 += c.massf * (dvx * frx + dvy * fry + dvz * frz) * (vr1.w / vr2.w);
          const float dot = dr.x * dvx + dr.y * dvy + dr.z * dvz;
          const float dot_rr2 = dot / (rr2 + c.eta2);
          visc = fmaxf(dot_rr2, visc);
        }
      }
    }
  if (arp1 != 0 || visc != 0) {
    ar[p1] += arp1;
    if (visc > viscdt[p1]) viscdt[p1] = visc;
  }
}

// Smagorinsky + Blin SPS stress tau/rho^2 (KerComputeSpsTau, JSphGpu_ker.cu:1387-1420).
__global__ void __launch_bounds__(kBs) k_sps_tau(unsigned n, unsigned pini, float smag,
    float blin, const float4* __restrict__ velrho, const float* __restrict__ s2,
    float* __restrict__ tau) {
  const unsigned p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const unsigned p1 = p + pini;
  const float* s = s2 + 6 * size_t(p1);
  const float xx = s[0], xy = s[1], xz = s[2], yy = s[3], yz = s[4], zz = s[5];
  const float pow1 = xx * xx + yy * yy + zz * zz;
  const float prr = xy * xy + xz * xz + yz * yz + pow1 + pow1;
  const float visc_sps = smag * sqrtf(prr);
  const float div_u = xx + yy + zz;
  const float sps_k = visc_sps * div_u;
  const float sps_blin = blin * prr;
  const float sumsps = -(sps_k + sps_blin);
  const float twovisc_sps = (visc_sps + visc_sps);
  const float one_rho = 1.0f / velrho[p1].w;
  float* t = tau + 6 * size_t(p1);
  t[0] = one_rho * (twovisc_sps * xx + sumsps);
  t[1] = one_rho * (visc_sps * xy);
  t[2] = one_rho * (visc_sps * xz);
  t[3] = one_rho * (twovisc_sps * yy + sumsps);
  t[4] = one_rho * (visc_sps * yz);
  t[5] = one_rho * (twovisc_sps * zz + sumsps);
}

// ---------------------------------------------------------------- mDBC
// Determinant and inverse of a 4x4 matrix in FP64 by expansion in 2x2 minors
// of the first two rows (Laplace expansion).
struct M4d { double m[4][4]; };

__device__ inline double det4_inv(const M4d& a, M4d& inv) {
  const double s0 = a.m[0][0] * a.m[1][1] - a.m[1][0] * a.m[0][1];
  const double s1 = a.m[0][0] * a.m[1][2] - a.m[1][0] * a.m[0][2];
  const double s2 = a.m[0][0] * a.m[1][3] - a.m[1][0] * a.m[0][3];
  const double s3 = a.m[0][1] * a.m[1][2] - a.m[1][1] * a.m[0][2];
  const double s4 = a.m[0][1] * a.m[1][3] - a.m[1][1] * a.m[0][3];
  const double s5 = a.m[0][2] * a.m[1][3] - a.m[1][2] * a.m[0][3];
  const double c5 = a.m[2][2] * a.m[3][3] - a.m[3][2] * a.m[2][3];
  const double c4 = a.m[2][1] * a.m[3][3] - a.m[3][1] * a.m[2][3];
  const double c3 = a.m[2][1] * a.m[3][2] - a.m[3][1] * a.m[2][2];
  const double c2 = a.m[2][0] * a.m[3][3] - a.m[3][0] * a.m[2][3];
  const double c1 = a.m[2][0] * a.m[3][2] - a.m[3][0] * a.m[2][2];
  const double c0 = a.m[2][0] * a.m[3][1] - a.m[3][0] * a.m[2][1];
  const double det = s0 * c5 - s1 * c4 + s2 * c3 + s3 * c2 - s4 * c1 + s5 * c0;
  if (det != 0) {
    const double id = 1.0 / det;
    inv.m[0][0] = (a.m[1][1] * c5 - a.m[1][2] * c4 + a.m[1][3] * c3) * id;
    inv.m[0][1] = (-a.m[0][1] * c5 + a.m[0][2] * c4 - a.m[0][3] * c3) * id;
    inv.m[0][2] = (a.m[3][1] * s5 - a.m[3][2] * s4 + a.m[3][3] * s3) * id;
    inv.m[0][3] = (-a.m[2][1] * s5 + a.m[2][2] * s4 - a.m[2][3] * s3) * id;
    inv.m[1][0] = (-a.m[1][0] * c5 + a.m[1][2] * c2 - a.m[1][3] * c1) * id;
    inv.m[1][1] = (a.m[0][0] * c5 - a.m[0][2] * c2 + a.m[0][3] * c1) * id;
    inv.m[1][2] = (-a.m[3][0] * s5 + a.m[3][2] * s2 - a.m[3][3] * s1) * id;
    inv.m[1][3] = (a.m[2][0] * s5 - a.m[2][2] * s2 + a.m[2][3] * s1) * id;
    inv.m[2][0] = (a.m[1][0] * c4 - a.m[1][1] * c2 + a.m[1][3] * c0) * id;
    inv.m[2][1] = (-a.m[0][0] * c4 + a.m[0][1] * c2 - a.m[0][3] * c0) * id;
    inv.m[2][2] = (a.m[3][0] * s4 - a.m[3][1] * s2 + a.m[3][3] * s0) * id;
    inv.m[2][3] = (-a.m[2][0] * s4 + a.m[2][1] * s2 - a.m[2][3] * s0) * id;
    inv.m[3][0] = (-a.m[1][0] * c3 + a.m[1][1] * c1 - a.m[1][2] * c0) * id;
    inv.m[3][1] = (a.m[0][0] * c3 - a.m[0][1] * c1 + a.m[0][2] * c0) * id;
    inv.m[3][2] = (-a.m[3][0] * s3 + a.m[3][1] * s1 - a.m[3][2] * s0) * id;
    inv.m[3][3] = (a.m[2][0] * s3 - a.m[2][1] * s1 + a.m[2][2] * s0) * id;
  }
  return det;
}

// mDBC, SlipMode SLIP_Vel0 (KerInteractionMdbcCorrection_Fast,
// JSphGpu_mdbc_iker.cu:34-150): Shepard density and gradient at the ghost node
// g = x_b + boundnor with a first-order (4x4) correction; mirrored back to
// the boundary particle by a linear extrapolation.
__global__ void __launch_bounds__(kBsForces) k_mdbc(unsigned n, DivData dvd, Cte c,
    double3 mapposmin, const double2* __restrict__ posxy, const double* __restrict__ posz,
    const float4* __restrict__ poscell, const typecode* __restrict__ code,
    const float* __restrict__ boundnor, float4* __restrict__ velrho) {
  const unsigned p1 = blockIdx.x * blockDim.x + threadIdx.x;
  if (p1 >= n) return;
  const float bnx = boundnor[3 * size_t(p1)], bny = boundnor[3 * size_t(p1) + 1],
              bnz = boundnor[3 * size_t(p1) + 2];
  if (bnx == 0 && bny == 0 && bnz == 0) return;
  const double gx = posxy[p1].x + bnx, gy = posxy[p1].y + bny, gz = posz[p1] + bnz;
  // PosCell of the ghost node (KerComputePosCell, JSphGpu_ker.cu:1185-1202).
  float4 gps;
  {
    const double dx = gx - mapposmin.x, dy = gy - mapposmin.y, dz = gz - mapposmin.z;
    const unsigned cx = unsigned(dx / c.pcs), cy = unsigned(dy / c.pcs), cz = unsigned(dz / c.pcs);
    gps = make_float4(float(dx - double(c.pcs) * cx), float(dy - double(c.pcs) * cy),
                      float(dz - double(c.pcs) * cz), __uint_as_float(pscel_code(cx, cy, cz)));
  }
  float rhop1 = 0, grx = 0, gry = 0, grz = 0;
  float A[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) A[i][j] = 0;
  const int cx = int((gx - dvd.domposminx) / dvd.scell) - dvd.czx;
  const int cy = int((gy - dvd.domposminy) / dvd.scell) - dvd.czy;
  const int cz = int((gz - dvd.domposminz) / dvd.scell) - dvd.czz;
  const NgLimits l = ng_limits(cx, cy, cz, dvd);
  for (int z = l.z0; z < l.z1; z++)
    for (int y = l.y0; y < l.y1; y++) {
      unsigned pini, pfin;
      row_range(dvd.begincell, int(dvd.cellfluid) + z * dvd.ncw + y * dvd.ncx, l.x0, l.x1,
                pini, pfin);
      for (unsigned p2 = pini; p2 < pfin; p2++) {
        const float4 ps2 = poscell[p2];
        const float3 dr = pscel_dr(gps, ps2, c.pcs);
        const float rr2 = dr.x * dr.x + dr.y * dr.y + dr.z * dr.z;
        if (rr2 <= c.ks2 && code_is_fluid(code[p2])) {
          float fac;
          const float wab = wendland_wab_fac(rr2, c, fac);
          const float frx = fac * dr.x, fry = fac * dr.y, frz = fac * dr.z;
          const float4 vr2 = velrho[p2];
          const float massp2 = c.massf;
          const float volp2 = massp2 / vr2.w;
          rhop1 += massp2 * wab;
          grx += massp2 * frx; gry += massp2 * fry; grz += massp2 * frz;
          const float vwab = wab * volp2, vfrx = frx * volp2, vfry = fry * volp2,
                      vfrz = frz * volp2;
          A[0][0] += vwab; A[0][1] += dr.x * vwab; A[0][2] += dr.y * vwab; A[0][3] += dr.z * vwab;
          A[1][0] += vfrx; A[1][1] += dr.x * vfrx; A[1][2] += dr.y * vfrx; A[1][3] += dr.z * vfrx;
          A[2][0] += vfry; A[2][1] += dr.x * vfry; A[2][2] += dr.y * vfry; A[2][3] += dr.z * vfry;
          A[3][0] += vfrz; A[3][1] += dr.x * vfrz; A[3][2] += dr.y * vfrz; A[3][3] += dr.z * vfrz;
        }
      }
    }
  float rhofinal = FLT_MAX;
  M4d Ad, inv;
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) Ad.m[i][j] = double(A[i][j]);
  const double det = det4_inv(Ad, inv);
  if (fabs(det) >= 1e-3) {
    float fi[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j < 4; j++) fi[i][j] = float(inv.m[i][j]);
    const float rhoghost = float(fi[0][0] * rhop1 + fi[0][1] * grx + fi[0][2] * gry + fi[0][3] * grz);
    const float gx_ = -float(fi[1][0] * rhop1 + fi[1][1] * grx + fi[1][2] * gry + fi[1][3] * grz);
    const float gy_ = -float(fi[2][0] * rhop1 + fi[2][1] * grx + fi[2][2] * gry + fi[2][3] * grz);
    const float gz_ = -float(fi[3][0] * rhop1 + fi[3][1] * grx + fi[3][2] * gry + fi[3][3] * grz);
    rhofinal = (rhoghost + gx_ * (-bnx) + gy_ * (-bny) + gz_ * (-bnz));
  } else if (A[0][0] > 0) {
    rhofinal = float(rhop1 / A[0][0]);
  }
  velrho[p1].w = (rhofinal != FLT_MAX ? rhofinal : c.rhopzero);
}

// ---------------------------------------------------------------- PIPS
// Real and checked neighbour counts (KerInteractionNg, JDsPips_ker.cu:54-124):
// boundary p1 in [pinitb,pinitb+nb) count fluid neighbours; fluid p1 in
// [pinitf,pinitf+nf) count fluid and boundary neighbours.
__global__ void __launch_bounds__(kBs) k_pips(unsigned nb, unsigned pinitb, unsigned nf,
    unsigned pinitf, DivData dvd, const unsigned* __restrict__ dcell,
    const float4* __restrict__ poscell, unsigned long long* __restrict__ res) {
  const unsigned p = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long rf = 0, rb = 0, cf = 0, cb = 0, rfb = 0;
  if (p < nb + nf) {
    const bool fluid = p >= nb;
    const unsigned p1 = fluid ? p - nb + pinitf : p + pinitb;
    const float4 ps1 = poscell[p1];
    const unsigned cel = dcell[p1];
    const int cx = int(dcel_cellx(dvd.domcellcode, cel)) - dvd.czx;
    const int cy = int(dcel_celly(dvd.domcellcode, cel)) - dvd.czy;
    const int cz = int(dcel_cellz(dvd.domcellcode, cel)) - dvd.czz;
    const NgLimits l = ng_limits(cx, cy, cz, dvd);
    unsigned numr[2] = {0, 0}, numc = 0;
    for (int pass = 0; pass < (fluid ? 2 : 1); pass++) {
      const int base = pass == 0 ? int(dvd.cellfluid) : 0;
      for (int z = l.z0; z < l.z1; z++)
        for (int y = l.y0; y < l.y1; y++) {
          unsigned pini, pfin;
          row_range(dvd.begincell, base + z * dvd.ncw + y * dvd.ncx, l.x0, l.x1, pini, pfin);
          for (unsigned p2 = pini; p2 < pfin; p2++) {
            const float3 dr = pscel_dr(ps1, poscell[p2], dvd.poscellsize);
            const float rr2 = dr.x * dr.x + dr.y * dr.y + dr.z * dr.z;
            if (rr2 <= dvd.kernelsize2 && rr2 >= kAlmostZero) numr[pass]++;
          }
          numc += pfin - pini;
        }
    }
    if (fluid) { rf = numr[0] + numr[1]; rfb = numr[1]; cf = numc; }
    else { rb = numr[0]; cb = numc; }
  }
  for (int off = 32; off > 0; off >>= 1) {
    rf += __shfl_xor(rf, off); rb += __shfl_xor(rb, off);
    cf += __shfl_xor(cf, off); cb += __shfl_xor(cb, off);
    rfb += __shfl_xor(rfb, off);
  }
  if ((threadIdx.x & 63) == 0) {
    if (rf) atomicAdd(res + 0, rf);
    if (rb) atomicAdd(res + 1, rb);
    if (cf) atomicAdd(res + 2, cf);
    if (cb) atomicAdd(res + 3, cb);
    if (rfb) atomicAdd(res + 4, rfb);
  }
}

__global__ void k_zero_u64(unsigned long long* a, unsigned n) {
  if (threadIdx.x < n) a[threadIdx.x] = 0ull;
}

template <int TVISCO, int TDDT, bool VELGRAD>
void launch_fluid(dsph_ctx* ctx, const dsph_inter_parms* t, const DivData& dvd, const Cte& c,
                  const float* press) {
  KTimer kt(ctx, "k_forces_fluid");
  const unsigned npf = t->np - t->npb;
  k_forces_fluid<TVISCO, TDDT, VELGRAD><<<grid_for(npf, kBsForces), kBsForces, 0, ctx->stream>>>(
      npf, t->npb, t->viscob, t->viscof, dvd, c, t->d_dcell, (const float4*)t->d_poscell,
      (const float4*)t->d_velrho, press, t->d_spstaurho2, t->d_viscdt, t->d_ar, t->d_ace,
      t->d_sps2strain, t->d_velgrad);
}

template <int TVISCO, bool VELGRAD>
int dispatch_ddt(dsph_ctx* ctx, const dsph_inter_parms* t, const DivData& dvd, const Cte& c,
                 const float* press) {
  switch (t->tdensity) {
    case DSPH_DDT_NONE: launch_fluid<TVISCO, DSPH_DDT_NONE, VELGRAD>(ctx, t, dvd, c, press); break;
    case DSPH_DDT_DDT: launch_fluid<TVISCO, DSPH_DDT_DDT, VELGRAD>(ctx, t, dvd, c, press); break;
    case DSPH_DDT_DDT2: launch_fluid<TVISCO, DSPH_DDT_DDT2, VELGRAD>(ctx, t, dvd, c, press); break;
    case DSPH_DDT_DDT2FULL:
      launch_fluid<TVISCO, DSPH_DDT_DDT2FULL, VELGRAD>(ctx, t, dvd, c, press);
      break;
    default: return set_error(ctx, DSPH_E_ARG, "dsph_interaction_forces: invalid tdensity");
  }
  return DSPH_OK;
}

}  // namespace

extern "C" {

int dsph_interaction_forces(dsph_ctx* ctx, const dsph_inter_parms* t) {
  if (!ctx || !t) return DSPH_E_ARG;
  if (!ctx->cte_set) return set_error(ctx, DSPH_E_STATE, "constants not set");
  DSPH_REQUIRE(ctx, t->tkernel == DSPH_KERNEL_WENDLAND, "only the Wendland kernel is built");
  DSPH_REQUIRE(ctx, t->np >= t->npb && t->npb >= t->npbok, "invalid particle counts");
  DSPH_REQUIRE(ctx, t->d_dcell && t->d_poscell && t->d_velrho && t->d_viscdt && t->d_ar &&
                        t->d_ace && t->divdata.d_beginendcell,
               "dsph_interaction_forces: null array");
  const bool sps = t->tvisco == DSPH_VISCO_LAMINARSPS;
  DSPH_REQUIRE(ctx, !sps || (t->d_spstaurho2 && t->d_sps2strain),
               "Laminar+SPS needs spstaurho2 and sps2strain");
  const Cte c = make_cte(ctx->cte);
  const DivData dvd = make_divdata(t->divdata);
  const unsigned npf = t->np - t->npb;
  if (npf) {
    int rc = ensure_redu_scratch(ctx, 0);
    if (rc) return rc;
    // Pressure of every particle (fluid neighbours read it; boundary too).
    float* press = nullptr;
    rc = ensure_sort_scratch(ctx, t->np);
    if (rc) return rc;
    press = reinterpret_cast<float*>(ctx->sort_keys_alt);  // sort scratch is free here
    {
      KTimer kt(ctx, "k_press");
      k_press<<<grid_for(t->np, kBs), kBs, 0, ctx->stream>>>(t->np, (const float4*)t->d_velrho,
                                                              press, c);
    }
    const bool vg = t->d_velgrad != nullptr;
    switch (t->tvisco) {
      case DSPH_VISCO_ARTIFICIAL:
        rc = vg ? dispatch_ddt<DSPH_VISCO_ARTIFICIAL, true>(ctx, t, dvd, c, press)
                : dispatch_ddt<DSPH_VISCO_ARTIFICIAL, false>(ctx, t, dvd, c, press);
        break;
      case DSPH_VISCO_LAMINAR:
        rc = vg ? dispatch_ddt<DSPH_VISCO_LAMINAR, true>(ctx, t, dvd, c, press)
                : dispatch_ddt<DSPH_VISCO_LAMINAR, false>(ctx, t, dvd, c, press);
        break;
      case DSPH_VISCO_LAMINARSPS:
        rc = vg ? dispatch_ddt<DSPH_VISCO_LAMINARSPS, true>(ctx, t, dvd, c, press)
                : dispatch_ddt<DSPH_VISCO_LAMINARSPS, false>(ctx, t, dvd, c, press);
        break;
      default: return set_error(ctx, DSPH_E_ARG, "dsph_interaction_forces: invalid tvisco");
    }
    if (rc) return rc;
    DSPH_LAUNCHED(ctx, "k_forces_fluid");
  }
  if (t->npbok) {
    KTimer kt(ctx, "k_forces_bound");
    k_forces_bound<<<grid_for(t->npbok, kBsForces), kBsForces, 0, ctx->stream>>>(
        t->npbok, dvd, c, t->d_dcell, (const float4*)t->d_poscell, (const float4*)t->d_velrho,
        t->d_viscdt, t->d_ar);
    DSPH_LAUNCHED(ctx, "k_forces_bound");
  }
  return DSPH_OK;
}

int dsph_sps_tau(dsph_ctx* ctx, uint32_t np, uint32_t npb, float smag, float blin,
                 const float* d_velrho, const float* d_sps2strain, float* d_tau) {
  if (!ctx) return DSPH_E_ARG;
  if (np <= npb) return DSPH_OK;
  DSPH_REQUIRE(ctx, d_velrho && d_sps2strain && d_tau, "dsph_sps_tau: null array");
  const unsigned npf = np - npb;
  k_sps_tau<<<grid_for(npf, kBs), kBs, 0, ctx->stream>>>(npf, npb, smag, blin,
                                                          (const float4*)d_velrho, d_sps2strain,
                                                          d_tau);
  DSPH_LAUNCHED(ctx, "k_sps_tau");
  return DSPH_OK;
}

int dsph_mdbc_correction(dsph_ctx* ctx, uint32_t n, const dsph_divdata* dvdp,
                         const double mapposmin[3], const double* d_posxy, const double* d_posz,
                         const float* d_poscell, const uint16_t* d_code, const float* d_boundnor,
                         float* d_velrho) {
  if (!ctx || !dvdp || !mapposmin) return DSPH_E_ARG;
  if (!ctx->cte_set) return set_error(ctx, DSPH_E_STATE, "constants not set");
  if (!n) return DSPH_OK;
  DSPH_REQUIRE(ctx, d_posxy && d_posz && d_poscell && d_code && d_boundnor && d_velrho,
               "dsph_mdbc_correction: null array");
  const Cte c = make_cte(ctx->cte);
  k_mdbc<<<grid_for(n, kBsForces), kBsForces, 0, ctx->stream>>>(
      n, make_divdata(*dvdp), c, make_double3(mapposmin[0], mapposmin[1], mapposmin[2]),
      (const double2*)d_posxy, d_posz, (const float4*)d_poscell, d_code, d_boundnor,
      (float4*)d_velrho);
  DSPH_LAUNCHED(ctx, "k_mdbc");
  return DSPH_OK;
}

int dsph_pips_count_split_sync(dsph_ctx* ctx, uint32_t nb, uint32_t pinitb, uint32_t nf,
                               uint32_t pinitf, const dsph_divdata* dvdp,
                               const uint32_t* d_dcell, const float* d_poscell,
                               uint64_t h_res[5]) {
  if (!ctx || !dvdp || !h_res) return DSPH_E_ARG;
  unsigned long long* res = reinterpret_cast<unsigned long long*>(ctx->small_dev);
  k_zero_u64<<<1, 64, 0, ctx->stream>>>(res, 5);
  if (nb + nf) {
    DSPH_REQUIRE(ctx, d_dcell && d_poscell, "dsph_pips_count_sync: null array");
    KTimer kt(ctx, "k_pips");
    k_pips<<<grid_for(nb + nf, kBs), kBs, 0, ctx->stream>>>(
        nb, pinitb, nf, pinitf, make_divdata(*dvdp), d_dcell, (const float4*)d_poscell, res);
    DSPH_LAUNCHED(ctx, "k_pips");
  }
  DSPH_HIP(ctx, hipMemcpyAsync(ctx->pinned, res, 5 * sizeof(uint64_t), hipMemcpyDeviceToHost,
                               ctx->stream));
  DSPH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  std::memcpy(h_res, ctx->pinned, 5 * sizeof(uint64_t));
  return DSPH_OK;
}

int dsph_pips_count_sync(dsph_ctx* ctx, uint32_t nb, uint32_t pinitb, uint32_t nf,
                         uint32_t pinitf, const dsph_divdata* dvdp, const uint32_t* d_dcell,
                         const float* d_poscell, uint64_t h_res[4]) {
  if (!h_res) return DSPH_E_ARG;
  uint64_t r[5];
  const int rc = dsph_pips_count_split_sync(ctx, nb, pinitb, nf, pinitf, dvdp, d_dcell,
                                            d_poscell, r);
  if (rc) return rc;
  for (int k = 0; k < 4; k++) h_res[k] = r[k];
  return DSPH_OK;
}

}  // extern "C"
