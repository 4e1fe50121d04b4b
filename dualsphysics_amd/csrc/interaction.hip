// Interaction kernels beside the force interaction (forces.hip): Smagorinsky
// SPS stress, mDBC ghost-node density and the PIPS pair counter (FP32 pair
// arithmetic, reference kernels cited at each kernel).
#include <cfloat>
#include <cstring>
#include "sph_math.h"

using namespace dsph;

namespace {

constexpr unsigned kBsForces = 128;
constexpr unsigned kBs = 256;

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

// 2-D: the y row and column of the moment sums vanish (all dy = 0) and the
// system is the 3x3 one over (1, x, z) of the reference's 2-D branches
// (Determinant3x3dbl by the rule of Sarrus and InverseMatrix3x3dbl by 2x2
// cofactors, FunctionsMath_iker.h:111-114 and 150-165; FP64 from the FP32
// sums); its inverse is returned in the 4x4 positions, the y row and column 0.
__device__ inline double mdbc_det_inv(bool sim2d, const M4d& a, M4d& inv) {
  if (!sim2d) return det4_inv(a, inv);
  const int ix[3] = {0, 1, 3};
  double m[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) m[i][j] = a.m[ix[i]][ix[j]];
  const double det = m[0][0] * m[1][1] * m[2][2] + m[0][1] * m[1][2] * m[2][0] +
                     m[0][2] * m[1][0] * m[2][1] - m[2][0] * m[1][1] * m[0][2] -
                     m[2][1] * m[1][2] * m[0][0] - m[2][2] * m[1][0] * m[0][1];
  double r[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  if (det != 0) {
    r[0][0] = (m[1][1] * m[2][2] - m[1][2] * m[2][1]) / det;
    r[0][1] = -(m[0][1] * m[2][2] - m[0][2] * m[2][1]) / det;
    r[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) / det;
    r[1][0] = -(m[1][0] * m[2][2] - m[1][2] * m[2][0]) / det;
    r[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) / det;
    r[1][2] = -(m[0][0] * m[1][2] - m[0][2] * m[1][0]) / det;
    r[2][0] = (m[1][0] * m[2][1] - m[1][1] * m[2][0]) / det;
    r[2][1] = -(m[0][0] * m[2][1] - m[0][1] * m[2][0]) / det;
    r[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) / det;
  }
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) inv.m[i][j] = 0;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) inv.m[ix[i]][ix[j]] = r[i][j];
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
  const double det = mdbc_det_inv(c.sim2d != 0, Ad, inv);
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

// mDBC2, SlipMode no-slip with the boundary velocities kept (MDBC2_KEEPVEL,
// DualSphDef.h:471); KerInteractionMdbc2Correction_Fast,
// JSphGpu_mdbc_iker.cu:289-487 (3-D). Shepard sums, the 4x4 first-order
// matrix (columns 1..3 with -dr) and the submerged test at the ghost node;
// density from the corrected system when it is invertible and
// dp^2 |A|_inf |A^-1|_inf <= 50 (determinant and inverse in FP64 from the FP32
// sums, inverse rounded to FP32), Shepard otherwise (clamped at rho0 when the
// kernel sum is < 0.1); pressure cloned to the particle with the hydrostatic
// gradient (KerMdbc2PressClone); tangential part of 2 v_motion - v_ghost
// (KerMdbc2TangenVel), or of v_ghost for free slip (tslip =
// DSPH_SLIP_FREESLIP, JSphGpu_mdbc_iker.cu:464-473). A dry ghost node
// switches the particle's mass off.
__device__ inline float3 mdbc2_tangential(float bx, float by, float bz, float3 v) {
  const float sn = sqrtf(bx * bx + by * by + bz * bz);
  const float nx = bx / sn, ny = by / sn, nz = bz / sn;
  const float vdn = v.x * nx + v.y * ny + v.z * nz;
  return make_float3(v.x - vdn * nx, v.y - vdn * ny, v.z - vdn * nz);
}

__global__ void __launch_bounds__(kBsForces) k_mdbc2(unsigned n, int tslip, DivData dvd, Cte c,
    float3 gravity, double3 mapposmin, const double2* __restrict__ posxy,
    const double* __restrict__ posz, const float4* __restrict__ poscell,
    const typecode* __restrict__ code, const float* __restrict__ boundnor,
    const float* __restrict__ motionvel, const float* __restrict__ motionace,
    float4* __restrict__ velrho, uint8_t* __restrict__ boundmode, float* __restrict__ tangenvel) {
  const unsigned p1 = blockIdx.x * blockDim.x + threadIdx.x;
  if (p1 >= n) return;
  const float bnx = boundnor[3 * size_t(p1)], bny = boundnor[3 * size_t(p1) + 1],
              bnz = boundnor[3 * size_t(p1) + 2];
  if (bnx == 0 && bny == 0 && bnz == 0) return;
  const float3 mv = motionvel ? make_float3(motionvel[3 * size_t(p1)], motionvel[3 * size_t(p1) + 1],
                                            motionvel[3 * size_t(p1) + 2])
                              : make_float3(0, 0, 0);
  const float3 ma = motionace ? make_float3(motionace[3 * size_t(p1)], motionace[3 * size_t(p1) + 1],
                                            motionace[3 * size_t(p1) + 2])
                              : make_float3(0, 0, 0);
  const double gx = posxy[p1].x + bnx, gy = posxy[p1].y + bny, gz = posz[p1] + bnz;
  float4 gps;
  {
    const double dx = gx - mapposmin.x, dy = gy - mapposmin.y, dz = gz - mapposmin.z;
    const unsigned cx = unsigned(dx / c.pcs), cy = unsigned(dy / c.pcs), cz = unsigned(dz / c.pcs);
    gps = make_float4(float(dx - double(c.pcs) * cx), float(dy - double(c.pcs) * cy),
                      float(dz - double(c.pcs) * cz), __uint_as_float(pscel_code(cx, cy, cz)));
  }
  float rhop1 = 0, grx = 0, gry = 0, grz = 0, sumwab = 0, submerged = 0;
  float vpx = 0, vpy = 0, vpz = 0;
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
          const float volp2 = c.massf / vr2.w;
          submerged -= volp2 * (dr.x * frx + dr.y * fry + dr.z * frz);
          rhop1 += c.massf * wab;
          grx += c.massf * frx; gry += c.massf * fry; grz += c.massf * frz;
          const float vwab = wab * volp2;
          sumwab += vwab;
          const float vfrx = frx * volp2, vfry = fry * volp2, vfrz = frz * volp2;
          vpx += vwab * vr2.x; vpy += vwab * vr2.y; vpz += vwab * vr2.z;
          A[0][0] += vwab; A[0][1] -= dr.x * vwab; A[0][2] -= dr.y * vwab; A[0][3] -= dr.z * vwab;
          A[1][0] += vfrx; A[1][1] -= dr.x * vfrx; A[1][2] -= dr.y * vfrx; A[1][3] -= dr.z * vfrx;
          A[2][0] += vfry; A[2][1] -= dr.x * vfry; A[2][2] -= dr.y * vfry; A[2][3] -= dr.z * vfry;
          A[3][0] += vfrz; A[3][1] -= dr.x * vfrz; A[3][2] -= dr.y * vfrz; A[3][3] -= dr.z * vfrz;
        }
      }
    }
  float* tv = tangenvel + 3 * size_t(p1);
  if (!(submerged > 0.f)) {
    boundmode[p1] = DSPH_BMODE_MDBC2OFF;
    velrho[p1].w = c.rhopzero;
    const float3 t = mdbc2_tangential(bnx, bny, bnz, mv);
    tv[0] = t.x; tv[1] = t.y; tv[2] = t.z;
    return;
  }
  boundmode[p1] = DSPH_BMODE_MDBC2;
  float rhofinal;
  if (sumwab < 0.1f) {
    rhofinal = fmaxf(c.rhopzero, rhop1 / A[0][0]);
  } else {
    M4d Ad, inv;
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j < 4; j++) Ad.m[i][j] = double(A[i][j]);
    const double det = mdbc_det_inv(c.sim2d != 0, Ad, inv);
    if (fabs(det) >= 0.001) {
      float fi[4][4];
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) fi[i][j] = float(inv.m[i][j]);
      float na = 0, ni = 0;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        na = fmaxf(na, fabsf(A[i][0]) + fabsf(A[i][1]) + fabsf(A[i][2]) + fabsf(A[i][3]));
        ni = fmaxf(ni, fabsf(fi[i][0]) + fabsf(fi[i][1]) + fabsf(fi[i][2]) + fabsf(fi[i][3]));
      }
      const float condinf = c.dp * c.dp * na * ni;
      rhofinal = (condinf <= 50.f)
                     ? fi[0][0] * rhop1 + fi[0][1] * grx + fi[0][2] * gry + fi[0][3] * grz
                     : rhop1 / A[0][0];
    } else {
      rhofinal = rhop1 / A[0][0];
    }
  }
  {
    const float sn = sqrtf(bnx * bnx + bny * bny + bnz * bnz);
    const float nx = bnx / sn, ny = bny / sn, nz = bnz / sn;
    const float normpos = (-bnx) * nx + (-bny) * ny + (-bnz) * nz;
    const float normforce = c.rhopzero * ((gravity.x - ma.x) * nx + (gravity.y - ma.y) * ny +
                                          (gravity.z - ma.z) * nz);
    const float cs2 = c.cs0 * c.cs0;
    const float pghost = cs2 * (rhofinal - c.rhopzero);
    const float pressfinal = pghost + normforce * normpos;
    rhofinal = c.rhopzero + pressfinal / cs2;
  }
  velrho[p1].w = rhofinal;
  const float ia = A[0][0];
  const float3 vg = make_float3(vpx / ia, vpy / ia, vpz / ia);
  const float3 v2 = tslip == DSPH_SLIP_FREESLIP
                        ? vg
                        : make_float3(mv.x + mv.x - vg.x, mv.y + mv.y - vg.y, mv.z + mv.z - vg.z);
  const float3 t = mdbc2_tangential(bnx, bny, bnz, v2);
  tv[0] = t.x; tv[1] = t.y; tv[2] = t.z;
}

// ---------------------------------------------------------------- gauges
// Velocity and mass at measurement points (KerInteractionGaugeVel and
// KerCalculeMass, JDsGauge_ker.cu:50-164): one lane per point, FP64 sums in
// the reference's row order over the fluid particles within 2h of
// wab*V*v (velocity, no Shepard normalisation) and wab*V*m (mass), with
// V = m/rho; distances float(FP64 difference); points outside the map give 0.
__device__ inline float gauge_wab(float rr2, const Cte& c) {
  const float qq = sqrtf(rr2) / c.h;
  const float wqq = qq + qq + 1.f;
  const float wqq1 = 1.f - 0.5f * qq;
  const float wqq2 = wqq1 * wqq1;
  return c.awen * wqq * wqq2 * wqq2;
}

__global__ void __launch_bounds__(kBs) k_gauge_points(unsigned n,
    const double* __restrict__ pts, DivData dvd, Cte c, double3 mapmin, double3 mapmax,
    const double2* __restrict__ posxy, const double* __restrict__ posz,
    const float4* __restrict__ velrho, const typecode* __restrict__ code,
    float* __restrict__ outvel, float* __restrict__ outmass) {
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double px = pts[3 * size_t(i)], py = pts[3 * size_t(i) + 1], pz = pts[3 * size_t(i) + 2];
  double sx = 0, sy = 0, sz = 0, sm = 0;
  const bool out = !(px >= mapmin.x && py >= mapmin.y && pz >= mapmin.z && px < mapmax.x &&
                     py < mapmax.y && pz < mapmax.z);
  if (!out) {
    const int cx = int((px - dvd.domposminx) / dvd.scell) - dvd.czx;
    const int cy = int((py - dvd.domposminy) / dvd.scell) - dvd.czy;
    const int cz = int((pz - dvd.domposminz) / dvd.scell) - dvd.czz;
    const NgLimits l = ng_limits(cx, cy, cz, dvd);
    for (int z = l.z0; z < l.z1; z++)
      for (int y = l.y0; y < l.y1; y++) {
        unsigned pini, pfin;
        row_range(dvd.begincell, int(dvd.cellfluid) + z * dvd.ncw + y * dvd.ncx, l.x0, l.x1,
                  pini, pfin);
        for (unsigned p2 = pini; p2 < pfin; p2++) {
          const double2 xy = posxy[p2];
          const float drx = float(px - xy.x), dry = float(py - xy.y), drz = float(pz - posz[p2]);
          const float rr2 = drx * drx + dry * dry + drz * drz;
          if (rr2 <= c.ks2 && rr2 >= kAlmostZero && code_is_fluid(code[p2])) {
            const float4 v = velrho[p2];
            const float wab = gauge_wab(rr2, c) * (c.massf / v.w);
            sx += wab * v.x; sy += wab * v.y; sz += wab * v.z;
            sm += wab * c.massf;
          }
        }
      }
  }
  if (outvel) {
    outvel[3 * size_t(i)] = float(sx);
    outvel[3 * size_t(i) + 1] = float(sy);
    outvel[3 * size_t(i) + 2] = float(sz);
  }
  if (outmass) outmass[i] = float(sm);
}

// Highest fluid particle near a vertical line (KerInteractionGaugeMaxz,
// JDsGauge_ker.cu:252-312; cell range JGaugeMaxZ::GetInteractionCellsMaxZ,
// JDsGaugeItem.cpp:1023-1040): one lane per gauge, cell layers from the top
// of [z0, z0+height] (widened by the distance limit) down to the first layer
// with a fluid particle within distlimit horizontally; result
// max(zmax, zfloor).
__global__ void __launch_bounds__(kBs) k_gauge_maxz(unsigned n, const double* __restrict__ pts,
    const float* __restrict__ hd, DivData dvd, double zfloor,
    const double2* __restrict__ posxy, const double* __restrict__ posz,
    const typecode* __restrict__ code, float* __restrict__ out) {
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double px = pts[3 * size_t(i)], py = pts[3 * size_t(i) + 1], pz = pts[3 * size_t(i) + 2];
  const float height = hd[2 * size_t(i)], dist = hd[2 * size_t(i) + 1];
  const float maxdist2 = dist * dist;
  const int cx = int((px - dvd.domposminx) / dvd.scell) - dvd.czx;
  const int cy = int((py - dvd.domposminy) / dvd.scell) - dvd.czy;
  const int cz = int((pz - dvd.domposminz) / dvd.scell) - dvd.czz;
  const int cz2 = int((pz + height - dvd.domposminz) / dvd.scell) - dvd.czz;
  const int ncel = int(ceilf(dist / dvd.scell));
  const int x0 = max(cx - ncel, 0), x1 = min(cx + ncel + 1, dvd.ncx);
  const int y0 = max(cy - ncel, 0), y1 = min(cy + ncel + 1, dvd.ncy);
  const int z0 = max(cz - ncel, 0), z1 = min(cz2 + ncel + 1, dvd.ncz);
  float zmax = -FLT_MAX;
  bool found = false;
  if (x0 < x1)
    for (int z = z1 - 1; z >= z0 && !found; z--)
      for (int y = y0; y < y1; y++) {
        unsigned pini, pfin;
        row_range(dvd.begincell, int(dvd.cellfluid) + z * dvd.ncw + y * dvd.ncx, x0, x1, pini,
                  pfin);
        for (unsigned p2 = pini; p2 < pfin; p2++) {
          const double z2 = posz[p2];
          if (z2 > zmax) {
            const double2 xy = posxy[p2];
            const float drx = float(px - xy.x), dry = float(py - xy.y);
            if (drx * drx + dry * dry <= maxdist2 && code_is_fluid(code[p2])) {
              zmax = float(z2);
              found = true;
            }
          }
        }
      }
  out[i] = fmaxf(zmax, float(zfloor));
}

// Pressure force of the fluid on boundary particles with id in
// [idbegin, idbegin+count) (KerInteractionGaugeForce, JDsGauge_ker.cu:497-545):
// per particle sum over fluid neighbours of -(P1+P2)/(rho1 rho2) m2 gradW,
// P = B((rho/rho0)^gamma - 1); ace[id-idbegin] (the caller sums and scales by
// the boundary mass).
__global__ void __launch_bounds__(kBs) k_gauge_force(unsigned npb, unsigned idbegin,
    unsigned count, DivData dvd, Cte c, const double2* __restrict__ posxy,
    const double* __restrict__ posz, const typecode* __restrict__ code,
    const unsigned* __restrict__ idp, const float4* __restrict__ velrho, float* __restrict__ ace) {
  const unsigned p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npb) return;
  const unsigned id = idp[p];
  if (id - idbegin >= count || !code_is_normal(code[p])) return;
  const double2 xy = posxy[p];
  const double px = xy.x, py = xy.y, pz = posz[p];
  const float rho1 = velrho[p].w;
  const float press1 = c.cteb * (powf(rho1 / c.rhopzero, c.gamma) - 1.f);
  float ax = 0, ay = 0, az = 0;
  const int cx = int((px - dvd.domposminx) / dvd.scell) - dvd.czx;
  const int cy = int((py - dvd.domposminy) / dvd.scell) - dvd.czy;
  const int cz = int((pz - dvd.domposminz) / dvd.scell) - dvd.czz;
  const NgLimits l = ng_limits(cx, cy, cz, dvd);
  for (int z = l.z0; z < l.z1; z++)
    for (int y = l.y0; y < l.y1; y++) {
      unsigned pini, pfin;
      row_range(dvd.begincell, int(dvd.cellfluid) + z * dvd.ncw + y * dvd.ncx, l.x0, l.x1, pini,
                pfin);
      for (unsigned p2 = pini; p2 < pfin; p2++) {
        const double2 xy2 = posxy[p2];
        const float drx = float(px - xy2.x), dry = float(py - xy2.y), drz = float(pz - posz[p2]);
        const float rr2 = drx * drx + dry * dry + drz * drz;
        if (rr2 <= c.ks2 && rr2 >= kAlmostZero && code_is_fluid(code[p2])) {
          const float qq = sqrtf(rr2) / c.h;
          const float w1 = 1.f - 0.5f * qq;
          const float fac = c.bwenh * w1 * w1 * w1;
          const float rho2 = velrho[p2].w;
          const float press2 = c.cteb * (powf(rho2 / c.rhopzero, c.gamma) - 1.f);
          const float prs = (press1 + press2) / (rho1 * rho2);
          const float p_vpm1 = -prs * c.massf;
          ax += p_vpm1 * fac * drx; ay += p_vpm1 * fac * dry; az += p_vpm1 * fac * drz;
        }
      }
    }
  float* a = ace + 3 * size_t(id - idbegin);
  a[0] = ax; a[1] = ay; a[2] = az;
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

}  // namespace

extern "C" {

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

int dsph_mdbc2_correction(dsph_ctx* ctx, uint32_t n, const float gravity[3],
                          const dsph_divdata* dvdp, const double mapposmin[3],
                          const double* d_posxy, const double* d_posz, const float* d_poscell,
                          const uint16_t* d_code, const float* d_boundnor,
                          const float* d_motionvel, const float* d_motionace, float* d_velrho,
                          uint8_t* d_boundmode, float* d_tangenvel) {
  return dsph_mdbc2_correction_slip(ctx, DSPH_SLIP_NOSLIP, n, gravity, dvdp, mapposmin, d_posxy,
                                    d_posz, d_poscell, d_code, d_boundnor, d_motionvel,
                                    d_motionace, d_velrho, d_boundmode, d_tangenvel);
}

int dsph_mdbc2_correction_slip(dsph_ctx* ctx, int32_t slipmode, uint32_t n,
                               const float gravity[3], const dsph_divdata* dvdp,
                               const double mapposmin[3], const double* d_posxy,
                               const double* d_posz, const float* d_poscell,
                               const uint16_t* d_code, const float* d_boundnor,
                               const float* d_motionvel, const float* d_motionace,
                               float* d_velrho, uint8_t* d_boundmode, float* d_tangenvel) {
  if (!ctx || !dvdp || !mapposmin || !gravity) return DSPH_E_ARG;
  if (slipmode != DSPH_SLIP_NOSLIP && slipmode != DSPH_SLIP_FREESLIP)
    return set_error(ctx, DSPH_E_ARG, "dsph_mdbc2_correction: slipmode must be 2 or 3");
  if (!ctx->cte_set) return set_error(ctx, DSPH_E_STATE, "constants not set");
  if (!n) return DSPH_OK;
  DSPH_REQUIRE(ctx, d_posxy && d_posz && d_poscell && d_code && d_boundnor && d_velrho &&
                        d_boundmode && d_tangenvel,
               "dsph_mdbc2_correction: null array");
  const Cte c = make_cte(ctx->cte);
  k_mdbc2<<<grid_for(n, kBsForces), kBsForces, 0, ctx->stream>>>(
      n, slipmode, make_divdata(*dvdp), c, make_float3(gravity[0], gravity[1], gravity[2]),
      make_double3(mapposmin[0], mapposmin[1], mapposmin[2]), (const double2*)d_posxy, d_posz,
      (const float4*)d_poscell, d_code, d_boundnor, d_motionvel, d_motionace, (float4*)d_velrho,
      d_boundmode, d_tangenvel);
  DSPH_LAUNCHED(ctx, "k_mdbc2");
  return DSPH_OK;
}

int dsph_gauge_points(dsph_ctx* ctx, uint32_t n, const double* d_points,
                      const dsph_divdata* dvdp, const double mapmin[3], const double mapmax[3],
                      const double* d_posxy, const double* d_posz, const float* d_velrho,
                      const uint16_t* d_code, float* d_vel, float* d_mass) {
  if (!ctx || !dvdp || !mapmin || !mapmax) return DSPH_E_ARG;
  if (!ctx->cte_set) return set_error(ctx, DSPH_E_STATE, "constants not set");
  if (!n) return DSPH_OK;
  DSPH_REQUIRE(ctx, d_points && d_posxy && d_posz && d_velrho && d_code && (d_vel || d_mass),
               "dsph_gauge_points: null array");
  k_gauge_points<<<grid_for(n, kBs), kBs, 0, ctx->stream>>>(
      n, d_points, make_divdata(*dvdp), make_cte(ctx->cte),
      make_double3(mapmin[0], mapmin[1], mapmin[2]), make_double3(mapmax[0], mapmax[1], mapmax[2]),
      (const double2*)d_posxy, d_posz, (const float4*)d_velrho, d_code, d_vel, d_mass);
  DSPH_LAUNCHED(ctx, "k_gauge_points");
  return DSPH_OK;
}

int dsph_gauge_maxz(dsph_ctx* ctx, uint32_t n, const double* d_points,
                    const float* d_height_dist, const dsph_divdata* dvdp, double zfloor,
                    const double* d_posxy, const double* d_posz, const uint16_t* d_code,
                    float* d_zmax) {
  if (!ctx || !dvdp) return DSPH_E_ARG;
  if (!n) return DSPH_OK;
  DSPH_REQUIRE(ctx, d_points && d_height_dist && d_posxy && d_posz && d_code && d_zmax,
               "dsph_gauge_maxz: null array");
  k_gauge_maxz<<<grid_for(n, kBs), kBs, 0, ctx->stream>>>(
      n, d_points, d_height_dist, make_divdata(*dvdp), zfloor, (const double2*)d_posxy, d_posz,
      d_code, d_zmax);
  DSPH_LAUNCHED(ctx, "k_gauge_maxz");
  return DSPH_OK;
}

int dsph_gauge_force(dsph_ctx* ctx, uint32_t npb, uint32_t idbegin, uint32_t count,
                     const dsph_divdata* dvdp, const double* d_posxy, const double* d_posz,
                     const uint16_t* d_code, const uint32_t* d_idp, const float* d_velrho,
                     float* d_ace) {
  if (!ctx || !dvdp) return DSPH_E_ARG;
  if (!ctx->cte_set) return set_error(ctx, DSPH_E_STATE, "constants not set");
  if (!npb || !count) return DSPH_OK;
  DSPH_REQUIRE(ctx, d_posxy && d_posz && d_code && d_idp && d_velrho && d_ace,
               "dsph_gauge_force: null array");
  DSPH_HIP(ctx, hipMemsetAsync(d_ace, 0, 12 * size_t(count), ctx->stream));
  k_gauge_force<<<grid_for(npb, kBs), kBs, 0, ctx->stream>>>(
      npb, idbegin, count, make_divdata(*dvdp), make_cte(ctx->cte), (const double2*)d_posxy,
      d_posz, d_code, d_idp, (const float4*)d_velrho, d_ace);
  DSPH_LAUNCHED(ctx, "k_gauge_force");
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
