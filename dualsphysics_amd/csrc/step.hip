// Reductions, variable dt and time integration (Symplectic, Verlet, FP64
// position update) of the SPH step on gfx950.
//
// References: ReduMaxFloat / ComputeVelMod / ComputeAceMod (JSphGpu_ker.cu:
// 55-312), DtVariable (JSphGpu.cpp:1103-1147), KerComputeStepSymplecticPre/Cor
// and KerComputeStepVerlet (JSphGpuSimple_ker.cu:132-513), KerUpdatePos /
// KerComputeStepPos(2) (JSphGpu_ker.cu:334-392, 1476-1570). Options out of
// scope (floatings, shifting, inlet/outlet, mDBC2, periodic) are not built.
#include <algorithm>
#include <cfloat>
#include <climits>
#include "dsph_ctx.h"
#include "ensemble.h"

using namespace dsph;

namespace {

constexpr unsigned kBs = 256;

// ---------------------------------------------------------------- max reduce
// Mode 0: plain values; 1: |float4.xyz|^2; 2: |float3|^2.
template <int MODE>
__device__ inline float redu_load(const float* __restrict__ d, unsigned i) {
  if (MODE == 0) return d[i];
  if (MODE == 1) {
    const float4 v = reinterpret_cast<const float4*>(d)[i];
    return v.x * v.x + v.y * v.y + v.z * v.z;
  }
  const float x = d[3 * size_t(i)], y = d[3 * size_t(i) + 1], z = d[3 * size_t(i) + 2];
  return x * x + y * y + z * z;
}

__device__ inline float block_max(float v) {
  __shared__ float sw[kBs / 64];
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off));
  if ((threadIdx.x & 63) == 0) sw[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0)
    for (unsigned i = 1; i < blockDim.x / 64; i++) v = fmaxf(v, sw[i]);
  return v;
}

// Each block reduces 8 x 256 values into partial[blockIdx.x].
template <int MODE>
__global__ void __launch_bounds__(kBs) k_max_partial(unsigned n, unsigned ini,
    const float* __restrict__ data, float* __restrict__ partial) {
  float v = -FLT_MAX;
  const unsigned base = blockIdx.x * kBs * 8 + threadIdx.x;
#pragma unroll
  for (unsigned k = 0; k < 8; k++) {
    const unsigned i = base + k * kBs;
    if (i < n) v = fmaxf(v, redu_load<MODE>(data, i + ini));
  }
  v = block_max(v);
  if (threadIdx.x == 0) partial[blockIdx.x] = v;
}

__global__ void __launch_bounds__(kBs) k_max_final(unsigned nparts, const float* __restrict__ partial,
                                                   float* __restrict__ out) {
  float v = -FLT_MAX;
  for (unsigned i = threadIdx.x; i < nparts; i += kBs) v = fmaxf(v, partial[i]);
  v = block_max(v);
  if (threadIdx.x == 0) *out = (nparts ? v : 0.f);
}

template <int MODE>
int redu_max(dsph_ctx* ctx, uint32_t n, uint32_t ini, const float* d, float* d_out) {
  if (!n) {
    k_max_final<<<1, kBs, 0, ctx->stream>>>(0u, ctx->redu, d_out);
    DSPH_LAUNCHED(ctx, "k_max_final");
    return DSPH_OK;
  }
  KTimer kt(ctx, "k_redu_max");
  const unsigned nparts = (n + kBs * 8 - 1) / (kBs * 8);
  if (nparts + 1 > ctx->redu_cap)
    return set_error(ctx, DSPH_E_NOMEM, "reduction scratch too small (dsph_ctx_reserve)");
  k_max_partial<MODE><<<nparts, kBs, 0, ctx->stream>>>(n, ini, d, ctx->redu);
  k_max_final<<<1, kBs, 0, ctx->stream>>>(nparts, ctx->redu, d_out);
  DSPH_LAUNCHED(ctx, "k_max");
  return DSPH_OK;
}

// ---------------------------------------------------------------- dt
// DtVariable (JSphGpu.cpp:1103-1147) on one lane, FP64. red = {max|v|^2,
// ViscDtMax, max|ace|^2}.
// Returns dt; sets bit 0 of *fl for NaN/inf and bit 1 when clamped to dtmin.
__device__ inline double dt_of(float red0, float red1, float red2, double kernelh, double cs0,
                               double cfl, double dtmin, unsigned& fl) {
  const double velmax = double(sqrtf(red0));
  const float viscdtmax = red1;
  const double acemax = sqrt(double(red2));
  const double dt1 = (acemax ? sqrt(kernelh / acemax) : DBL_MAX);
  const double dt2 = kernelh / (fmax(cs0, velmax * 10.) + kernelh * double(viscdtmax));
  double dt = cfl * fmin(dt1, dt2);
  fl = 0;
  if (dt != dt || isinf(dt)) fl |= 1u;
  if (dt < dtmin) {
    dt = dtmin;
    fl |= 2u;
  }
  return dt;
}

__global__ void k_dt_variable(const float* __restrict__ red, double kernelh, double cs0,
                              double cfl, double dtmin, double* __restrict__ dtout,
                              unsigned slot, unsigned* __restrict__ flags) {
  if (threadIdx.x != 0) return;
  unsigned fl;
  const double dt = dt_of(red[0], red[1], red[2], kernelh, cs0, cfl, dtmin, fl);
  if (fl && flags) atomicOr(flags, fl);
  dtout[slot] = dt;
}

__global__ void k_dt_advance(double* __restrict__ d) {
  if (threadIdx.x != 0) return;
  const double dt = d[2];
  d[3] += dt;
  d[4] = dt;
  d[2] = fmin(d[0], d[1]);
}

// ---------------------------------------------------------------- Symplectic
__device__ inline double dev_dt(double dtm_host, const double* d_dt, unsigned slot, double scale) {
  return d_dt ? d_dt[slot] * scale : dtm_host;
}

__device__ inline void sym_pre_one(unsigned p, bool bound, double dtm,
    const float4* __restrict__ velrhopre, const float* __restrict__ ar,
    const float* __restrict__ ace, float rhopzero, float rhopoutmin, float rhopoutmax,
    float3 gravity, typecode* __restrict__ code, double2* __restrict__ movxy,
    double* __restrict__ movz, float4* __restrict__ velrho) {
  if (bound) {
    float4 vr = velrhopre[p];
    vr.w = float(double(vr.w) + dtm * ar[p]);
    vr.w = (vr.w < rhopzero ? rhopzero : vr.w);
    velrho[p] = vr;
    return;
  }
  const unsigned rcode = code[p];
  const float4 pre = velrhopre[p];
  float4 nw = pre;
  nw.w = float(double(pre.w) + dtm * ar[p]);
  if (code_type(rcode) == kCodeTypeFloating) {
    // Floating particle: its velocity and position stay (RunFloating moves
    // it); density clamped at rho0 (JSphGpuSimple_ker.cu:341-344).
    nw.w = (nw.w < rhopzero ? rhopzero : nw.w);
    movxy[p] = make_double2(0.0, 0.0);
    movz[p] = 0.0;
    velrho[p] = nw;
    return;
  }
  const double dx = double(pre.x) * dtm, dy = double(pre.y) * dtm, dz = double(pre.z) * dtm;
  const bool outrho = (nw.w < rhopoutmin || nw.w > rhopoutmax);
  const float ax = ace[3 * size_t(p)], ay = ace[3 * size_t(p) + 1], az = ace[3 * size_t(p) + 2];
  nw.x = float(double(pre.x) + (double(ax) + gravity.x) * dtm);
  nw.y = float(double(pre.y) + (double(ay) + gravity.y) * dtm);
  nw.z = float(double(pre.z) + (double(az) + gravity.z) * dtm);
  movxy[p] = make_double2(dx, dy);
  movz[p] = dz;
  if (outrho && code_is_normal(rcode)) code[p] = typecode(code_set(rcode, kCodeOutRho));
  velrho[p] = nw;
}

// mDBC2: a boundary particle with boundmode >= 1 keeps the state the step
// starts from (JSphGpuSimple_ker.cu:142, 276, 406-412).
__device__ inline bool mdbc2_keep(unsigned p, unsigned npb, const uint8_t* __restrict__ boundmode) {
  return boundmode && p < npb && boundmode[p] >= DSPH_BMODE_MDBC2;
}

// Shifting displacement added to a fluid particle's displacement
// (JSphGpuSimple_ker.cu:164-168, 303-308, 460-465).
__device__ inline void add_shift(unsigned p, const float4* __restrict__ shiftposfs,
                                 double2* __restrict__ movxy, double* __restrict__ movz) {
  const float4 s = shiftposfs[p];
  const double2 m = movxy[p];
  movxy[p] = make_double2(m.x + double(s.x), m.y + double(s.y));
  movz[p] = movz[p] + double(s.z);
}

__global__ void __launch_bounds__(kBs) k_sym_pre(unsigned n, unsigned npb,
    const float4* __restrict__ velrhopre, const uint8_t* __restrict__ boundmode,
    const float4* __restrict__ shiftposfs, const float* __restrict__ ar,
    const float* __restrict__ ace, double dtm_h, const double* __restrict__ d_dt, unsigned slot,
    float rhopzero, float rhopoutmin, float rhopoutmax, float3 gravity,
    typecode* __restrict__ code, double2* __restrict__ movxy, double* __restrict__ movz,
    float4* __restrict__ velrho) {
  const unsigned p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  if (mdbc2_keep(p, npb, boundmode)) {
    velrho[p] = velrhopre[p];
    return;
  }
  sym_pre_one(p, p < npb, dev_dt(dtm_h, d_dt, slot, 0.5), velrhopre, ar, ace, rhopzero,
              rhopoutmin, rhopoutmax, gravity, code, movxy, movz, velrho);
  if (shiftposfs && p >= npb) add_shift(p, shiftposfs, movxy, movz);
}

__device__ inline void sym_cor_one(unsigned p, bool bound, double dt, double dtm,
    const float4* __restrict__ velrhopre, const float* __restrict__ ar,
    const float* __restrict__ ace, float rhopzero, float rhopoutmin, float rhopoutmax,
    float3 gravity, typecode* __restrict__ code, double2* __restrict__ movxy,
    double* __restrict__ movz, float4* __restrict__ velrho) {
  if (bound) {
    const double eps = (-double(ar[p]) / double(velrho[p].w)) * dt;
    float rrho = float(double(velrhopre[p].w) * (2. - eps) / (2. + eps));
    rrho = (rrho < rhopzero ? rhopzero : rrho);
    velrho[p] = make_float4(0, 0, 0, rrho);
    return;
  }
  const unsigned rcode = code[p];
  const double eps = (-double(ar[p]) / double(velrho[p].w)) * dt;
  const float4 pre = velrhopre[p];
  float4 nw = pre;
  nw.w = float(double(pre.w) * (2. - eps) / (2. + eps));
  if (code_type(rcode) == kCodeTypeFloating) {  // JSphGpuSimple_ker.cu:501-504
    nw.w = (nw.w < rhopzero ? rhopzero : nw.w);
    movxy[p] = make_double2(0.0, 0.0);
    movz[p] = 0.0;
    velrho[p] = nw;
    return;
  }
  const float ax = ace[3 * size_t(p)], ay = ace[3 * size_t(p) + 1], az = ace[3 * size_t(p) + 2];
  nw.x = float(double(pre.x) + (double(ax) + gravity.x) * dt);
  nw.y = float(double(pre.y) + (double(ay) + gravity.y) * dt);
  nw.z = float(double(pre.z) + (double(az) + gravity.z) * dt);
  const double dx = (double(pre.x) + double(nw.x)) * dtm;
  const double dy = (double(pre.y) + double(nw.y)) * dtm;
  const double dz = (double(pre.z) + double(nw.z)) * dtm;
  const bool outrho = (nw.w < rhopoutmin || nw.w > rhopoutmax);
  movxy[p] = make_double2(dx, dy);
  movz[p] = dz;
  if (outrho && code_is_normal(rcode)) code[p] = typecode(code_set(rcode, kCodeOutRho));
  velrho[p] = nw;
}

__global__ void __launch_bounds__(kBs) k_sym_cor(unsigned n, unsigned npb,
    const float4* __restrict__ velrhopre, const uint8_t* __restrict__ boundmode,
    const float4* __restrict__ shiftposfs, const float* __restrict__ ar,
    const float* __restrict__ ace, double dtm_h, double dt_h, const double* __restrict__ d_dt,
    unsigned slot, float rhopzero, float rhopoutmin, float rhopoutmax, float3 gravity,
    typecode* __restrict__ code, double2* __restrict__ movxy, double* __restrict__ movz,
    float4* __restrict__ velrho) {
  const unsigned p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  if (mdbc2_keep(p, npb, boundmode)) {
    velrho[p] = velrhopre[p];
    return;
  }
  sym_cor_one(p, p < npb, dev_dt(dt_h, d_dt, slot, 1.0), dev_dt(dtm_h, d_dt, slot, 0.5),
              velrhopre, ar, ace, rhopzero, rhopoutmin, rhopoutmax, gravity, code, movxy, movz,
              velrho);
  if (shiftposfs && p >= npb) add_shift(p, shiftposfs, movxy, movz);
}

__global__ void __launch_bounds__(kBs) k_verlet(unsigned n, unsigned npb,
    const float4* __restrict__ velrho1, const float4* __restrict__ velrho2,
    const uint8_t* __restrict__ boundmode, const float4* __restrict__ shiftposfs,
    const float* __restrict__ ar, const float* __restrict__ ace, double dt, double dt205,
    double dt2, float rhopzero, float rhopoutmin, float rhopoutmax, float3 gravity,
    typecode* __restrict__ code, double2* __restrict__ movxy, double* __restrict__ movz,
    float4* __restrict__ velrhonew) {
  const unsigned p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  if (mdbc2_keep(p, npb, boundmode)) {
    velrhonew[p] = velrho2[p];
    return;
  }
  if (p < npb) {
    float rrho = float(double(velrho2[p].w) + dt2 * ar[p]);
    rrho = (rrho < rhopzero ? rhopzero : rrho);
    velrhonew[p] = make_float4(0, 0, 0, rrho);
    return;
  }
  const unsigned rcode = code[p];
  const float4 v2 = velrho2[p];
  const float rhonew = float(double(v2.w) + dt2 * ar[p]);
  const float4 v1 = velrho1[p];
  if (code_type(rcode) == kCodeTypeFloating) {
    // Floating particle (JSphGpuSimple_ker.cu:217-221): the current velocity
    // and the new density clamped at rho0; RunFloating moves it.
    velrhonew[p] = make_float4(v1.x, v1.y, v1.z, rhonew < rhopzero ? rhopzero : rhonew);
    return;
  }
  const double ax = double(ace[3 * size_t(p)]) + gravity.x;
  const double ay = double(ace[3 * size_t(p) + 1]) + gravity.y;
  const double az = double(ace[3 * size_t(p) + 2]) + gravity.z;
  double dx = double(v1.x) * dt + ax * dt205;
  double dy = double(v1.y) * dt + ay * dt205;
  double dz = double(v1.z) * dt + az * dt205;
  if (shiftposfs) {
    const float4 s = shiftposfs[p];
    dx += double(s.x); dy += double(s.y); dz += double(s.z);
  }
  const bool outrho = (rhonew < rhopoutmin || rhonew > rhopoutmax);
  const float4 nw = make_float4(float(double(v2.x) + ax * dt2), float(double(v2.y) + ay * dt2),
                                float(double(v2.z) + az * dt2), rhonew);
  movxy[p] = make_double2(dx, dy);
  movz[p] = dz;
  if (outrho && code_is_normal(rcode)) code[p] = typecode(code_set(rcode, kCodeOutRho));
  velrhonew[p] = nw;
}

// mDBC2 no penetration (JSphGpuSimple_ker.cu:177-191, 447-461): after the
// Symplectic corrector or the Verlet step, each axis of a flagged fluid
// particle (w > 5) with a non-zero correction takes velocity
// base + correction and displacement velocity * dt; other axes keep the
// step's values.
__global__ void __launch_bounds__(kBs) k_nopen_step(unsigned np, unsigned npb,
    const float4* __restrict__ nopenshift, const float4* __restrict__ velbase, double dt_h,
    const double* __restrict__ d_dt, unsigned slot, float4* __restrict__ velrho,
    double2* __restrict__ movxy, double* __restrict__ movz) {
  const unsigned p = npb + blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const float4 ns = nopenshift[p];
  if (!(ns.w > 5.f)) return;
  const double dt = dev_dt(dt_h, d_dt, slot, 1.0);
  const float4 vb = velbase[p];
  float4 v = velrho[p];
  double2 m = movxy[p];
  if (ns.x != 0) { v.x = vb.x + ns.x; m.x = double(v.x) * dt; }
  if (ns.y != 0) { v.y = vb.y + ns.y; m.y = double(v.y) * dt; }
  if (ns.z != 0) { v.z = vb.z + ns.z; movz[p] = double(v.z) * dt; }
  velrho[p] = v;
  movxy[p] = m;
}

// mDBC2 with moving boundaries (KerCopyMotionVelAce,
// JSphGpu_mdbc_iker.cu:587-603): each moving particle (boundary p with
// idp - idbegin < count) gets motion acceleration float((v - motionvel)/dt)
// in FP64 and motion velocity v, after the step's motion.
__global__ void __launch_bounds__(kBs) k_motion_vel_ace(unsigned npb, unsigned idbegin,
    unsigned count, double dt_h, const double* __restrict__ d_dt, unsigned slot,
    const unsigned* __restrict__ idp, const float4* __restrict__ velrho,
    float* __restrict__ motionvel, float* __restrict__ motionace) {
  const unsigned p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npb || idp[p] - idbegin >= count) return;
  const double dt = dev_dt(dt_h, d_dt, slot, 1.0);
  const float4 v = velrho[p];
  float* mv = motionvel + 3 * size_t(p);
  float* ma = motionace + 3 * size_t(p);
  ma[0] = float((double(v.x) - mv[0]) / dt);
  ma[1] = float((double(v.y) - mv[1]) / dt);
  ma[2] = float((double(v.z) - mv[2]) / dt);
  mv[0] = v.x; mv[1] = v.y; mv[2] = v.z;
}

// ---------------------------------------------------------------- positions

// KerUpdatePos (JSphGpu_ker.cu:334-392). With periodic axes a particle
// that leaves the real map through a periodic face re-enters by the axis'
// increment, and its cell is measured from DomPosMin.
__device__ inline void update_pos(double2 rxy, double rz, double mx, double my, double mz,
                                  bool outrho, unsigned p, const PosCte& c,
                                  double2* __restrict__ posxy, double* __restrict__ posz,
                                  unsigned* __restrict__ dcell, typecode* __restrict__ code) {
  const bool outmov =
      (fmaxf(fabsf(float(mx)), fmaxf(fabsf(float(my)), fabsf(float(mz)))) > c.movlimit);
  double x = rxy.x + mx, y = rxy.y + my, z = rz + mz;
  double dx = x - c.mapx, dy = y - c.mapy, dz = z - c.mapz;
  bool out = (dx != dx || dy != dy || dz != dz || dx < 0 || dy < 0 || dz < 0 ||
              dx >= c.sizex || dy >= c.sizey || dz >= c.sizez);
  if (c.periactive && out) {
    const bool px = (c.periactive & 1), py = (c.periactive & 2), pz = (c.periactive & 4);
    if (px) {
      if (dx < 0) { dx -= c.perinc[0][0]; dy -= c.perinc[0][1]; dz -= c.perinc[0][2]; }
      if (dx >= c.sizex) { dx += c.perinc[0][0]; dy += c.perinc[0][1]; dz += c.perinc[0][2]; }
    }
    if (py) {
      if (dy < 0) { dx -= c.perinc[1][0]; dy -= c.perinc[1][1]; dz -= c.perinc[1][2]; }
      if (dy >= c.sizey) { dx += c.perinc[1][0]; dy += c.perinc[1][1]; dz += c.perinc[1][2]; }
    }
    if (pz) {
      if (dz < 0) { dx -= c.perinc[2][0]; dy -= c.perinc[2][1]; dz -= c.perinc[2][2]; }
      if (dz >= c.sizez) { dx += c.perinc[2][0]; dy += c.perinc[2][1]; dz += c.perinc[2][2]; }
    }
    out = (!px && (dx < 0 || dx >= c.sizex)) || (!py && (dy < 0 || dy >= c.sizey)) ||
          (!pz && (dz < 0 || dz >= c.sizez));
    x = dx + c.mapx; y = dy + c.mapy; z = dz + c.mapz;
  }
  posxy[p] = make_double2(x, y);
  posz[p] = z;
  if (c.periactive) { dx = x - c.domx; dy = y - c.domy; dz = z - c.domz; }
  if (outrho || outmov || out) {
    unsigned rc = code[p];
    if (out) rc = code_set(rc, kCodeOutPos);
    else if (outrho) rc = code_set(rc, kCodeOutRho);
    else rc = code_set(rc, kCodeOutMov);
    code[p] = typecode(rc);
    dcell[p] = kDcelMapOut;
  } else {
    const unsigned cx = unsigned(dx / c.scell), cy = unsigned(dy / c.scell),
                   cz = unsigned(dz / c.scell);
    dcell[p] = dcel_cell(c.cellcode, cx, cy, cz);
  }
}

// ---------------------------------------------------------------- moving boundaries
// Displacement of a rectilinear sinusoid over [t, t+dt] (JMotionObj.cpp:558-575).
__host__ __device__ inline void rect_sinu_mov(const dsph_motion& m, double t, double dt,
                                              double mv[3]) {
  const double twopi = 6.283185307179586476925286766559;
  for (int k = 0; k < 3; k++) {
    mv[k] = 0;
    if (m.ampl[k] != 0)
      mv[k] = m.ampl[k] * sin(m.phase[k] + twopi * m.freq[k] * (t + dt)) -
              m.ampl[k] * sin(m.phase[k] + twopi * m.freq[k] * t);
  }
}

// KerMoveLinBound (JSphGpu_ker.cu:1671-1685): boundary particles of the moving
// object (selected by id) are displaced with KerUpdatePos and take the
// imposed velocity. With a device dt slot array the displacement of the step
// is computed here from the motion law.
__global__ void __launch_bounds__(kBs) k_move_lin_bound(unsigned npb, unsigned idbegin,
    unsigned count, PosCte c, double3 mvh, float3 velh, dsph_motion m, double t0,
    const double* __restrict__ d_dt, const unsigned* __restrict__ idp,
    double2* __restrict__ posxy, double* __restrict__ posz, unsigned* __restrict__ dcell,
    float4* __restrict__ velrho, typecode* __restrict__ code) {
  const unsigned p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npb) return;
  if (idp[p] - idbegin >= count) return;
  double3 mv = mvh;
  float3 vel = velh;
  if (d_dt) {
    const double dt = d_dt[2], t = t0 + d_dt[3];
    if (!(t >= m.start && t < m.finish) || dt <= 0) return;
    double d[3];
    rect_sinu_mov(m, t, dt, d);
    mv = make_double3(d[0], d[1], d[2]);
    vel = make_float3(float(d[0] / dt), float(d[1] / dt), float(d[2] / dt));
  }
  update_pos(posxy[p], posz[p], mv.x, mv.y, mv.z, false, p, c, posxy, posz, dcell, code);
  const float4 vr = velrho[p];
  velrho[p] = make_float4(vel.x, vel.y, vel.z, vr.w);
}

// Rotation by ang degrees about the axis p1->p2 as JMatrix4::MatrixRot
// (JMatrix4.h:465-497) composes it (translate, align the axis with z,
// rotate, undo): a rotation by -ang about the unit direction u = (p2-p1)/|p2-p1|
// in the right-hand sense, here in Rodrigues' closed form
// R = cos(t) I + sin(t) [u]x + (1 - cos(t)) u u^T with t = -ang; x' = R (x - p1) + p1.
// m: row-major 3x4.
__host__ __device__ inline void rot_matrix(double ang, const double p1[3], const double p2[3],
                                           double m[12]) {
  const double torad = 3.14159265358979323846 / 180.0;
  double u[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
  const double l = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  for (int k = 0; k < 3; k++) u[k] /= l;
  const double t = -ang * torad;
  const double c = cos(t), sn = sin(t), oc = 1.0 - c;
  const double r[3][3] = {
      {c + oc * u[0] * u[0], oc * u[0] * u[1] - sn * u[2], oc * u[0] * u[2] + sn * u[1]},
      {oc * u[1] * u[0] + sn * u[2], c + oc * u[1] * u[1], oc * u[1] * u[2] - sn * u[0]},
      {oc * u[2] * u[0] - sn * u[1], oc * u[2] * u[1] + sn * u[0], c + oc * u[2] * u[2]}};
  for (int i = 0; i < 3; i++) {
    m[4 * i] = r[i][0]; m[4 * i + 1] = r[i][1]; m[4 * i + 2] = r[i][2];
    m[4 * i + 3] = p1[i] - (r[i][0] * p1[0] + r[i][1] * p1[1] + r[i][2] * p1[2]);
  }
}

// Matrix of a rotation motion over [t, t+dt] (JMotionObj.cpp:518-524, 577-587).
__host__ __device__ inline void motion_matrix(const dsph_motion& m, double t, double dt,
                                              double mat[12]) {
  const double twopi = 6.283185307179586476925286766559;
  double ang = 0;
  if (m.type == 2)
    ang = m.angampl * sin(m.angphase + twopi * m.angfreq * (t + dt)) -
          m.angampl * sin(m.angphase + twopi * m.angfreq * t);
  else if (m.type == 3)
    ang = m.angvel * dt;
  rot_matrix(ang, m.axisp1, m.axisp2, mat);
}

struct Mat34 {
  double a[12];
};

// KerMoveMatBound (JSphGpu_ker.cu:1705-1740) for the object's particles
// (selected by id). With a device dt slot array the matrix of the step is
// computed here from the motion law.
__global__ void __launch_bounds__(kBs) k_move_mat_bound(unsigned npb, unsigned idbegin,
    unsigned count, PosCte c, Mat34 mh, double dth, dsph_motion mo, double t0,
    const double* __restrict__ d_dt, const unsigned* __restrict__ idp,
    double2* __restrict__ posxy, double* __restrict__ posz, unsigned* __restrict__ dcell,
    float4* __restrict__ velrho, typecode* __restrict__ code, float* __restrict__ boundnor) {
  const unsigned p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npb) return;
  if (idp[p] - idbegin >= count) return;
  Mat34 m = mh;
  double dt = dth;
  if (d_dt) {
    dt = d_dt[2];
    const double t = t0 + d_dt[3];
    if (!(t >= mo.start && t < mo.finish) || dt <= 0) return;
    motion_matrix(mo, t, dt, m.a);
  }
  const double2 rxy = posxy[p];
  const double rx = rxy.x, ry = rxy.y, rz = posz[p];
  const double x2 = rx * m.a[0] + ry * m.a[1] + rz * m.a[2] + m.a[3];
  // 2-D: y kept (KerMoveMatBound<.., simulate2d>, JSphGpu_ker.cu:1721).
  const double y2 = c.sim2d ? ry : rx * m.a[4] + ry * m.a[5] + rz * m.a[6] + m.a[7];
  const double z2 = rx * m.a[8] + ry * m.a[9] + rz * m.a[10] + m.a[11];
  const double dx = x2 - rx, dy = y2 - ry, dz = z2 - rz;
  update_pos(rxy, rz, dx, dy, dz, false, p, c, posxy, posz, dcell, code);
  const float4 vr = velrho[p];
  velrho[p] = make_float4(float(dx / dt), float(dy / dt), float(dz / dt), vr.w);
  if (boundnor) {
    float* bn = boundnor + 3 * size_t(p);
    const double gx = rx + bn[0], gy = ry + bn[1], gz = rz + bn[2];
    const double gx2 = gx * m.a[0] + gy * m.a[1] + gz * m.a[2] + m.a[3];
    const double gy2 = gx * m.a[4] + gy * m.a[5] + gz * m.a[6] + m.a[7];
    const double gz2 = gx * m.a[8] + gy * m.a[9] + gz * m.a[10] + m.a[11];
    bn[0] = float(gx2 - x2); bn[1] = float(gy2 - y2); bn[2] = float(gz2 - z2);
  }
}

// Shifting displacement of fluid particles (KerRunShifting,
// JSphShifting_ker.cu:123-151): umagn = dt*coef*h*|v| (FP64), scaled by the
// free-surface factor (w - tfs)/(3 - tfs) above the threshold and 0 below,
// shift = sum*umagn clamped to +-maxdist; a cancelled particle gets 0.
__global__ void __launch_bounds__(kBs) k_run_shifting(unsigned n, unsigned pini, double dt_h,
    const double* __restrict__ d_dt, unsigned slot, double scale, double shiftcoef,
    double kernelh, float tfs, float coeftfs, float maxdist, const float4* __restrict__ velrho,
    float4* __restrict__ shiftposfs) {
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const unsigned p = i + pini;
  const double dt = d_dt ? d_dt[slot] * scale : dt_h;
  const double coefumagn = dt * shiftcoef * kernelh;
  float4 rs = shiftposfs[p];
  if (rs.x != FLT_MAX) {
    const float4 v = velrho[p];
    const double vx = v.x, vy = v.y, vz = v.z;
    double umagn = coefumagn * sqrt(vx * vx + vy * vy + vz * vz);
    if (tfs) {
      if (rs.w < tfs) umagn = 0;
      else umagn *= (double(rs.w) - tfs) / coeftfs;
    }
    const float sx = float(double(rs.x) * umagn), sy = float(double(rs.y) * umagn),
                sz = float(double(rs.z) * umagn);
    rs.x = (fabsf(sx) < maxdist ? sx : (sx >= 0 ? maxdist : -maxdist));
    rs.y = (fabsf(sy) < maxdist ? sy : (sy >= 0 ? maxdist : -maxdist));
    rs.z = (fabsf(sz) < maxdist ? sz : (sz >= 0 ? maxdist : -maxdist));
  } else {
    rs = make_float4(0, 0, 0, 0);
  }
  shiftposfs[p] = rs;
}

__global__ void __launch_bounds__(kBs) k_step_pos2(unsigned n, unsigned pini, PosCte c,
    const double2* __restrict__ posxypre, const double* __restrict__ poszpre,
    const double2* __restrict__ movxy, const double* __restrict__ movz,
    double2* __restrict__ posxy, double* __restrict__ posz, unsigned* __restrict__ dcell,
    typecode* __restrict__ code) {
  const unsigned pt = blockIdx.x * blockDim.x + threadIdx.x;
  if (pt >= n) return;
  const unsigned p = pt + pini;
  const unsigned rcode = code[p];
  const bool outrho = code_is_outrho(rcode);
  if (c.periactive && !outrho && !code_is_normal(rcode)) return;  // periodic copies stay
  const double2 m = movxy[p];
  update_pos(posxypre[p], poszpre[p], m.x, m.y, movz[p], outrho, p, c, posxy, posz, dcell, code);
}

__global__ void __launch_bounds__(kBs) k_step_pos(unsigned n, unsigned pini, PosCte c,
    const double2* __restrict__ movxy, const double* __restrict__ movz,
    double2* __restrict__ posxy, double* __restrict__ posz, unsigned* __restrict__ dcell,
    typecode* __restrict__ code) {
  const unsigned pt = blockIdx.x * blockDim.x + threadIdx.x;
  if (pt >= n) return;
  const unsigned p = pt + pini;
  const unsigned rcode = code[p];
  const bool outrho = code_is_outrho(rcode);
  if (c.periactive && !outrho && !code_is_normal(rcode)) return;  // periodic copies stay
  if (code_type(rcode) == kCodeTypeFloating) return;  // RunFloating moves floating particles
  const double2 m = movxy[p];
  update_pos(posxy[p], posz[p], m.x, m.y, movz[p], outrho, p, c, posxy, posz, dcell, code);
}

PosCte make_posc(const dsph_cte& c) {
  PosCte r{};
  r.movlimit = c.movlimit;
  r.mapx = c.maprealposminx; r.mapy = c.maprealposminy; r.mapz = c.maprealposminz;
  r.sizex = c.maprealsizex; r.sizey = c.maprealsizey; r.sizez = c.maprealsizez;
  r.scell = c.scell;
  r.cellcode = c.cellcode;
  r.periactive = c.periactive;
  for (int k = 0; k < 3; k++) {
    r.perinc[0][k] = c.xperinc[k]; r.perinc[1][k] = c.yperinc[k]; r.perinc[2][k] = c.zperinc[k];
  }
  r.domx = c.domposminx; r.domy = c.domposminy; r.domz = c.domposminz;
  r.sim2d = c.simulate2d;
  return r;
}

// ---------------------------------------------------------------- floating bodies
// Index of each floating particle by id (KerCalcRidp, JSphGpu_ker.cu:1582-1628):
// ridp[id - idini] = p for the particles p of [pini, pini+n) with ids in
// [idini, idfin).
__global__ void __launch_bounds__(kBs) k_ft_calc_ridp(unsigned n, unsigned pini, unsigned idini,
    unsigned idfin, const unsigned* __restrict__ idp, unsigned* __restrict__ ridp) {
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const unsigned p = i + pini;
  const unsigned id = idp[p];
  if (id >= idini && id < idfin) ridp[id - idini] = p;
}

// Sums over the particles of body blockIdx.x of ace and of (pos - centre) x ace
// (KerFtPartsSumAce, JSphGpu_ker.cu:1938-2012): thread t accumulates the
// body's particles t, t+256, ... in order, thread 0 then adds the 256
// partial sums in order, all in FP32 without contraction.
constexpr unsigned kFtBs = 256;
__global__ void __launch_bounds__(kFtBs) k_ft_sum_ace(const unsigned* __restrict__ ftbody,
    const double* __restrict__ center, const unsigned* __restrict__ ridp,
    const double2* __restrict__ posxy, const double* __restrict__ posz,
    const float* __restrict__ ace, float* __restrict__ acelinang) {
#pragma clang fp contract(off)
  __shared__ float part[6][kFtBs];
  const unsigned cf = blockIdx.x, tid = threadIdx.x;
  const unsigned fpini = ftbody[2 * cf], fnp = ftbody[2 * cf + 1];
  const double cx = center[3 * cf], cy = center[3 * cf + 1], cz = center[3 * cf + 2];
  float l[3] = {0.f, 0.f, 0.f}, g[3] = {0.f, 0.f, 0.f};
  for (unsigned fp = tid; fp < fnp; fp += kFtBs) {
    const unsigned p = ridp[fpini + fp];
    if (p == UINT_MAX) continue;   // particle not present (excluded)
    const float ax = ace[3 * size_t(p)], ay = ace[3 * size_t(p) + 1], az = ace[3 * size_t(p) + 2];
    l[0] += ax; l[1] += ay; l[2] += az;
    const double2 rxy = posxy[p];
    const float dx = float(rxy.x - cx), dy = float(rxy.y - cy), dz = float(posz[p] - cz);
    g[0] += (az * dy - ay * dz);
    g[1] += (ax * dz - az * dx);
    g[2] += (ay * dx - ax * dy);
  }
  for (int k = 0; k < 3; k++) { part[k][tid] = l[k]; part[3 + k][tid] = g[k]; }
  __syncthreads();
  if (tid == 0) {
    const unsigned ntid = min(fnp, kFtBs);
    float r[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (unsigned t = 0; t < ntid; t++)
      for (int k = 0; k < 6; k++) r[k] += part[k][t];
    for (int k = 0; k < 6; k++) acelinang[6 * cf + k] = r[k];
  }
}

// Moves the particles of one body (KerFtPartsUpdate, JSphGpu_ker.cu:2027-2075,
// without periodic axes or mDBC normals): displacement dt times the
// particle's velocity (KerUpdatePos), then the rigid velocity
// fvel + fomega x (pos - fcenter), in FP32 without contraction.
__global__ void __launch_bounds__(kBs) k_ft_parts_update(unsigned n, unsigned fpini, double dt,
    PosCte c, float3 fvel, float3 fomega, double3 fcenter, const unsigned* __restrict__ ridp,
    double2* __restrict__ posxy, double* __restrict__ posz, float4* __restrict__ velrho,
    unsigned* __restrict__ dcell, typecode* __restrict__ code) {
#pragma clang fp contract(off)
  const unsigned fp = blockIdx.x * blockDim.x + threadIdx.x;
  if (fp >= n) return;
  const unsigned p = ridp[fpini + fp];
  if (p == UINT_MAX) return;   // particle not present (excluded)
  float4 vr = velrho[p];
  const double dx = dt * double(vr.x), dy = dt * double(vr.y), dz = dt * double(vr.z);
  update_pos(posxy[p], posz[p], dx, dy, dz, false, p, c, posxy, posz, dcell, code);
  const double2 rxy = posxy[p];
  const float distx = float(rxy.x - fcenter.x), disty = float(rxy.y - fcenter.y),
              distz = float(posz[p] - fcenter.z);
  vr.x = fvel.x + (fomega.y * distz - fomega.z * disty);
  vr.y = fvel.y + (fomega.z * distx - fomega.x * distz);
  vr.z = fvel.z + (fomega.x * disty - fomega.y * distx);
  velrho[p] = vr;
}

// ComputeAceMod with codes (JSphGpu_ker.cu:260-275): |ace|^2 of normal
// particles, 0 for periodic copies.
__global__ void __launch_bounds__(kBs) k_acemod_normal(unsigned n, unsigned pini,
    const float* __restrict__ ace, const typecode* __restrict__ code, float* __restrict__ out) {
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const unsigned p = i + pini;
  const float ax = ace[3 * size_t(p)], ay = ace[3 * size_t(p) + 1], az = ace[3 * size_t(p) + 2];
  out[i] = code_is_normal(code[p]) ? ax * ax + ay * ay + az * az : 0.f;
}

// ---------------------------------------------------------------- ensembles
__global__ void __launch_bounds__(kBs) k_ens_zero(unsigned n, float* __restrict__ ace,
                                                  float* __restrict__ ar,
                                                  float* __restrict__ viscdt) {
  for (unsigned p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
    ace[3 * size_t(p)] = 0.f;
    ace[3 * size_t(p) + 1] = 0.f;
    ace[3 * size_t(p) + 2] = 0.f;
    ar[p] = 0.f;
    viscdt[p] = 0.f;
  }
}

// WHICH 0: max |v|^2 over fluid; 1: max viscdt over all; 2: max |ace|^2 over fluid.
template <int WHICH>
__global__ void __launch_bounds__(kBs) k_ens_reduce(const EnsCase* __restrict__ cases,
    EnsState* __restrict__ st, const EnsBlock* __restrict__ pmap,
    const float4* __restrict__ velrho, const float* __restrict__ viscdt,
    const float* __restrict__ ace) {
  const EnsBlock e = pmap[blockIdx.x];
  if (!st[e.x].active) return;
  const EnsCase& ec = cases[e.x];
  const unsigned l = e.y + threadIdx.x;
  const unsigned lo = (WHICH == 1 ? 0u : ec.npb);
  const bool in = l >= lo && l < st[e.x].np;
  float v = -FLT_MAX;
  if (in) {
    const unsigned p = ec.p0 + l;
    if (WHICH == 0) {
      const float4 vr = velrho[p];
      v = vr.x * vr.x + vr.y * vr.y + vr.z * vr.z;
    } else if (WHICH == 1) {
      v = viscdt[p];
    } else {
      const float x = ace[3 * size_t(p)], y = ace[3 * size_t(p) + 1], z = ace[3 * size_t(p) + 2];
      v = x * x + y * y + z * z;
    }
  }
  v = block_max(v);
  // The values are >= 0, so their FP32 bits order like the values.
  if (threadIdx.x == 0 && v >= 0.f) atomicMax(&st[e.x].red[WHICH], __float_as_uint(v));
}

__global__ void k_ens_dt(const EnsCase* __restrict__ cases, EnsState* __restrict__ st,
                         unsigned ncases, unsigned slot) {
  const unsigned k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ncases) return;
  EnsState& s = st[k];
  if (!s.active) return;
  const EnsCase& ec = cases[k];
  unsigned fl;
  s.dt[slot] = dt_of(__uint_as_float(s.red[0]), __uint_as_float(s.red[1]),
                     __uint_as_float(s.red[2]), ec.kernelh, ec.cs0, ec.cfl, ec.dtmin, fl);
  s.red[0] = s.red[1] = s.red[2] = 0u;
  if (fl) s.flags |= fl;
  if (fl & 1u) s.active = 0;
}

__global__ void __launch_bounds__(kBs) k_ens_sym(const EnsCase* __restrict__ cases,
    const EnsState* __restrict__ st, const EnsBlock* __restrict__ pmap, bool corrector,
    const float4* __restrict__ velrhopre, const float* __restrict__ ar,
    const float* __restrict__ ace, typecode* __restrict__ code, double2* __restrict__ movxy,
    double* __restrict__ movz, float4* __restrict__ velrho) {
  const EnsBlock e = pmap[blockIdx.x];
  const unsigned l = e.y + threadIdx.x;
  const EnsCase& ec = cases[e.x];
  if (!st[e.x].active) {
    // The predictor swapped the state arrays: a stopped case keeps its state.
    if (!corrector && l < ec.np0) velrho[ec.p0 + l] = velrhopre[ec.p0 + l];
    return;
  }
  if (l >= st[e.x].np) return;
  const double dt = st[e.x].dtpre;
  const float3 g = make_float3(ec.gx, ec.gy, ec.gz);
  if (corrector)
    sym_cor_one(ec.p0 + l, l < ec.npb, dt * 1.0, dt * 0.5, velrhopre, ar, ace, ec.c.rhopzero,
                ec.rhopoutmin, ec.rhopoutmax, g, code, movxy, movz, velrho);
  else
    sym_pre_one(ec.p0 + l, l < ec.npb, dt * 0.5, velrhopre, ar, ace, ec.c.rhopzero,
                ec.rhopoutmin, ec.rhopoutmax, g, code, movxy, movz, velrho);
}

__global__ void __launch_bounds__(kBs) k_ens_pos2(const EnsCase* __restrict__ cases,
    const EnsState* __restrict__ st, const EnsBlock* __restrict__ pmap, bool copy_bound,
    const double2* __restrict__ posxypre, const double* __restrict__ poszpre,
    const double2* __restrict__ movxy, const double* __restrict__ movz,
    double2* __restrict__ posxy, double* __restrict__ posz, unsigned* __restrict__ dcell,
    typecode* __restrict__ code) {
  const EnsBlock e = pmap[blockIdx.x];
  const unsigned l = e.y + threadIdx.x;
  const EnsCase& ec = cases[e.x];
  const unsigned p = ec.p0 + l;
  if (!st[e.x].active) {
    // The predictor swapped the position arrays: a stopped case keeps them.
    if (copy_bound && l < ec.np0) {
      posxy[p] = posxypre[p];
      posz[p] = poszpre[p];
    }
    return;
  }
  if (l >= st[e.x].np) return;
  if (l < ec.npb) {
    if (copy_bound) {
      posxy[p] = posxypre[p];
      posz[p] = poszpre[p];
    }
    return;
  }
  const bool outrho = code_is_outrho(code[p]);
  const double2 m = movxy[p];
  update_pos(posxypre[p], poszpre[p], m.x, m.y, movz[p], outrho, p, ec.pc, posxy, posz, dcell,
             code);
}

__global__ void k_ens_advance(const EnsCase* __restrict__ cases, EnsState* __restrict__ st,
                              unsigned ncases) {
  const unsigned k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ncases) return;
  EnsState& s = st[k];
  if (!s.active) return;
  const double dt = s.dtpre;
  s.time += dt;
  s.dtlast = dt;
  s.dtpre = fmin(s.dt[0], s.dt[1]);
  s.nstep++;
  if (s.time >= cases[k].tend) s.active = 0;
}

}  // namespace

namespace dsph {

int ens_zero_forces(dsph_ctx* ctx, unsigned n, float* ace, float* ar, float* viscdt) {
  if (!n) return DSPH_OK;
  k_ens_zero<<<std::min(grid_for(n, kBs), 4096u), kBs, 0, ctx->stream>>>(n, ace, ar, viscdt);
  DSPH_LAUNCHED(ctx, "k_ens_zero");
  return DSPH_OK;
}

int ens_reduce(dsph_ctx* ctx, const EnsCase* cases, EnsState* st, const EnsBlock* pmap,
               unsigned npblocks, int which, const float4* velrho, const float* viscdt,
               const float* ace) {
  if (!npblocks) return DSPH_OK;
  KTimer kt(ctx, "k_redu_max");
  switch (which) {
    case 0: k_ens_reduce<0><<<npblocks, kBs, 0, ctx->stream>>>(cases, st, pmap, velrho, viscdt, ace); break;
    case 1: k_ens_reduce<1><<<npblocks, kBs, 0, ctx->stream>>>(cases, st, pmap, velrho, viscdt, ace); break;
    default: k_ens_reduce<2><<<npblocks, kBs, 0, ctx->stream>>>(cases, st, pmap, velrho, viscdt, ace); break;
  }
  DSPH_LAUNCHED(ctx, "k_ens_reduce");
  return DSPH_OK;
}

int ens_dt(dsph_ctx* ctx, const EnsCase* cases, EnsState* st, unsigned ncases, unsigned slot) {
  if (!ncases) return DSPH_OK;
  k_ens_dt<<<grid_for(ncases, 64), 64, 0, ctx->stream>>>(cases, st, ncases, slot);
  DSPH_LAUNCHED(ctx, "k_ens_dt");
  return DSPH_OK;
}

int ens_sym_pre(dsph_ctx* ctx, const EnsCase* cases, const EnsState* st, const EnsBlock* pmap,
                unsigned npblocks, const float4* velrhopre, const float* ar, const float* ace,
                uint16_t* code, double2* movxy, double* movz, float4* velrho) {
  if (!npblocks) return DSPH_OK;
  KTimer kt(ctx, "k_sym_pre");
  k_ens_sym<<<npblocks, kBs, 0, ctx->stream>>>(cases, st, pmap, false, velrhopre, ar, ace, code,
                                               movxy, movz, velrho);
  DSPH_LAUNCHED(ctx, "k_ens_sym_pre");
  return DSPH_OK;
}

int ens_sym_cor(dsph_ctx* ctx, const EnsCase* cases, const EnsState* st, const EnsBlock* pmap,
                unsigned npblocks, const float4* velrhopre, const float* ar, const float* ace,
                uint16_t* code, double2* movxy, double* movz, float4* velrho) {
  if (!npblocks) return DSPH_OK;
  KTimer kt(ctx, "k_sym_cor");
  k_ens_sym<<<npblocks, kBs, 0, ctx->stream>>>(cases, st, pmap, true, velrhopre, ar, ace, code,
                                               movxy, movz, velrho);
  DSPH_LAUNCHED(ctx, "k_ens_sym_cor");
  return DSPH_OK;
}

int ens_pos2(dsph_ctx* ctx, const EnsCase* cases, const EnsState* st, const EnsBlock* pmap,
             unsigned npblocks, bool copy_bound, const double2* posxypre, const double* poszpre,
             const double2* movxy, const double* movz, double2* posxy, double* posz,
             unsigned* dcell, uint16_t* code) {
  if (!npblocks) return DSPH_OK;
  KTimer kt(ctx, "k_step_pos2");
  k_ens_pos2<<<npblocks, kBs, 0, ctx->stream>>>(cases, st, pmap, copy_bound, posxypre, poszpre,
                                                movxy, movz, posxy, posz, dcell, code);
  DSPH_LAUNCHED(ctx, "k_ens_pos2");
  return DSPH_OK;
}

int ens_advance(dsph_ctx* ctx, EnsState* st, const EnsCase* cases, unsigned ncases) {
  if (!ncases) return DSPH_OK;
  k_ens_advance<<<grid_for(ncases, 64), 64, 0, ctx->stream>>>(cases, st, ncases);
  DSPH_LAUNCHED(ctx, "k_ens_advance");
  return DSPH_OK;
}

}  // namespace dsph

extern "C" {

int dsph_velmax2(dsph_ctx* ctx, uint32_t n, uint32_t pini, const float* d_velrho, float* d_out) {
  if (!ctx || !d_out) return DSPH_E_ARG;
  return redu_max<1>(ctx, n, pini, d_velrho, d_out);
}

int dsph_acemax2(dsph_ctx* ctx, uint32_t n, uint32_t pini, const float* d_ace, float* d_out) {
  if (!ctx || !d_out) return DSPH_E_ARG;
  return redu_max<2>(ctx, n, pini, d_ace, d_out);
}

int dsph_acemax2_normal(dsph_ctx* ctx, uint32_t n, uint32_t pini, const float* d_ace,
                        const uint16_t* d_code, float* d_work, float* d_out) {
  if (!ctx || !d_out) return DSPH_E_ARG;
  if (!n) return redu_max<0>(ctx, 0, 0, d_work, d_out);
  DSPH_REQUIRE(ctx, d_ace && d_code && d_work, "dsph_acemax2_normal: null argument");
  k_acemod_normal<<<grid_for(n, kBs), kBs, 0, ctx->stream>>>(n, pini, d_ace, d_code, d_work);
  DSPH_LAUNCHED(ctx, "k_acemod_normal");
  return redu_max<0>(ctx, n, 0, d_work, d_out);
}

int dsph_redu_max_f32(dsph_ctx* ctx, uint32_t n, uint32_t ini, const float* d_data,
                      float* d_out) {
  if (!ctx || !d_out) return DSPH_E_ARG;
  return redu_max<0>(ctx, n, ini, d_data, d_out);
}

int dsph_redu_max_f32_sync(dsph_ctx* ctx, uint32_t n, uint32_t ini, const float* d_data,
                           float* h_out) {
  if (!ctx || !h_out) return DSPH_E_ARG;
  float* d_out = reinterpret_cast<float*>(ctx->small_dev) + 32;
  int rc = redu_max<0>(ctx, n, ini, d_data, d_out);
  if (rc) return rc;
  DSPH_HIP(ctx, hipMemcpyAsync(ctx->pinned, d_out, sizeof(float), hipMemcpyDeviceToHost,
                               ctx->stream));
  DSPH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *h_out = *static_cast<float*>(ctx->pinned);
  return DSPH_OK;
}

int dsph_dt_variable(dsph_ctx* ctx, const float* d_red, double kernelh, double cs0, double cfl,
                     double dtmin, double* d_dt, uint32_t slot, uint32_t* d_flags) {
  if (!ctx || !d_red || !d_dt) return DSPH_E_ARG;
  k_dt_variable<<<1, 64, 0, ctx->stream>>>(d_red, kernelh, cs0, cfl, dtmin, d_dt, slot, d_flags);
  DSPH_LAUNCHED(ctx, "k_dt_variable");
  return DSPH_OK;
}

int dsph_dt_advance(dsph_ctx* ctx, double* d_dt) {
  if (!ctx || !d_dt) return DSPH_E_ARG;
  k_dt_advance<<<1, 64, 0, ctx->stream>>>(d_dt);
  DSPH_LAUNCHED(ctx, "k_dt_advance");
  return DSPH_OK;
}

int dsph_step_symplectic_pre(dsph_ctx* ctx, uint32_t np, uint32_t npb, const float* d_velrhopre,
                             const uint8_t* d_boundmode, const float* d_ar, const float* d_ace,
                             const float* d_shiftposfs, double dtm,
                             const double* d_dt, uint32_t dt_slot, float rhopzero,
                             float rhopoutmin, float rhopoutmax, const float gravity[3],
                             uint16_t* d_code, double* d_movxy, double* d_movz,
                             float* d_velrho) {
  if (!ctx || !gravity) return DSPH_E_ARG;
  if (!np) return DSPH_OK;
  DSPH_REQUIRE(ctx, d_velrhopre && d_ar && d_ace && d_code && d_movxy && d_movz && d_velrho,
               "dsph_step_symplectic_pre: null array");
  KTimer kt(ctx, "k_sym_pre");
  k_sym_pre<<<grid_for(np, kBs), kBs, 0, ctx->stream>>>(
      np, npb, (const float4*)d_velrhopre, d_boundmode, (const float4*)d_shiftposfs, d_ar, d_ace,
      dtm, d_dt, dt_slot,
      rhopzero, rhopoutmin,
      rhopoutmax, make_float3(gravity[0], gravity[1], gravity[2]), d_code, (double2*)d_movxy,
      d_movz, (float4*)d_velrho);
  DSPH_LAUNCHED(ctx, "k_sym_pre");
  return DSPH_OK;
}

int dsph_step_symplectic_cor(dsph_ctx* ctx, uint32_t np, uint32_t npb, const float* d_velrhopre,
                             const uint8_t* d_boundmode, const float* d_ar, const float* d_ace,
                             const float* d_shiftposfs, double dtm, double dt,
                             const double* d_dt, uint32_t dt_slot, float rhopzero,
                             float rhopoutmin, float rhopoutmax, const float gravity[3],
                             uint16_t* d_code, double* d_movxy, double* d_movz,
                             float* d_velrho) {
  if (!ctx || !gravity) return DSPH_E_ARG;
  if (!np) return DSPH_OK;
  DSPH_REQUIRE(ctx, d_velrhopre && d_ar && d_ace && d_code && d_movxy && d_movz && d_velrho,
               "dsph_step_symplectic_cor: null array");
  KTimer kt(ctx, "k_sym_cor");
  k_sym_cor<<<grid_for(np, kBs), kBs, 0, ctx->stream>>>(
      np, npb, (const float4*)d_velrhopre, d_boundmode, (const float4*)d_shiftposfs, d_ar, d_ace,
      dtm, dt, d_dt, dt_slot,
      rhopzero,
      rhopoutmin, rhopoutmax, make_float3(gravity[0], gravity[1], gravity[2]), d_code,
      (double2*)d_movxy, d_movz, (float4*)d_velrho);
  DSPH_LAUNCHED(ctx, "k_sym_cor");
  return DSPH_OK;
}

int dsph_step_verlet(dsph_ctx* ctx, uint32_t np, uint32_t npb, const float* d_velrho1,
                     const float* d_velrho2, const uint8_t* d_boundmode, const float* d_ar,
                     const float* d_ace, const float* d_shiftposfs, double dt,
                     double dt2, float rhopzero, float rhopoutmin, float rhopoutmax,
                     const float gravity[3], uint16_t* d_code, double* d_movxy, double* d_movz,
                     float* d_velrhonew) {
  if (!ctx || !gravity) return DSPH_E_ARG;
  if (!np) return DSPH_OK;
  DSPH_REQUIRE(ctx, d_velrho1 && d_velrho2 && d_ar && d_ace && d_code && d_movxy && d_movz &&
                        d_velrhonew,
               "dsph_step_verlet: null array");
  const double dt205 = 0.5 * dt * dt;
  KTimer kt(ctx, "k_verlet");
  k_verlet<<<grid_for(np, kBs), kBs, 0, ctx->stream>>>(
      np, npb, (const float4*)d_velrho1, (const float4*)d_velrho2, d_boundmode,
      (const float4*)d_shiftposfs, d_ar, d_ace, dt,
      dt205, dt2,
      rhopzero, rhopoutmin, rhopoutmax, make_float3(gravity[0], gravity[1], gravity[2]), d_code,
      (double2*)d_movxy, d_movz, (float4*)d_velrhonew);
  DSPH_LAUNCHED(ctx, "k_verlet");
  return DSPH_OK;
}

int dsph_mdbc2_nopen_step(dsph_ctx* ctx, uint32_t np, uint32_t npb,
                          const float* d_nopenshift, const float* d_velbase, double dt,
                          const double* d_dt, uint32_t dt_slot, float* d_velrho,
                          double* d_movxy, double* d_movz) {
  if (!ctx) return DSPH_E_ARG;
  if (np <= npb) return DSPH_OK;
  DSPH_REQUIRE(ctx, d_nopenshift && d_velbase && d_velrho && d_movxy && d_movz,
               "dsph_mdbc2_nopen_step: null array");
  KTimer kt(ctx, "k_nopen_step");
  k_nopen_step<<<grid_for(np - npb, kBs), kBs, 0, ctx->stream>>>(
      np, npb, (const float4*)d_nopenshift, (const float4*)d_velbase, dt, d_dt, dt_slot,
      (float4*)d_velrho, (double2*)d_movxy, d_movz);
  DSPH_LAUNCHED(ctx, "k_nopen_step");
  return DSPH_OK;
}

int dsph_mdbc2_motion_vel_ace(dsph_ctx* ctx, uint32_t npb, uint32_t idbegin, uint32_t count,
                              double dt, const double* d_dt, uint32_t dt_slot,
                              const uint32_t* d_idp, const float* d_velrho, float* d_motionvel,
                              float* d_motionace) {
  if (!ctx) return DSPH_E_ARG;
  if (!npb || !count) return DSPH_OK;
  DSPH_REQUIRE(ctx, d_idp && d_velrho && d_motionvel && d_motionace,
               "dsph_mdbc2_motion_vel_ace: null array");
  DSPH_REQUIRE(ctx, d_dt || dt > 0, "dsph_mdbc2_motion_vel_ace: dt must be > 0");
  KTimer kt(ctx, "k_motion_vel_ace");
  k_motion_vel_ace<<<grid_for(npb, kBs), kBs, 0, ctx->stream>>>(
      npb, idbegin, count, dt, d_dt, dt_slot, d_idp, (const float4*)d_velrho, d_motionvel,
      d_motionace);
  DSPH_LAUNCHED(ctx, "k_motion_vel_ace");
  return DSPH_OK;
}

int dsph_run_shifting(dsph_ctx* ctx, uint32_t n, uint32_t pini, double dt, const double* d_dt,
                      uint32_t dt_slot, double dt_scale, double shiftcoef, double kernelh,
                      float shifttfs, double dp, const float* d_velrho, float* d_shiftposfs) {
  if (!ctx) return DSPH_E_ARG;
  if (!n) return DSPH_OK;
  DSPH_REQUIRE(ctx, d_velrho && d_shiftposfs, "dsph_run_shifting: null array");
  KTimer kt(ctx, "k_run_shifting");
  k_run_shifting<<<grid_for(n, kBs), kBs, 0, ctx->stream>>>(
      n, pini, dt, d_dt, dt_slot, dt_scale, shiftcoef, kernelh, shifttfs,
      ((ctx->cte_set && ctx->cte.simulate2d) ? 2.0f : 3.0f) - shifttfs,
      float(dp * 0.1), (const float4*)d_velrho, (float4*)d_shiftposfs);
  DSPH_LAUNCHED(ctx, "k_run_shifting");
  return DSPH_OK;
}

int dsph_step_pos2(dsph_ctx* ctx, uint32_t np, uint32_t npb, const double* d_posxypre,
                   const double* d_poszpre, const double* d_movxy, const double* d_movz,
                   double* d_posxy, double* d_posz, uint32_t* d_dcell, uint16_t* d_code) {
  if (!ctx) return DSPH_E_ARG;
  if (!ctx->cte_set) return set_error(ctx, DSPH_E_STATE, "constants not set");
  if (np <= npb) return DSPH_OK;
  const unsigned n = np - npb;
  KTimer kt(ctx, "k_step_pos2");
  k_step_pos2<<<grid_for(n, kBs), kBs, 0, ctx->stream>>>(
      n, npb, make_posc(ctx->cte), (const double2*)d_posxypre, d_poszpre, (const double2*)d_movxy,
      d_movz, (double2*)d_posxy, d_posz, d_dcell, d_code);
  DSPH_LAUNCHED(ctx, "k_step_pos2");
  return DSPH_OK;
}

int dsph_step_pos(dsph_ctx* ctx, uint32_t np, uint32_t npb, const double* d_movxy,
                  const double* d_movz, double* d_posxy, double* d_posz, uint32_t* d_dcell,
                  uint16_t* d_code) {
  if (!ctx) return DSPH_E_ARG;
  if (!ctx->cte_set) return set_error(ctx, DSPH_E_STATE, "constants not set");
  if (np <= npb) return DSPH_OK;
  const unsigned n = np - npb;
  k_step_pos<<<grid_for(n, kBs), kBs, 0, ctx->stream>>>(
      n, npb, make_posc(ctx->cte), (const double2*)d_movxy, d_movz, (double2*)d_posxy, d_posz,
      d_dcell, d_code);
  DSPH_LAUNCHED(ctx, "k_step_pos");
  return DSPH_OK;
}

int dsph_ft_calc_ridp(dsph_ctx* ctx, uint32_t n, uint32_t pini, uint32_t idini,
                      uint32_t idfin, const uint32_t* d_idp, uint32_t* d_ridp) {
  if (!ctx) return DSPH_E_ARG;
  if (!n) return DSPH_OK;
  DSPH_REQUIRE(ctx, d_idp && d_ridp && idfin >= idini, "dsph_ft_calc_ridp: invalid argument");
  KTimer kt(ctx, "k_ft_calc_ridp");
  k_ft_calc_ridp<<<grid_for(n, kBs), kBs, 0, ctx->stream>>>(n, pini, idini, idfin, d_idp, d_ridp);
  DSPH_LAUNCHED(ctx, "k_ft_calc_ridp");
  return DSPH_OK;
}

int dsph_ft_parts_sum_ace(dsph_ctx* ctx, uint32_t ftcount, const uint32_t* d_ftbody,
                          const double* d_center, const uint32_t* d_ridp, const double* d_posxy,
                          const double* d_posz, const float* d_ace, float* d_acelinang) {
  if (!ctx) return DSPH_E_ARG;
  if (!ftcount) return DSPH_OK;
  DSPH_REQUIRE(ctx, d_ftbody && d_center && d_ridp && d_posxy && d_posz && d_ace && d_acelinang,
               "dsph_ft_parts_sum_ace: null argument");
  KTimer kt(ctx, "k_ft_sum_ace");
  k_ft_sum_ace<<<ftcount, kFtBs, 0, ctx->stream>>>(d_ftbody, d_center, d_ridp,
                                                   (const double2*)d_posxy, d_posz, d_ace,
                                                   d_acelinang);
  DSPH_LAUNCHED(ctx, "k_ft_sum_ace");
  return DSPH_OK;
}

int dsph_ft_parts_update(dsph_ctx* ctx, double dt, uint32_t count, uint32_t fpini,
                         const float fvel[3], const float fomega[3], const double fcenter[3],
                         const uint32_t* d_ridp, double* d_posxy, double* d_posz,
                         float* d_velrho, uint32_t* d_dcell, uint16_t* d_code) {
  if (!ctx) return DSPH_E_ARG;
  if (!ctx->cte_set) return set_error(ctx, DSPH_E_STATE, "constants not set");
  if (!count) return DSPH_OK;
  DSPH_REQUIRE(ctx, fvel && fomega && fcenter && d_ridp && d_posxy && d_posz && d_velrho &&
                        d_dcell && d_code,
               "dsph_ft_parts_update: null argument");
  DSPH_REQUIRE(ctx, !ctx->cte.periactive, "dsph_ft_parts_update: periodic axes are not built");
  KTimer kt(ctx, "k_ft_parts_update");
  k_ft_parts_update<<<grid_for(count, kBs), kBs, 0, ctx->stream>>>(
      count, fpini, dt, make_posc(ctx->cte), make_float3(fvel[0], fvel[1], fvel[2]),
      make_float3(fomega[0], fomega[1], fomega[2]),
      make_double3(fcenter[0], fcenter[1], fcenter[2]), d_ridp, (double2*)d_posxy, d_posz,
      (float4*)d_velrho, d_dcell, d_code);
  DSPH_LAUNCHED(ctx, "k_ft_parts_update");
  return DSPH_OK;
}

int dsph_move_lin_bound(dsph_ctx* ctx, uint32_t npb, uint32_t idbegin, uint32_t count,
                        const double mvpos[3], const float mvvel[3], const uint32_t* d_idp,
                        double* d_posxy, double* d_posz, uint32_t* d_dcell, float* d_velrho,
                        uint16_t* d_code) {
  if (!ctx) return DSPH_E_ARG;
  if (!ctx->cte_set) return set_error(ctx, DSPH_E_STATE, "constants not set");
  if (!npb || !count) return DSPH_OK;
  DSPH_REQUIRE(ctx, mvpos && mvvel && d_idp && d_posxy && d_posz && d_dcell && d_velrho && d_code,
               "dsph_move_lin_bound: null argument");
  dsph_motion m = {};
  KTimer kt(ctx, "k_move_lin_bound");
  k_move_lin_bound<<<grid_for(npb, kBs), kBs, 0, ctx->stream>>>(
      npb, idbegin, count, make_posc(ctx->cte), make_double3(mvpos[0], mvpos[1], mvpos[2]),
      make_float3(mvvel[0], mvvel[1], mvvel[2]), m, 0.0, nullptr, d_idp, (double2*)d_posxy,
      d_posz, d_dcell, (float4*)d_velrho, d_code);
  DSPH_LAUNCHED(ctx, "k_move_lin_bound");
  return DSPH_OK;
}

int dsph_move_mat_bound(dsph_ctx* ctx, uint32_t npb, uint32_t idbegin, uint32_t count,
                        const double mat[12], double dt, const uint32_t* d_idp, double* d_posxy,
                        double* d_posz, uint32_t* d_dcell, float* d_velrho, uint16_t* d_code,
                        float* d_boundnor) {
  if (!ctx) return DSPH_E_ARG;
  if (!ctx->cte_set) return set_error(ctx, DSPH_E_STATE, "constants not set");
  if (!npb || !count) return DSPH_OK;
  DSPH_REQUIRE(ctx, mat && d_idp && d_posxy && d_posz && d_dcell && d_velrho && d_code,
               "dsph_move_mat_bound: null argument");
  DSPH_REQUIRE(ctx, dt > 0, "dsph_move_mat_bound: dt must be > 0");
  Mat34 m;
  for (int k = 0; k < 12; k++) m.a[k] = mat[k];
  dsph_motion mo = {};
  KTimer kt(ctx, "k_move_mat_bound");
  k_move_mat_bound<<<grid_for(npb, kBs), kBs, 0, ctx->stream>>>(
      npb, idbegin, count, make_posc(ctx->cte), m, dt, mo, 0.0, nullptr, d_idp,
      (double2*)d_posxy, d_posz, d_dcell, (float4*)d_velrho, d_code, d_boundnor);
  DSPH_LAUNCHED(ctx, "k_move_mat_bound");
  return DSPH_OK;
}

int dsph_motion_matrix(const dsph_motion* m, double t, double dt, double mat[12]) {
  if (!m || !mat || (m->type != 2 && m->type != 3)) return DSPH_E_ARG;
  motion_matrix(*m, t, dt, mat);
  return DSPH_OK;
}

int dsph_move_motion_device(dsph_ctx* ctx, uint32_t npb, const dsph_motion* m, double t0,
                            const double* d_dt, const uint32_t* d_idp, double* d_posxy,
                            double* d_posz, uint32_t* d_dcell, float* d_velrho,
                            uint16_t* d_code, float* d_boundnor) {
  if (!ctx || !m) return DSPH_E_ARG;
  if (!ctx->cte_set) return set_error(ctx, DSPH_E_STATE, "constants not set");
  if (!npb || !m->count || m->type == 0) return DSPH_OK;
  DSPH_REQUIRE(ctx, m->type >= 1 && m->type <= 3, "dsph_move_motion_device: invalid motion type");
  DSPH_REQUIRE(ctx, d_dt && d_idp && d_posxy && d_posz && d_dcell && d_velrho && d_code,
               "dsph_move_motion_device: null argument");
  if (m->type != 1) {
    Mat34 mz = {};
    KTimer kt(ctx, "k_move_mat_bound");
    k_move_mat_bound<<<grid_for(npb, kBs), kBs, 0, ctx->stream>>>(
        npb, m->idbegin, m->count, make_posc(ctx->cte), mz, 0.0, *m, t0, d_dt, d_idp,
        (double2*)d_posxy, d_posz, d_dcell, (float4*)d_velrho, d_code, d_boundnor);
    DSPH_LAUNCHED(ctx, "k_move_mat_bound");
    return DSPH_OK;
  }
  KTimer kt(ctx, "k_move_lin_bound");
  k_move_lin_bound<<<grid_for(npb, kBs), kBs, 0, ctx->stream>>>(
      npb, m->idbegin, m->count, make_posc(ctx->cte), make_double3(0, 0, 0),
      make_float3(0, 0, 0), *m, t0, d_dt, d_idp, (double2*)d_posxy, d_posz, d_dcell,
      (float4*)d_velrho, d_code);
  DSPH_LAUNCHED(ctx, "k_move_lin_bound");
  return DSPH_OK;
}

}  // extern "C"

// Host-side displacement of the same law (the reference's path: motion on the
// host, JDsMotion::ProcesTime), for the host driver.
namespace dsph {
__attribute__((visibility("hidden"))) void rect_sinu_mov_host(const dsph_motion* m, double t,
                                                               double dt, double mv[3]) {
  rect_sinu_mov(*m, t, dt, mv);
}
}  // namespace dsph
