// Learned SPS closure: a per-particle MLP predicts the eddy viscosity of the
// sub-particle-scale stress from the local velocity gradient, density and
// particle spacing; it replaces the Smagorinsky viscosity of
// KerComputeSpsTau (JSphGpu_ker.cu:1387-1436) and keeps the rest of the tau
// formula. Network: 11 features -> 64 -> 64 -> 1, SiLU hidden layers,
// softplus output times a scale. Products run on the matrix cores as
// v_mfma_f32_16x16x16_bf16 (BF16 operands, FP32 accumulation); weights are
// kept in FP32 (master, Adam state) and used as BF16 copies.
//
// MFMA operand layout (16x16x16, wave of 64 lanes, g = lane/16, c = lane%16):
//   A (16x16, row-major MxK): lane holds A[c][4g..4g+3]
//   B (16x16, KxN):           lane holds B[4g..4g+3][c]
//   D (16x16, MxN):           lane holds D[4g..4g+3][c]
// A tile is 16 samples (particles). A D-layout block of Z = X W^T (rows =
// samples, cols = units) is directly the A operand of the transposed block
// (units x samples) and the B operand of (samples x units): the weight
// gradients need no data movement; only dH1 = dZ2 W2 goes through LDS.
//
// Also here: the probe filter that builds training data from a fine-
// resolution run (Shepard-filtered velocity, velocity products, density and
// velocity gradient at coarse-scale probe points).
#include <algorithm>
#include <cfloat>
#include <cmath>
#include "sph_math.h"

using namespace dsph;

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kIn = 11, kInPad = 16, kHid = 64;
// Offsets of the parameter block (FP32 master and gradients).
constexpr int kOffW1 = 0;                        // [64][16] (cols >= 11 unused)
constexpr int kOffB1 = kOffW1 + kHid * kInPad;   // [64]
constexpr int kOffW2 = kOffB1 + kHid;            // [64][64]
constexpr int kOffB2 = kOffW2 + kHid * kHid;     // [64]
constexpr int kOffW3 = kOffB2 + kHid;            // [64]
constexpr int kOffB3 = kOffW3 + kHid;            // [1]
constexpr int kNParam = kOffB3 + 1;              // 5313
// BF16 copies: w1 [64][16], w2 [64][64], w2t [64][64] (transposed), w3 [64].
constexpr int kBfW1 = 0, kBfW2 = kBfW1 + kHid * kInPad, kBfW2T = kBfW2 + kHid * kHid,
              kBfW3 = kBfW2T + kHid * kHid, kNBf = kBfW3 + kHid;

__device__ inline unsigned short f2bf(float x) {
  // Round to nearest even (finite inputs).
  unsigned u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ inline float bf2f(unsigned short b) { return __uint_as_float(unsigned(b) << 16); }
__device__ inline float silu(float z) { return z / (1.f + __expf(-z)); }
__device__ inline float dsilu(float z) {
  const float s = 1.f / (1.f + __expf(-z));
  return s * (1.f + z * (1.f - s));
}
__device__ inline float softplus(float z) { return z > 20.f ? z : log1pf(__expf(z)); }
__device__ inline float sigmoid(float z) { return 1.f / (1.f + __expf(-z)); }

__device__ inline f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ inline bf16x4 pack4(float a, float b, float c, float d) {
  bf16x4 r;
  r[0] = (short)f2bf(a); r[1] = (short)f2bf(b); r[2] = (short)f2bf(c); r[3] = (short)f2bf(d);
  return r;
}
__device__ inline bf16x4 load_bf4(const unsigned short* p) {
  bf16x4 r;
  r[0] = (short)p[0]; r[1] = (short)p[1]; r[2] = (short)p[2]; r[3] = (short)p[3];
  return r;
}
// Sum over the 16 lanes of a lane group (same g).
__device__ inline float group16_sum(float v) {
  for (int off = 1; off < 16; off <<= 1) v += __shfl_xor(v, off);
  return v;
}
// Sum over the 4 lane groups (same c).
__device__ inline float across_groups_sum(float v) {
  v += __shfl_xor(v, 16);
  v += __shfl_xor(v, 32);
  return v;
}

struct Norm {
  float mean[kInPad], scale[kInPad];
  float out_scale;
};

inline Norm make_norm(const dsph_mlp* m) {
  Norm n;
  for (int k = 0; k < kInPad; k++) {
    n.mean[k] = k < kIn ? m->feat_mean[k] : 0.f;
    n.scale[k] = k < kIn ? m->feat_scale[k] : 0.f;
  }
  n.out_scale = m->out_scale;
  return n;
}

// Forward pass of one tile of 16 samples whose (raw) features are given by
// `feat(i, k)`; keeps what the backward pass needs. Returns the network
// output o (before softplus) for rows 4g..4g+3 in o[0..3] (every lane of the
// group holds them).
struct Fwd {
  f32x4 z1[4], z2[4];   // pre-activations, D layout, blocks of 16 units
  float o[4];
};

template <typename FeatFn>
__device__ inline void forward_tile(const unsigned short* __restrict__ wbf,
                                    const float* __restrict__ par, const Norm& nm, FeatFn feat,
                                    unsigned short* __restrict__ lds_h1, Fwd& f) {
  const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
  // A operand of layer 1: X[c][4g..4g+3] (normalised, BF16).
  float x[4];
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int k = 4 * g + r;
    x[r] = k < kIn ? (feat(c, k) - nm.mean[k]) * nm.scale[k] : 0.f;
  }
  const bf16x4 ax = pack4(x[0], x[1], x[2], x[3]);
  // Layer 1: Z1 = X W1^T + b1 (4 blocks of 16 units).
#pragma unroll
  for (int jb = 0; jb < 4; jb++) {
    const bf16x4 bw = load_bf4(wbf + kBfW1 + (jb * 16 + c) * kInPad + 4 * g);
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma(ax, bw, acc);
    const float b = par[kOffB1 + jb * 16 + c];
#pragma unroll
    for (int r = 0; r < 4; r++) {
      acc[r] += b;
      // H1 to LDS as [sample][unit] BF16 for the A operand of layer 2.
      lds_h1[(4 * g + r) * kHid + jb * 16 + c] = f2bf(silu(acc[r]));
    }
    f.z1[jb] = acc;
  }
  __syncthreads();
  // Layer 2: Z2 = H1 W2^T + b2.
  bf16x4 ah[4];
#pragma unroll
  for (int kb = 0; kb < 4; kb++) ah[kb] = load_bf4(lds_h1 + c * kHid + kb * 16 + 4 * g);
#pragma unroll
  for (int jb = 0; jb < 4; jb++) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < 4; kb++)
      acc = mfma(ah[kb], load_bf4(wbf + kBfW2 + (jb * 16 + c) * kHid + kb * 16 + 4 * g), acc);
    const float b = par[kOffB2 + jb * 16 + c];
#pragma unroll
    for (int r = 0; r < 4; r++) acc[r] += b;
    f.z2[jb] = acc;
  }
  // Layer 3 (64 -> 1) on the VALU: o[row] = sum_j silu(Z2[row][j]) w3[j] + b3.
  float part[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int jb = 0; jb < 4; jb++) {
    const float w = bf2f(wbf[kBfW3 + jb * 16 + c]);
#pragma unroll
    for (int r = 0; r < 4; r++) part[r] += silu(f.z2[jb][r]) * w;
  }
  const float b3 = par[kOffB3];
#pragma unroll
  for (int r = 0; r < 4; r++) f.o[r] = group16_sum(part[r]) + b3;
}

// ---------------------------------------------------------------- inference
// nu[i] = out_scale * softplus(MLP(features[i])) for i in [0,n); features are
// [n][11] FP32 rows.
__global__ void __launch_bounds__(64) k_mlp_forward(unsigned n, const float* __restrict__ feat,
    const unsigned short* __restrict__ wbf, const float* __restrict__ par, Norm nm,
    float* __restrict__ out) {
  __shared__ unsigned short lds_h1[16 * kHid];
  const unsigned ntiles = (n + 15) / 16;
  const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
  for (unsigned t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const unsigned base = t * 16;
    auto fe = [&](int i, int k) -> float {
      const unsigned s = base + i;
      return s < n ? feat[size_t(s) * kIn + k] : 0.f;
    };
    Fwd f;
    forward_tile(wbf, par, nm, fe, lds_h1, f);
    if (c == 0) {
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const unsigned s = base + 4 * g + r;
        if (s < n) out[s] = nm.out_scale * softplus(f.o[r]);
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------- training
// Loss L = mean_i (softplus(o_i) - t_i/out_scale)^2 over the n samples.
// Each block (one wave) accumulates the gradient of its tiles and writes it
// to partial[block][kNParam] and its loss sum to lossp[block]; a second
// kernel reduces the partials in block order (deterministic).
__global__ void __launch_bounds__(64) k_mlp_grad(unsigned n, const float* __restrict__ feat,
    const float* __restrict__ target, const unsigned short* __restrict__ wbf,
    const float* __restrict__ par, Norm nm, float* __restrict__ partial,
    float* __restrict__ lossp) {
  __shared__ unsigned short lds_h1[16 * kHid];
  __shared__ unsigned short lds_dz2[16 * kHid];
  const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
  const unsigned ntiles = (n + 15) / 16;
  const float invn = 1.f / float(n);
  // Gradient accumulators (D layout for the MFMA-produced blocks).
  f32x4 gw2[4][4], gw1[4];
#pragma unroll
  for (int a = 0; a < 4; a++) {
    gw1[a] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int b = 0; b < 4; b++) gw2[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  }
  float gb1[4] = {0, 0, 0, 0}, gb2[4] = {0, 0, 0, 0}, gw3[4] = {0, 0, 0, 0}, gb3 = 0, loss = 0;
  for (unsigned t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const unsigned base = t * 16;
    auto fe = [&](int i, int k) -> float {
      const unsigned s = base + i;
      return s < n ? feat[size_t(s) * kIn + k] : 0.f;
    };
    Fwd f;
    forward_tile(wbf, par, nm, fe, lds_h1, f);
    // dL/do per row (rows 4g+r); samples beyond n contribute nothing.
    float dout[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const unsigned s = base + 4 * g + r;
      if (s < n) {
        const float y = softplus(f.o[r]), tt = target[s] / nm.out_scale;
        const float e = y - tt;
        if (c == 0) loss += e * e;
        dout[r] = 2.f * e * invn * sigmoid(f.o[r]);
      } else {
        dout[r] = 0.f;
      }
    }
    // Layer 3 and dZ2 (D layout).
    f32x4 dz2[4];
#pragma unroll
    for (int jb = 0; jb < 4; jb++) {
      const float w = bf2f(wbf[kBfW3 + jb * 16 + c]);
      float gw = 0.f, gb = 0.f;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const float z = f.z2[jb][r];
        gw += dout[r] * silu(z);
        const float d = dout[r] * w * dsilu(z);
        dz2[jb][r] = d;
        gb += d;
        lds_dz2[(4 * g + r) * kHid + jb * 16 + c] = f2bf(d);
      }
      gw3[jb] += gw;
      gb2[jb] += gb;
    }
    {
      float d3 = (c == 0) ? dout[0] + dout[1] + dout[2] + dout[3] : 0.f;
      gb3 += across_groups_sum(d3);
    }
    // dW2[j][k] += sum_i dZ2[i][j] H1[i][k]: A = dZ2^T block (D layout as is),
    // B = H1 block (D layout of layer 1 as is, after SiLU).
#pragma unroll
    for (int jb = 0; jb < 4; jb++) {
      const bf16x4 a = pack4(dz2[jb][0], dz2[jb][1], dz2[jb][2], dz2[jb][3]);
#pragma unroll
      for (int kb = 0; kb < 4; kb++) {
        const bf16x4 b = pack4(silu(f.z1[kb][0]), silu(f.z1[kb][1]), silu(f.z1[kb][2]),
                               silu(f.z1[kb][3]));
        gw2[jb][kb] = mfma(a, b, gw2[jb][kb]);
      }
    }
    __syncthreads();
    // dH1 = dZ2 W2: A = dZ2 (samples x units, from LDS), B[j][k] = W2[j][k]
    // read from the transposed copy W2T[k][j].
    bf16x4 adz[4];
#pragma unroll
    for (int jb = 0; jb < 4; jb++) adz[jb] = load_bf4(lds_dz2 + c * kHid + jb * 16 + 4 * g);
    f32x4 dz1[4];
#pragma unroll
    for (int kb = 0; kb < 4; kb++) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int jb = 0; jb < 4; jb++)
        acc = mfma(adz[jb], load_bf4(wbf + kBfW2T + (kb * 16 + c) * kHid + jb * 16 + 4 * g), acc);
      float gb = 0.f;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        acc[r] *= dsilu(f.z1[kb][r]);
        gb += acc[r];
      }
      dz1[kb] = acc;
      gb1[kb] += gb;
    }
    // dW1[j][k] += sum_i dZ1[i][j] X[i][k]: A = dZ1^T block, B[i][k] = X[i][k]
    // (lane: samples 4g..4g+3, feature c).
    {
      float xb[4];
#pragma unroll
      for (int r = 0; r < 4; r++)
        xb[r] = c < kIn ? (fe(4 * g + r, c) - nm.mean[c]) * nm.scale[c] : 0.f;
      const bf16x4 b = pack4(xb[0], xb[1], xb[2], xb[3]);
#pragma unroll
      for (int jb = 0; jb < 4; jb++)
        gw1[jb] = mfma(pack4(dz1[jb][0], dz1[jb][1], dz1[jb][2], dz1[jb][3]), b, gw1[jb]);
    }
    __syncthreads();
  }
  // Write the block's partial gradient.
  float* pg = partial + size_t(blockIdx.x) * kNParam;
#pragma unroll
  for (int jb = 0; jb < 4; jb++) {
#pragma unroll
    for (int r = 0; r < 4; r++) {
      // gw1 block jb: D[j = 4g+r within block][k = c]
      pg[kOffW1 + (jb * 16 + 4 * g + r) * kInPad + c] = c < kIn ? gw1[jb][r] : 0.f;
#pragma unroll
      for (int kb = 0; kb < 4; kb++)
        pg[kOffW2 + (jb * 16 + 4 * g + r) * kHid + kb * 16 + c] = gw2[jb][kb][r];
    }
  }
#pragma unroll
  for (int jb = 0; jb < 4; jb++) {
    const float b1 = across_groups_sum(gb1[jb]);
    const float b2 = across_groups_sum(gb2[jb]);
    const float w3 = across_groups_sum(gw3[jb]);
    if (g == 0) {
      pg[kOffB1 + jb * 16 + c] = b1;
      pg[kOffB2 + jb * 16 + c] = b2;
      pg[kOffW3 + jb * 16 + c] = w3;
    }
  }
  if (lane == 0) pg[kOffB3] = gb3;
  const float l = across_groups_sum(loss);
  if (lane == 0) lossp[blockIdx.x] = l * invn;
}

// Sums the block partials in block order.
__global__ void __launch_bounds__(256) k_mlp_reduce(unsigned nblocks,
    const float* __restrict__ partial, const float* __restrict__ lossp, float* __restrict__ grad,
    float* __restrict__ loss) {
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < unsigned(kNParam)) {
    float s = 0.f;
    for (unsigned b = 0; b < nblocks; b++) s += partial[size_t(b) * kNParam + i];
    grad[i] = s;
  }
  if (i == 0 && loss) {
    float s = 0.f;
    for (unsigned b = 0; b < nblocks; b++) s += lossp[b];
    *loss = s;
  }
}

// Adam on the FP32 master parameters (Kingma & Ba 2015), then the BF16 copies.
__global__ void __launch_bounds__(256) k_mlp_adam(float* __restrict__ par,
    float* __restrict__ m1, float* __restrict__ m2, const float* __restrict__ grad, float lr,
    float beta1, float beta2, float eps, float c1, float c2) {
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= unsigned(kNParam)) return;
  float gr = grad[i];
  if (i < unsigned(kOffB1) && (i % kInPad) >= unsigned(kIn)) gr = 0.f;  // padding columns
  const float a = beta1 * m1[i] + (1.f - beta1) * gr;
  const float b = beta2 * m2[i] + (1.f - beta2) * gr * gr;
  m1[i] = a;
  m2[i] = b;
  par[i] -= lr * (a * c1) / (sqrtf(b * c2) + eps);
}

__global__ void __launch_bounds__(256) k_mlp_to_bf16(const float* __restrict__ par,
                                                     unsigned short* __restrict__ wbf) {
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < unsigned(kHid * kInPad)) {
    const unsigned k = i % kInPad;
    wbf[kBfW1 + i] = k < unsigned(kIn) ? f2bf(par[kOffW1 + i]) : 0;
  }
  if (i < unsigned(kHid * kHid)) {
    const unsigned j = i / kHid, k = i % kHid;
    const unsigned short v = f2bf(par[kOffW2 + i]);
    wbf[kBfW2 + i] = v;
    wbf[kBfW2T + k * kHid + j] = v;
  }
  if (i < unsigned(kHid)) wbf[kBfW3 + i] = f2bf(par[kOffW3 + i]);
}

// ---------------------------------------------------------------- closure
// SPS tau of fluid particles with the learned eddy viscosity: features are
// the velocity gradient G[9] of the interaction, rho and dp; nu_t replaces
// smag*sqrt(prr) of KerComputeSpsTau (JSphGpu_ker.cu:1387-1436); the
// isotropic blin term is kept.
__global__ void __launch_bounds__(64) k_sps_tau_learned(unsigned n, unsigned pini, float blin,
    float dp, const float4* __restrict__ velrho, const float* __restrict__ velgrad,
    const float* __restrict__ s2, const unsigned short* __restrict__ wbf,
    const float* __restrict__ par, Norm nm, float* __restrict__ tau, float* __restrict__ nu_out) {
  __shared__ unsigned short lds_h1[16 * kHid];
  const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
  const unsigned ntiles = (n + 15) / 16;
  for (unsigned t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const unsigned base = t * 16;
    auto fe = [&](int i, int k) -> float {
      const unsigned s = base + i;
      if (s >= n) return 0.f;
      const unsigned p = s + pini;
      if (k < 9) return velgrad[9 * size_t(p) + k];
      return k == 9 ? velrho[p].w : dp;
    };
    Fwd f;
    forward_tile(wbf, par, nm, fe, lds_h1, f);
    if (c == 0) {
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const unsigned s = base + 4 * g + r;
        if (s >= n) continue;
        const unsigned p = s + pini;
        const float visc_sps = nm.out_scale * softplus(f.o[r]);
        const float* sp = s2 + 6 * size_t(p);
        const float xx = sp[0], xy = sp[1], xz = sp[2], yy = sp[3], yz = sp[4], zz = sp[5];
        const float pow1 = xx * xx + yy * yy + zz * zz;
        const float prr = xy * xy + xz * xz + yz * yz + pow1 + pow1;
        const float div_u = xx + yy + zz;
        const float sumsps = -(visc_sps * div_u + blin * prr);
        const float two = visc_sps + visc_sps;
        const float one_rho = 1.0f / velrho[p].w;
        float* tp = tau + 6 * size_t(p);
        tp[0] = one_rho * (two * xx + sumsps);
        tp[1] = one_rho * (visc_sps * xy);
        tp[2] = one_rho * (visc_sps * xz);
        tp[3] = one_rho * (two * yy + sumsps);
        tp[4] = one_rho * (visc_sps * yz);
        tp[5] = one_rho * (two * zz + sumsps);
        if (nu_out) nu_out[p] = visc_sps;
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------- training data
// Shepard filter of a fine-resolution state at probe points with the
// Wendland kernel of smoothing length hf (support 2hf): per probe
//   out[0]      sum V W                 (V = m/rho)
//   out[1..3]   sum V W u
//   out[4..9]   sum V W u_a u_b  {xx,xy,xz,yy,yz,zz}
//   out[10]     sum V W rho
//   out[11..19] sum V u_a dW/dx_b       (row-major a,b)
//   out[20..22] sum V dW/dx_b
// in FP64, over the fluid particles of the cell rows within 2hf. One block
// per probe.
__global__ void __launch_bounds__(256) k_filter_probes(unsigned nprobe, const double* __restrict__ probes,
    double hf, int reach, DivData dvd, float massf, const double2* __restrict__ posxy,
    const double* __restrict__ posz, const float4* __restrict__ velrho,
    const typecode* __restrict__ code, double* __restrict__ out) {
  __shared__ double red[256];
  const unsigned pr = blockIdx.x;
  if (pr >= nprobe) return;
  const double px = probes[3 * size_t(pr)], py = probes[3 * size_t(pr) + 1],
               pz = probes[3 * size_t(pr) + 2];
  const double r2max = 4.0 * hf * hf;
  const double awen = 21.0 / (16.0 * 3.14159265358979323846 * hf * hf * hf);
  double acc[23];
  for (int k = 0; k < 23; k++) acc[k] = 0;
  const int cx = int((px - dvd.domposminx) / dvd.scell) - dvd.czx;
  const int cy = int((py - dvd.domposminy) / dvd.scell) - dvd.czy;
  const int cz = int((pz - dvd.domposminz) / dvd.scell) - dvd.czz;
  const int x0 = max(cx - reach, 0), x1 = min(cx + reach, dvd.ncx - 1) + 1;
  const int y0 = max(cy - reach, 0), y1 = min(cy + reach, dvd.ncy - 1) + 1;
  const int z0 = max(cz - reach, 0), z1 = min(cz + reach, dvd.ncz - 1) + 1;
  for (int z = z0; z < z1; z++)
    for (int y = y0; y < y1; y++) {
      unsigned pini, pfin;
      if (x0 >= x1) continue;
      row_range(dvd.begincell, int(dvd.cellfluid) + z * dvd.ncw + y * dvd.ncx, x0, x1, pini, pfin);
      for (unsigned p = pini + threadIdx.x; p < pfin; p += blockDim.x) {
        if (!code_is_fluid(code[p]) || !code_is_normal(code[p])) continue;
        const double2 xy = posxy[p];
        const double dx = px - xy.x, dy = py - xy.y, dz = pz - posz[p];
        const double rr2 = dx * dx + dy * dy + dz * dz;
        if (rr2 >= r2max) continue;
        const double q = sqrt(rr2) / hf;
        const double w1 = 1.0 - 0.5 * q;
        const double w = awen * (2.0 * q + 1.0) * w1 * w1 * w1 * w1;
        const double fac = -5.0 * awen / (hf * hf) * w1 * w1 * w1;  // (dW/dr)/r
        const float4 vr = velrho[p];
        const double vol = double(massf) / double(vr.w);
        const double u[3] = {vr.x, vr.y, vr.z};
        const double gw[3] = {fac * dx, fac * dy, fac * dz};
        acc[0] += vol * w;
        for (int a = 0; a < 3; a++) acc[1 + a] += vol * w * u[a];
        acc[4] += vol * w * u[0] * u[0]; acc[5] += vol * w * u[0] * u[1];
        acc[6] += vol * w * u[0] * u[2]; acc[7] += vol * w * u[1] * u[1];
        acc[8] += vol * w * u[1] * u[2]; acc[9] += vol * w * u[2] * u[2];
        acc[10] += vol * w * double(vr.w);
        for (int a = 0; a < 3; a++)
          for (int b = 0; b < 3; b++) acc[11 + 3 * a + b] += vol * u[a] * gw[b];
        for (int b = 0; b < 3; b++) acc[20 + b] += vol * gw[b];
      }
    }
  for (int k = 0; k < 23; k++) {
    red[threadIdx.x] = acc[k];
    __syncthreads();
    for (unsigned s = blockDim.x / 2; s > 0; s >>= 1) {
      if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[23 * size_t(pr) + k] = red[0];
    __syncthreads();
  }
}

}  // namespace

extern "C" {

int dsph_mlp_param_count(void) { return kNParam; }
int dsph_mlp_bf16_count(void) { return kNBf; }

int dsph_mlp_sync(dsph_ctx* ctx, const dsph_mlp* m) {
  if (!ctx || !m) return DSPH_E_ARG;
  DSPH_REQUIRE(ctx, m->d_params && m->d_wbf16, "dsph_mlp_sync: null buffers");
  k_mlp_to_bf16<<<grid_for(kHid * kHid, 256), 256, 0, ctx->stream>>>(m->d_params, m->d_wbf16);
  DSPH_LAUNCHED(ctx, "k_mlp_to_bf16");
  return DSPH_OK;
}

int dsph_mlp_forward(dsph_ctx* ctx, const dsph_mlp* m, uint32_t n, const float* d_features,
                     float* d_out) {
  if (!ctx || !m) return DSPH_E_ARG;
  if (!n) return DSPH_OK;
  DSPH_REQUIRE(ctx, m->d_params && m->d_wbf16 && d_features && d_out,
               "dsph_mlp_forward: null argument");
  const unsigned ntiles = (n + 15) / 16;
  KTimer kt(ctx, "k_mlp_forward");
  k_mlp_forward<<<std::min(ntiles, 8192u), 64, 0, ctx->stream>>>(n, d_features, m->d_wbf16,
                                                                 m->d_params, make_norm(m), d_out);
  DSPH_LAUNCHED(ctx, "k_mlp_forward");
  return DSPH_OK;
}

int dsph_mlp_grad(dsph_ctx* ctx, const dsph_mlp* m, uint32_t n, const float* d_features,
                  const float* d_targets, float* d_scratch, size_t scratch_bytes,
                  float* d_grad, float* d_loss) {
  if (!ctx || !m) return DSPH_E_ARG;
  DSPH_REQUIRE(ctx, n > 0 && m->d_params && m->d_wbf16 && d_features && d_targets && d_scratch &&
                        d_grad,
               "dsph_mlp_grad: null argument");
  const unsigned ntiles = (n + 15) / 16;
  unsigned nblocks = std::min(ntiles, 1024u);
  const size_t per = sizeof(float) * (size_t(kNParam) + 1);
  if (scratch_bytes / per < nblocks) nblocks = unsigned(scratch_bytes / per);
  DSPH_REQUIRE(ctx, nblocks > 0, "dsph_mlp_grad: scratch too small (need >= 21256 bytes)");
  float* partial = d_scratch;
  float* lossp = d_scratch + size_t(nblocks) * kNParam;
  KTimer kt(ctx, "k_mlp_grad");
  k_mlp_grad<<<nblocks, 64, 0, ctx->stream>>>(n, d_features, d_targets, m->d_wbf16, m->d_params,
                                              make_norm(m), partial, lossp);
  DSPH_LAUNCHED(ctx, "k_mlp_grad");
  k_mlp_reduce<<<grid_for(kNParam, 256), 256, 0, ctx->stream>>>(nblocks, partial, lossp, d_grad,
                                                                 d_loss);
  DSPH_LAUNCHED(ctx, "k_mlp_reduce");
  return DSPH_OK;
}

int dsph_mlp_adam(dsph_ctx* ctx, const dsph_mlp* m, const float* d_grad, float lr, float beta1,
                  float beta2, float eps, uint32_t step) {
  if (!ctx || !m) return DSPH_E_ARG;
  DSPH_REQUIRE(ctx, m->d_params && m->d_adam_m && m->d_adam_v && d_grad && step > 0,
               "dsph_mlp_adam: null argument or step 0");
  const float c1 = float(1.0 / (1.0 - std::pow(double(beta1), double(step))));
  const float c2 = float(1.0 / (1.0 - std::pow(double(beta2), double(step))));
  k_mlp_adam<<<grid_for(kNParam, 256), 256, 0, ctx->stream>>>(
      m->d_params, m->d_adam_m, m->d_adam_v, d_grad, lr, beta1, beta2, eps, c1, c2);
  DSPH_LAUNCHED(ctx, "k_mlp_adam");
  return dsph_mlp_sync(ctx, m);
}

int dsph_sps_tau_learned(dsph_ctx* ctx, const dsph_mlp* m, uint32_t np, uint32_t npb,
                         float blin, float dp, const float* d_velrho, const float* d_velgrad,
                         const float* d_sps2strain, float* d_tau, float* d_nu) {
  if (!ctx || !m) return DSPH_E_ARG;
  if (np <= npb) return DSPH_OK;
  DSPH_REQUIRE(ctx, m->d_params && m->d_wbf16 && d_velrho && d_velgrad && d_sps2strain && d_tau,
               "dsph_sps_tau_learned: null argument");
  const unsigned n = np - npb, ntiles = (n + 15) / 16;
  KTimer kt(ctx, "k_sps_tau_learned");
  k_sps_tau_learned<<<std::min(ntiles, 8192u), 64, 0, ctx->stream>>>(
      n, npb, blin, dp, (const float4*)d_velrho, d_velgrad, d_sps2strain, m->d_wbf16,
      m->d_params, make_norm(m), d_tau, d_nu);
  DSPH_LAUNCHED(ctx, "k_sps_tau_learned");
  return DSPH_OK;
}

int dsph_filter_probes(dsph_ctx* ctx, uint32_t nprobe, const double* d_probes, double hf,
                       const dsph_divdata* dvd, float massf, const double* d_posxy,
                       const double* d_posz, const float* d_velrho, const uint16_t* d_code,
                       double* d_out) {
  if (!ctx || !dvd) return DSPH_E_ARG;
  if (!nprobe) return DSPH_OK;
  DSPH_REQUIRE(ctx, d_probes && d_posxy && d_posz && d_velrho && d_code && d_out && hf > 0,
               "dsph_filter_probes: invalid argument");
  const DivData d = make_divdata(*dvd);
  const int reach = int(std::ceil(2.0 * hf / double(d.scell)));
  KTimer kt(ctx, "k_filter_probes");
  k_filter_probes<<<nprobe, 256, 0, ctx->stream>>>(nprobe, d_probes, hf, reach, d, massf,
                                                   (const double2*)d_posxy, d_posz,
                                                   (const float4*)d_velrho, d_code, d_out);
  DSPH_LAUNCHED(ctx, "k_filter_probes");
  return DSPH_OK;
}

}  // extern "C"
