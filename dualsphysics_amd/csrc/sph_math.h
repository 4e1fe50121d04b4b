// SPH device math shared by the interaction kernels: constants view, Tait EOS,
// Wendland C2 kernel, and the Fourtakas hydrostatic density term.
#pragma once
#include <cfloat>
#include "dsph_ctx.h"

namespace dsph {

struct Cte {
  float massb, massf, h, invh, ks2, pcs, awen, bwenh, cs0, eta2, ddtkh, ddtgz, rhopzero, invgamma,
      cteb, gamma, ovrhopzero, dp;
  unsigned cellcode;
  unsigned tboundary;  // DSPH_BC_*: the Molteni-Colagrossi term skips boundaries only with DBC
  unsigned sim2d;      // 2-D simulation: the mDBC systems are 3x3 over (1, x, z)
  // Binomial-series coefficients of (1+u)^(1/gamma), k = 1..5.
  float pc1, pc2, pc3, pc4, pc5;
};

inline Cte make_cte(const dsph_cte& c) {
  Cte r;
  r.massb = c.massb; r.massf = c.massf; r.h = c.kernelh; r.invh = 1.f / c.kernelh; r.ks2 = c.kernelsize2;
  r.pcs = c.poscellsize; r.awen = c.awen; r.bwenh = c.bwenh; r.cs0 = c.cs0; r.eta2 = c.eta2;
  r.ddtkh = c.ddtkh; r.ddtgz = c.ddtgz; r.rhopzero = c.rhopzero; r.invgamma = 1.f / c.gamma;
  r.cteb = c.cteb; r.gamma = c.gamma; r.ovrhopzero = c.ovrhopzero; r.dp = c.dp;
  r.cellcode = c.cellcode;
  r.tboundary = c.tboundary;
  r.sim2d = c.simulate2d;
  // binom(a,k) = binom(a,k-1) * (a-k+1)/k with a = invgamma (as the reference's 1.f/gamma).
  const double a = double(r.invgamma);
  double b = 1.0;
  float* pc[5] = {&r.pc1, &r.pc2, &r.pc3, &r.pc4, &r.pc5};
  for (int k = 1; k <= 5; k++) {
    b = b * (a - (k - 1)) / k;
    *pc[k - 1] = float(b);
  }
  return r;
}

// x^7 in FP64 rounded once to FP32: x^2 is exact in FP64 and x^3, x^4, x^7
// round once each, so the FP64 value is within 3 ulp(FP64) of x^7 and its
// FP32 rounding is the correctly rounded power except when x^7 lies within
// that distance of an FP32 rounding boundary (about 6e-9 of the ratios).
__device__ inline float pow7_f64(float x) {
  const double xd = double(x);
  const double x2 = xd * xd;
  return float((x2 * x2) * (x2 * xd));
}

// Tait EOS (FunSphEos_iker.h:37-59): B*((rho*(1/rho0))^gamma - 1). The power
// is evaluated in FP64 and rounded once, i.e. the correctly rounded FP32
// power of the FP32 ratio (the reference's powf is platform dependent; its
// correctly rounded value is the common target of the kernels and the oracle);
// for gamma = 7 by four FP64 products (pow7_f64), cheap enough for the fluid
// kernel to evaluate per pair instead of loading the neighbour's pressure.
__device__ inline float eos_press(float rho, const Cte& c) {
  const float x = rho * c.ovrhopzero;
  const float pw = c.gamma == 7.f ? pow7_f64(x) : float(pow(double(x), double(c.gamma)));
  return c.cteb * (pw - 1.0f);
}

// (1+u)^(1/gamma) for the Fourtakas term rh = 1 + ddtgz*dz (JSphGpu_ker.cu:668-674),
// |u| <= ddtgz*2h ~ 1e-3: binomial series to u^4 (truncation < 1e-11 for
// |u| < 1e-2, below the rounding of the FP32 sum), rounded once at 1+s, which
// is the correctly rounded power except when the exact value lies within
// ~1e-10 of a rounding boundary. Outside |u| < 1e-2 the library powf is used.
__device__ inline float pow_invgamma(float rh, const Cte& c) {
  const float u = rh - 1.f;
  if (fabsf(u) < 1e-2f) {
    const float s = u * (c.pc1 + u * (c.pc2 + u * (c.pc3 + u * c.pc4)));
    return 1.f + s;
  }
  return powf(rh, c.invgamma);
}

// Quotient and square root of the pair arithmetic. The reference's GPU build
// is compiled with -use_fast_math (Makefile:186-187), i.e. approximate
// reciprocal-based division and square root; here the same: v_rcp_f32 and
// v_sqrt_f32 (1 ulp each) instead of the correctly rounded sequences.
__device__ inline float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ inline float fast_div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ inline float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// Wendland C2 gradient factor (FunSphKernel_iker.h:117-124).
__device__ inline float wendland_fac(float rr2, const Cte& c) {
  const float rad = fast_sqrt(rr2);
  const float qq = rad * c.invh;
  const float wqq1 = 1.f - 0.5f * qq;
  return c.bwenh * wqq1 * wqq1 * wqq1;
}
// Wendland C2 kernel value and gradient factor (FunSphKernel_iker.h:106-140).
__device__ inline float wendland_wab_fac(float rr2, const Cte& c, float& fac) {
  const float rad = fast_sqrt(rr2);
  const float qq = rad * c.invh;
  const float wqq1 = 1.f - 0.5f * qq;
  const float wqq2 = wqq1 * wqq1;
  fac = c.bwenh * wqq2 * wqq1;
  const float wqq = qq + qq + 1.f;
  return c.awen * wqq * wqq2 * wqq2;
}

}  // namespace dsph
