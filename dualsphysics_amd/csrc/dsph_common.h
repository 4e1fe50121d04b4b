// Device-side definitions shared by the HIP kernels of the SPH step.
//
// Bit layouts are part of the drop-in contract and keep the reference's values:
//  - particle code (u16): DualSphDef.h:150-241 (3 special bits, 2 type bits,
//    11 mk bits);
//  - dcell (u32): JDsDcellDef.h:28-45, a per-domain split of 31 bits between
//    cx, cy, cz described by the 32-bit DomCellCode;
//  - PosCell (u32 stored in float4.w): DualSphDef.h:661-700, 13/10/9 bits.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "dsph_abi.h"

namespace dsph {

// ---- particle code bits (typecode = u16) ----
typedef uint16_t typecode;
constexpr unsigned kCodeMaskSpecial = 0xe000u;
constexpr unsigned kCodeNormal = 0x0000u;
constexpr unsigned kCodePeriodic = 0x2000u;
constexpr unsigned kCodeOutIgnore = 0x4000u;
constexpr unsigned kCodeOutMov = 0x6000u;
constexpr unsigned kCodeOutPos = 0x8000u;
constexpr unsigned kCodeOutRho = 0xa000u;
constexpr unsigned kCodeMaskType = 0x1800u;
constexpr unsigned kCodeTypeFixed = 0x0000u;
constexpr unsigned kCodeTypeMoving = 0x0800u;
constexpr unsigned kCodeTypeFloating = 0x1000u;
constexpr unsigned kCodeMaskValue = 0x07ffu;   // CODE_MASKVALUE: the object of its type
constexpr unsigned kCodeTypeFluid = 0x1800u;

__host__ __device__ inline unsigned code_special(unsigned c) { return c & kCodeMaskSpecial; }
__host__ __device__ inline unsigned code_type(unsigned c) { return c & kCodeMaskType; }
__host__ __device__ inline bool code_is_normal(unsigned c) { return code_special(c) == kCodeNormal; }
__host__ __device__ inline bool code_is_fluid(unsigned c) { return code_type(c) == kCodeTypeFluid; }
__host__ __device__ inline bool code_is_outrho(unsigned c) { return code_special(c) == kCodeOutRho; }
__host__ __device__ inline unsigned code_set(unsigned c, unsigned special) {
  return (c & ~kCodeMaskSpecial) | special;
}

// ---- dcell coding (31 bits split by DomCellCode) ----
constexpr unsigned kDcelMapOut = 0xffffffffu;
__host__ __device__ inline unsigned dcel_cellx(unsigned dcc, unsigned cel) {
  return cel >> ((dcc >> 10) & 31u);
}
__host__ __device__ inline unsigned dcel_celly(unsigned dcc, unsigned cel) {
  return (cel << (dcc >> 25)) >> ((dcc >> 5) & 31u);
}
__host__ __device__ inline unsigned dcel_cellz(unsigned dcc, unsigned cel) {
  return (cel << (dcc & 31u)) >> (dcc & 31u);
}
__host__ __device__ inline unsigned dcel_cell(unsigned dcc, unsigned cx, unsigned cy, unsigned cz) {
  return (cx << ((dcc >> 10) & 31u)) | (cy << ((dcc >> 15) & 31u)) | cz;
}

// ---- PosCell coding (13/10/9 bits) ----
constexpr unsigned kPscelMovX = 19u;
constexpr unsigned kPscelMovY = 9u;
constexpr unsigned kPscelMaskY = 0x0007fe00u;
constexpr unsigned kPscelMaskZ = 0x000001ffu;
__host__ __device__ inline unsigned pscel_code(unsigned cx, unsigned cy, unsigned cz) {
  return (cx << kPscelMovX) | (cy << kPscelMovY) | cz;
}
__device__ inline int pscel_fx(float w) { return int(__float_as_uint(w) >> kPscelMovX); }
__device__ inline int pscel_fy(float w) { return int((__float_as_uint(w) & kPscelMaskY) >> kPscelMovY); }
__device__ inline int pscel_fz(float w) { return int(__float_as_uint(w) & kPscelMaskZ); }

constexpr float kAlmostZero = 1e-18f;  // ALMOSTZERO (DualSphDef.h:114)

// Cell-search view of dsph_divdata passed by value to kernels.
struct DivData {
  int scelldiv;
  int ncx, ncy, ncz, ncw;
  unsigned cellfluid;
  int czx, czy, czz;
  const int2* begincell;
  unsigned domcellcode;
  float kernelsize2;
  float poscellsize;
  double domposminx, domposminy, domposminz;
  float scell;
};

inline DivData make_divdata(const dsph_divdata& d) {
  DivData r;
  r.scelldiv = d.scelldiv;
  r.ncx = d.ncx; r.ncy = d.ncy; r.ncz = d.ncz; r.ncw = d.ncw;
  r.cellfluid = d.cellfluid;
  r.czx = d.cellzerox; r.czy = d.cellzeroy; r.czz = d.cellzeroz;
  r.begincell = reinterpret_cast<const int2*>(d.d_beginendcell);
  r.domcellcode = d.domcellcode;
  r.kernelsize2 = d.kernelsize2;
  r.poscellsize = d.poscellsize;
  r.domposminx = d.domposminx; r.domposminy = d.domposminy; r.domposminz = d.domposminz;
  r.scell = d.scell;
  return r;
}

// Neighbour-search limits for a particle in cell (cx,cy,cz) of the current
// domain (cell coordinates already relative to CellDomainMin). Rows are
// x-contiguous ranges of cells [x0,x1) for y in [y0,y1), z in [z0,z1); the
// same clamping as cunsearch::InitCte (JCellSearch_iker.h:34-58).
struct NgLimits { int x0, x1, y0, y1, z0, z1; };

__device__ inline NgLimits ng_limits(int cx, int cy, int cz, const DivData& d) {
  NgLimits l;
  const int s = d.scelldiv;
  l.x0 = cx - min(cx, s); l.x1 = cx + min(d.ncx - cx - 1, s) + 1;
  l.y0 = cy - min(cy, s); l.y1 = cy + min(d.ncy - cy - 1, s) + 1;
  l.z0 = cz - min(cz, s); l.z1 = cz + min(d.ncz - cz - 1, s) + 1;
  return l;
}

// Contiguous particle range [pini,pfin) of one cell row: first begin and last
// end over the non-empty cells c0..c1-1 of the row starting at cell index
// `rowbase` (same result as cunsearch::ParticleRange, JCellSearch_iker.h:155).
__device__ inline void row_range(const int2* bc, int rowbase, int x0, int x1,
                                 unsigned& pini, unsigned& pfin) {
  pini = 0; pfin = 0;
  for (int c = x0; c < x1; c++) {
    const int2 be = bc[rowbase + c];
    if (be.y) {
      if (!pfin) pini = unsigned(be.x);
      pfin = unsigned(be.y);
    }
  }
}

// Relative distance p1-p2 from cell-relative FP32 positions
// (KerInteractionForcesFluidBox, JSphGpu_ker.cu:559-561).
__device__ inline float3 pscel_dr(const float4& a, const float4& b, float pcs) {
  float3 r;
  r.x = a.x - b.x + pcs * float(pscel_fx(a.w) - pscel_fx(b.w));
  r.y = a.y - b.y + pcs * float(pscel_fy(a.w) - pscel_fy(b.w));
  r.z = a.z - b.z + pcs * float(pscel_fz(a.w) - pscel_fz(b.w));
  return r;
}

}  // namespace dsph
