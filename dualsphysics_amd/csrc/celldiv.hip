// Cell-linked neighbour list on gfx950: cell limits, cell keys, stable radix
// sort by cell, BeginEndCell ranges, reorder of particle arrays, PosCell.
//
// Semantics follow the reference's single-GPU divide (JCellDivGpuSingle.cpp:
// 152-236, JCellDivGpuSingle_ker.cu:41-165, JCellDivGpu_ker.cu:120-127,
// 436-489, 533-712) with the reference's -stable ordering: equal keys keep
// their current array order, so the permutation equals the CPU counting sort
// (JCellDivCpuSingle.cpp:207-262).
#include <algorithm>
#include <climits>
#include <cstring>
#include <string>
#include "dsph_ctx.h"
#include "ensemble.h"

using namespace dsph;

namespace {

constexpr unsigned kBs = 256;

// ------------------------------------------------------------- cell limits
// Min/max cell (x,y,z) over non-excluded particles (exact and order
// independent): each block writes its six extremes to part[6 block + k]
// (k 0..2 min, 3..5 max), one block combines them into res[0..5]. Partials
// instead of global atomics: thousands of atomics on six words serialise in
// L2 and cost ten times the 6 bytes per particle that are read.
// Block extremes: on return, threads 0..2 hold min (a) and max (b) of axis
// threadIdx.x over the block.
__device__ inline void limits_block_reduce(unsigned mn[3], unsigned mx[3], unsigned& a,
                                           unsigned& b) {
  __shared__ unsigned smin[3][kBs / 64], smax[3][kBs / 64];
  for (int k = 0; k < 3; k++) {
    for (int off = 32; off > 0; off >>= 1) {
      mn[k] = min(mn[k], (unsigned)__shfl_xor((int)mn[k], off));
      mx[k] = max(mx[k], (unsigned)__shfl_xor((int)mx[k], off));
    }
  }
  const unsigned w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0)
    for (int k = 0; k < 3; k++) { smin[k][w] = mn[k]; smax[k][w] = mx[k]; }
  __syncthreads();
  a = UINT_MAX;
  b = 0;
  if (threadIdx.x < 3)
    for (unsigned i = 0; i < blockDim.x / 64; i++) {
      a = min(a, smin[threadIdx.x][i]);
      b = max(b, smax[threadIdx.x][i]);
    }
}

__global__ void __launch_bounds__(kBs) k_limits_final(unsigned nparts,
                                                      const unsigned* __restrict__ part,
                                                      unsigned* __restrict__ res) {
  unsigned mn[3] = {UINT_MAX, UINT_MAX, UINT_MAX}, mx[3] = {0u, 0u, 0u};
  for (unsigned b = threadIdx.x; b < nparts; b += blockDim.x)
    for (int k = 0; k < 3; k++) {
      mn[k] = min(mn[k], part[6 * b + k]);
      mx[k] = max(mx[k], part[6 * b + 3 + k]);
    }
  unsigned a, b;
  limits_block_reduce(mn, mx, a, b);
  if (threadIdx.x < 3) {
    res[threadIdx.x] = a;
    res[3 + threadIdx.x] = b;
  }
}

__global__ void __launch_bounds__(kBs) k_limits(unsigned n, unsigned pini, unsigned cellcode,
                                                const unsigned* __restrict__ dcell,
                                                const typecode* __restrict__ code,
                                                unsigned* __restrict__ part) {
  unsigned mn[3] = {UINT_MAX, UINT_MAX, UINT_MAX}, mx[3] = {0u, 0u, 0u};
  for (unsigned p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
    const unsigned pp = p + pini;
    if (code_special(code[pp]) < kCodeOutIgnore) {
      const unsigned c = dcell[pp];
      const unsigned cx = dcel_cellx(cellcode, c), cy = dcel_celly(cellcode, c),
                     cz = dcel_cellz(cellcode, c);
      mn[0] = min(mn[0], cx); mn[1] = min(mn[1], cy); mn[2] = min(mn[2], cz);
      mx[0] = max(mx[0], cx); mx[1] = max(mx[1], cy); mx[2] = max(mx[2], cz);
    }
  }
  unsigned a, b;
  limits_block_reduce(mn, mx, a, b);
  if (threadIdx.x < 3) {
    part[6 * blockIdx.x + threadIdx.x] = a;
    part[6 * blockIdx.x + 3 + threadIdx.x] = b;
  }
}

// ------------------------------------------------------------- cell keys
// Box layout of JCellDivGpuSingle::PrepareNct (JCellDivGpuSingle.cpp:109-128):
// [bound cells 0..nct) | BoundIgnore=nct | fluid cells nct+1.. | BoundOut |
// FluidOut | BoundOutIgnore | FluidOutIgnore.
__device__ inline unsigned presort_full_box(unsigned p, unsigned cellcode,
    const unsigned* __restrict__ dcell, const typecode* __restrict__ code, uint3 cellzero,
    uint3 ncells) {
  const unsigned nsheet = ncells.x * ncells.y;
  const unsigned boxboundignore = nsheet * ncells.z;
  const unsigned boxfluid = boxboundignore + 1;
  const unsigned boxboundout = boxfluid + boxboundignore;
  const unsigned boxfluidout = boxboundout + 1;
  const unsigned boxboundoutignore = boxfluidout + 1;
  const unsigned boxfluidoutignore = boxboundoutignore + 1;
  const unsigned c = dcell[p];
  const unsigned cx = dcel_cellx(cellcode, c) - cellzero.x;
  const unsigned cy = dcel_celly(cellcode, c) - cellzero.y;
  const unsigned cz = dcel_cellz(cellcode, c) - cellzero.z;
  const unsigned key = cx + cy * ncells.x + cz * nsheet;
  const unsigned rc = code[p];
  const unsigned type = code_type(rc), out = code_special(rc);
  unsigned box;
  if (type < kCodeTypeFloating) {
    box = (out < kCodeOutIgnore)
              ? ((cx < ncells.x && cy < ncells.y && cz < ncells.z) ? key : boxboundignore)
              : (out == kCodeOutIgnore ? boxboundoutignore : boxboundout);
  } else {
    box = (out <= kCodeOutIgnore) ? (out < kCodeOutIgnore ? boxfluid + key : boxfluidoutignore)
                                  : (type == kCodeTypeFloating ? boxboundout : boxfluidout);
  }
  return box;
}

__global__ void __launch_bounds__(kBs) k_presort_full(unsigned np, unsigned cellcode,
    const unsigned* __restrict__ dcell, const typecode* __restrict__ code,
    uint3 cellzero, uint3 ncells, unsigned* __restrict__ cellpart,
    unsigned* __restrict__ sortpart) {
  const unsigned p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  sortpart[p] = p;
  cellpart[p] = presort_full_box(p, cellcode, dcell, code, cellzero, ncells);
}

__global__ void __launch_bounds__(kBs) k_presort_fluid(unsigned n, unsigned pini,
    unsigned cellcode, const unsigned* __restrict__ dcell,
    const typecode* __restrict__ code, uint3 cellzero, uint3 ncells,
    unsigned* __restrict__ cellpart, unsigned* __restrict__ sortpart) {
  unsigned p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  p += pini;
  sortpart[p] = p;
  const unsigned nsheet = ncells.x * ncells.y;
  const unsigned boxfluid = nsheet * ncells.z + 1;
  const unsigned boxfluidout = boxfluid + boxfluid;
  const unsigned boxboundout = boxfluidout - 1;
  const unsigned boxfluidoutignore = boxfluidout + 2;
  const unsigned c = dcell[p];
  const unsigned cx = dcel_cellx(cellcode, c) - cellzero.x;
  const unsigned cy = dcel_celly(cellcode, c) - cellzero.y;
  const unsigned cz = dcel_cellz(cellcode, c) - cellzero.z;
  const unsigned rc = code[p];
  const unsigned type = code_type(rc), out = code_special(rc);
  cellpart[p] = (out <= kCodeOutIgnore)
                    ? (out < kCodeOutIgnore ? boxfluid + cx + cy * ncells.x + cz * nsheet
                                            : boxfluidoutignore)
                    : (type == kCodeTypeFloating ? boxboundout : boxfluidout);
}

// ------------------------------------------------------------- radix sort
// LSD radix sort of (key,value) pairs, 8-bit digits, tiles of 4096 keys per
// 256-thread workgroup (4 waves x 1024 keys). Stability inside a tile comes
// from ranking keys in tile order: each wave walks its 1024-key segment in 16
// rounds of 64 contiguous keys, finds the lanes sharing its digit with one
// ballot per digit bit, and keeps per-wave digit counters in LDS. A per-digit
// scan over waves and over tiles (digit-major histogram) gives each key its
// global slot.
constexpr unsigned kTileItems = 16;
constexpr unsigned kTile = kBs * kTileItems;  // 4096

__global__ void __launch_bounds__(kBs) k_radix_hist(const unsigned* __restrict__ keys,
    unsigned n, unsigned shift, unsigned mask, unsigned ntiles, unsigned* __restrict__ hist) {
  __shared__ unsigned cnt[256];
  const unsigned tile = blockIdx.x;
  cnt[threadIdx.x] = 0;
  __syncthreads();
  const unsigned base = tile * kTile;
#pragma unroll
  for (unsigned i = 0; i < kTileItems; i++) {
    const unsigned idx = base + i * kBs + threadIdx.x;
    if (idx < n) atomicAdd(&cnt[(keys[idx] >> shift) & mask], 1u);
  }
  __syncthreads();
  if (threadIdx.x <= mask) hist[threadIdx.x * ntiles + tile] = cnt[threadIdx.x];
}

// Exclusive scan over tiles of one digit (one workgroup per digit); digit
// totals go to totals[digit].
__global__ void __launch_bounds__(kBs) k_radix_scan_digit(unsigned* __restrict__ hist,
    unsigned ntiles, unsigned* __restrict__ totals) {
  __shared__ unsigned wsum[kBs / 64];
  unsigned* h = hist + size_t(blockIdx.x) * ntiles;
  const unsigned per = (ntiles + kBs - 1) / kBs;
  const unsigned beg = threadIdx.x * per;
  const unsigned end = min(beg + per, ntiles);
  unsigned s = 0;
  for (unsigned i = beg; i < end; i++) s += h[i];
  // Block exclusive scan of s.
  const unsigned lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  unsigned incl = s;
  for (int off = 1; off < 64; off <<= 1) {
    const unsigned v = __shfl_up((int)incl, off);
    if (lane >= (unsigned)off) incl += v;
  }
  if (lane == 63) wsum[w] = incl;
  __syncthreads();
  unsigned wbase = 0;
  for (unsigned i = 0; i < w; i++) wbase += wsum[i];
  unsigned run = wbase + incl - s;
  for (unsigned i = beg; i < end; i++) {
    const unsigned v = h[i];
    h[i] = run;
    run += v;
  }
  if (threadIdx.x == kBs - 1) totals[blockIdx.x] = wbase + incl;
}

// Scatter of one pass. Each tile ranks its keys stably (per wave: lanes with
// the same digit found by one ballot per digit bit, per-wave digit counters
// in LDS), stages the pairs in LDS in digit order (tile-local positions), and
// then writes every digit run with consecutive threads on consecutive
// addresses: one coalesced store per run instead of one store per key at its
// own global slot.
__global__ void __launch_bounds__(kBs) k_radix_scatter(const unsigned* __restrict__ kin,
    const unsigned* __restrict__ vin, unsigned* __restrict__ kout, unsigned* __restrict__ vout,
    unsigned n, unsigned shift, unsigned bits, unsigned ntiles,
    const unsigned* __restrict__ hist, const unsigned* __restrict__ totals) {
  __shared__ unsigned dbase[256];           // global slot of the tile's first key of digit d
  __shared__ unsigned lstart[256];          // tile-local position of the first key of digit d
  __shared__ unsigned wcnt[kBs / 64][256];  // per-wave digit counters, then wave offsets
  __shared__ unsigned ws[kBs / 64];
  __shared__ unsigned skey[kTile], sval[kTile];
  const unsigned mask = (1u << bits) - 1u;
  const unsigned tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const unsigned tile = blockIdx.x;
  const unsigned tbase = tile * kTile;
  const unsigned tcount = min(kTile, n - tbase);
  // Global base of every digit for this tile: exclusive scan of the digit
  // totals plus the tile's offset within the digit.
  {
    const unsigned t = (tid <= mask) ? totals[tid] : 0u;
    unsigned incl = t;
    for (int off = 1; off < 64; off <<= 1) {
      const unsigned v = __shfl_up((int)incl, off);
      if (lane >= (unsigned)off) incl += v;
    }
    if (lane == 63) ws[w] = incl;
#pragma unroll
    for (unsigned k = 0; k < kBs / 64; k++) wcnt[k][tid] = 0;
    __syncthreads();
    unsigned wb = 0;
    for (unsigned i = 0; i < w; i++) wb += ws[i];
    if (tid <= mask) dbase[tid] = wb + incl - t + hist[tid * ntiles + tile];
  }
  __syncthreads();
  volatile unsigned* cnt = wcnt[w];
  const unsigned seg = w * (kTile / (kBs / 64));   // wave segment, tile-local
  unsigned key[kTileItems], val[kTileItems], rank[kTileItems];
  const unsigned long long lanelt = (1ull << lane) - 1ull;
#pragma unroll
  for (unsigned r = 0; r < kTileItems; r++) {
    const unsigned li = seg + r * 64 + lane;
    const bool valid = li < tcount;
    unsigned k = 0, v = 0;
    if (valid) { k = kin[tbase + li]; v = vin[tbase + li]; }
    const unsigned d = (k >> shift) & mask;
    unsigned long long peers = __ballot(valid);
    for (unsigned b = 0; b < bits; b++) {
      const bool bit = (d >> b) & 1u;
      const unsigned long long bal = __ballot(valid && bit);
      peers &= bit ? bal : ~bal;
    }
    const unsigned before = (unsigned)__popcll(peers & lanelt);
    unsigned rk = 0;
    if (valid) rk = cnt[d] + before;
    // The highest lane of each peer group advances the digit counter.
    const bool leader = valid && ((peers >> lane) == 1ull);
    if (leader) cnt[d] = rk + 1;
    key[r] = k; val[r] = v; rank[r] = valid ? rk : 0xffffffffu;
  }
  __syncthreads();
  // Per digit: offsets of the waves (exclusive scan over waves) and the
  // tile's count of the digit.
  unsigned dcount = 0;
  {
    const unsigned d = tid;   // kBs == 256 digits
    unsigned s = 0;
#pragma unroll
    for (unsigned k = 0; k < kBs / 64; k++) {
      const unsigned c = wcnt[k][d];
      wcnt[k][d] = s;
      s += c;
    }
    dcount = (d <= mask) ? s : 0u;
  }
  // Exclusive scan of the digit counts over digits: tile-local digit starts.
  {
    unsigned incl = dcount;
    for (int off = 1; off < 64; off <<= 1) {
      const unsigned v = __shfl_up((int)incl, off);
      if (lane >= (unsigned)off) incl += v;
    }
    __syncthreads();
    if (lane == 63) ws[w] = incl;
    __syncthreads();
    unsigned wb = 0;
    for (unsigned i = 0; i < w; i++) wb += ws[i];
    lstart[tid] = wb + incl - dcount;
  }
  __syncthreads();
#pragma unroll
  for (unsigned r = 0; r < kTileItems; r++) {
    if (rank[r] != 0xffffffffu) {
      const unsigned d = (key[r] >> shift) & mask;
      const unsigned lp = lstart[d] + wcnt[w][d] + rank[r];
      skey[lp] = key[r];
      sval[lp] = val[r];
    }
  }
  __syncthreads();
  for (unsigned i = tid; i < tcount; i += kBs) {
    const unsigned k = skey[i];
    const unsigned d = (k >> shift) & mask;
    const unsigned pos = dbase[d] + (i - lstart[d]);
    kout[pos] = k;
    vout[pos] = sval[i];
  }
}

// ------------------------------------------------------------- BeginEndCell
__global__ void k_fill_int2(int2* __restrict__ a, unsigned n, int2 v) {
  const unsigned p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) a[p] = v;
}

__global__ void __launch_bounds__(kBs) k_begin_end(unsigned n, unsigned pini,
    const unsigned* __restrict__ cellpart, int2* __restrict__ begcell) {
  const unsigned pt = blockIdx.x * blockDim.x + threadIdx.x;
  if (pt >= n) return;
  const unsigned p = pt + pini;
  const unsigned cel = cellpart[p];
  if (pt == 0) {
    begcell[cel].x = int(p);
  } else {
    const unsigned prev = cellpart[p - 1];
    if (cel != prev) {
      begcell[cel].x = int(p);
      begcell[prev].y = int(p);
    }
  }
  if (pt == n - 1) begcell[cel].y = int(p + 1);
}

// ------------------------------------------------------------- gathers
__device__ inline float4 poscell_of(double x, double y, double z, double3 posmin, float pcs) {
  const double dx = x - posmin.x, dy = y - posmin.y, dz = z - posmin.z;
  const unsigned cx = unsigned(dx / pcs), cy = unsigned(dy / pcs), cz = unsigned(dz / pcs);
  const float px = float(dx - double(pcs) * cx);
  const float py = float(dy - double(pcs) * cy);
  const float pz = float(dz - double(pcs) * cz);
  return make_float4(px, py, pz, __uint_as_float(pscel_code(cx, cy, cz)));
}

template <bool kPosCell>
__global__ void __launch_bounds__(kBs) k_sort_basic(unsigned n, unsigned pini,
    const unsigned* __restrict__ sortpart, const unsigned* __restrict__ idp,
    const typecode* __restrict__ code, const unsigned* __restrict__ dcell,
    const double2* __restrict__ posxy, const double* __restrict__ posz,
    const float4* __restrict__ velrho, unsigned* __restrict__ idp2,
    typecode* __restrict__ code2, unsigned* __restrict__ dcell2,
    double2* __restrict__ posxy2, double* __restrict__ posz2,
    float4* __restrict__ velrho2, double3 posmin, float pcs,
    float4* __restrict__ poscell) {
  const unsigned p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const unsigned o = (p < pini ? p : sortpart[p]);
  idp2[p] = idp[o];
  code2[p] = code[o];
  dcell2[p] = dcell[o];
  const double2 xy = posxy[o];
  const double z = posz[o];
  posxy2[p] = xy;
  posz2[p] = z;
  velrho2[p] = velrho[o];
  if (kPosCell) poscell[p] = poscell_of(xy.x, xy.y, z, posmin, pcs);
}

template <typename T>
__global__ void __launch_bounds__(kBs) k_sort_one(unsigned n, unsigned pini,
    const unsigned* __restrict__ sortpart, const T* __restrict__ a, T* __restrict__ a2) {
  const unsigned p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  a2[p] = a[p < pini ? p : sortpart[p]];
}

struct Sym6 { float v[6]; };
struct F3 { float x, y, z; };

__global__ void __launch_bounds__(kBs) k_sort_d2_d_f4(unsigned n, unsigned pini,
    const unsigned* __restrict__ sortpart, const double2* __restrict__ a,
    const double* __restrict__ b, const float4* __restrict__ c, double2* __restrict__ a2,
    double* __restrict__ b2, float4* __restrict__ c2) {
  const unsigned p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const unsigned o = (p < pini ? p : sortpart[p]);
  a2[p] = a[o];
  b2[p] = b[o];
  c2[p] = c[o];
}

__global__ void __launch_bounds__(kBs) k_update_poscell(unsigned np, double3 posmin, float pcs,
    const double2* __restrict__ posxy, const double* __restrict__ posz,
    float4* __restrict__ poscell) {
  const unsigned p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const double2 xy = posxy[p];
  poscell[p] = poscell_of(xy.x, xy.y, posz[p], posmin, pcs);
}

// ---------------------------------------------------------------- ensembles
// Keys of every particle of every case: the case's box key (full divide over
// its fixed grid) plus the case's key base.
__global__ void __launch_bounds__(kBs) k_ens_presort(const EnsCase* __restrict__ cases,
    const EnsBlock* __restrict__ pmap, const unsigned* __restrict__ dcell,
    const typecode* __restrict__ code, unsigned* __restrict__ cellpart,
    unsigned* __restrict__ sortpart) {
  const EnsBlock e = pmap[blockIdx.x];
  const EnsCase& ec = cases[e.x];
  const unsigned l = e.y + threadIdx.x;
  if (l >= ec.np0) return;
  const unsigned p = ec.p0 + l;
  sortpart[p] = p;
  cellpart[p] = ec.key0 + presort_full_box(p, ec.c.cellcode, dcell, code, make_uint3(0, 0, 0),
                                           make_uint3(ec.ncx, ec.ncy, ec.ncz));
}

// NpFinal and NpbOk of each case from the sizes of its out boxes
// (JCellDivGpuSingle.cpp:218-236).
__global__ void k_ens_counts(const EnsCase* __restrict__ cases, EnsState* __restrict__ st,
                             unsigned ncases, const int2* __restrict__ bc) {
  const unsigned k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ncases) return;
  const EnsCase& ec = cases[k];
  const int2* b = bc + ec.key0;
  auto size = [&](unsigned box) { return unsigned(b[box].y - b[box].x); };
  const unsigned boundignore = ec.nct, boundout = 2 * ec.nct + 1;
  const unsigned nbout = size(boundout), nfout = size(boundout + 1);
  const unsigned nboutign = size(boundout + 2), nfoutign = size(boundout + 3);
  EnsState& s = st[k];
  s.np = ec.np0 - nbout - nfout - nboutign - nfoutign;
  s.npbok = ec.npb - nboutign - size(boundignore);
  if (nbout || nboutign) {
    s.flags |= 4u;
    s.active = 0;
  }
}

__global__ void __launch_bounds__(kBs) k_ens_sort_basic(const EnsCase* __restrict__ cases,
    const EnsBlock* __restrict__ pmap, const unsigned* __restrict__ sortpart,
    const unsigned* __restrict__ idp, const typecode* __restrict__ code,
    const unsigned* __restrict__ dcell, const double2* __restrict__ posxy,
    const double* __restrict__ posz, const float4* __restrict__ velrho,
    unsigned* __restrict__ idp2, typecode* __restrict__ code2, unsigned* __restrict__ dcell2,
    double2* __restrict__ posxy2, double* __restrict__ posz2, float4* __restrict__ velrho2,
    float4* __restrict__ poscell) {
  const EnsBlock e = pmap[blockIdx.x];
  const EnsCase& ec = cases[e.x];
  const unsigned l = e.y + threadIdx.x;
  if (l >= ec.np0) return;
  const unsigned p = ec.p0 + l;
  const unsigned o = sortpart[p];
  idp2[p] = idp[o];
  code2[p] = code[o];
  dcell2[p] = dcell[o];
  const double2 xy = posxy[o];
  const double z = posz[o];
  posxy2[p] = xy;
  posz2[p] = z;
  velrho2[p] = velrho[o];
  poscell[p] = poscell_of(xy.x, xy.y, z, make_double3(ec.posminx, ec.posminy, ec.posminz),
                          ec.c.pcs);
}


// ---------------------------------------------------------------- periodic boundaries
// KerPeriodicIgnore (JSphGpu_ker.cu:2113-2122).
__global__ void __launch_bounds__(kBs) k_periodic_ignore(unsigned n, uint16_t* __restrict__ code) {
  const unsigned p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const unsigned c = code[p];
  if (code_special(c) == kCodePeriodic) code[p] = uint16_t(code_set(c, kCodeOutIgnore));
}

// Entries a particle adds to the periodic list (KerPeriodicMakeList,
// JSphGpu_ker.cu:2142-2186): bit 0 plus, bit 1 minus.
__device__ inline unsigned periodic_flags(unsigned p2, double3 mn, double3 mx, double3 inc,
                                          const double2* posxy, const double* posz,
                                          const uint16_t* code) {
  if (code_special(code[p2]) > kCodePeriodic) return 0;
  const double2 rxy = posxy[p2];
  const double rz = posz[p2];
  unsigned f = 0;
  double x = rxy.x + inc.x, y = rxy.y + inc.y, z = rz + inc.z;
  if (mn.x <= x && mn.y <= y && mn.z <= z && x < mx.x && y < mx.y && z < mx.z) f |= 1;
  x = rxy.x - inc.x; y = rxy.y - inc.y; z = rz - inc.z;
  if (mn.x <= x && mn.y <= y && mn.z <= z && x < mx.x && y < mx.y && z < mx.z) f |= 2;
  return f;
}

// Inclusive prefix sum over the block of kBs lanes (wave scan + LDS).
__device__ inline unsigned block_incl_scan(unsigned v, unsigned* lds) {
  const unsigned lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
  for (unsigned o = 1; o < 64; o <<= 1) {
    const unsigned t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  if (lane == 63) lds[wv] = v;
  __syncthreads();
  unsigned add = 0;
  for (unsigned w = 0; w < wv; w++) add += lds[w];
  __syncthreads();
  return v + add;
}

// Pass 1: entries per block of kBs particles.
__global__ void __launch_bounds__(kBs) k_periodic_count(unsigned n, unsigned pini, double3 mn,
    double3 mx, double3 inc, const double2* __restrict__ posxy, const double* __restrict__ posz,
    const uint16_t* __restrict__ code, unsigned* __restrict__ blockcount) {
  __shared__ unsigned lds[kBs / 64];
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  const unsigned f = (i < n) ? periodic_flags(i + pini, mn, mx, inc, posxy, posz, code) : 0u;
  const unsigned s = block_incl_scan((f & 1) + (f >> 1), lds);
  if (threadIdx.x == kBs - 1) blockcount[blockIdx.x] = s;
}

// Pass 2: exclusive scan of the block counts (one block; total at [nb]).
__global__ void __launch_bounds__(kBs) k_periodic_scan(unsigned nb, unsigned* __restrict__ blockcount) {
  __shared__ unsigned lds[kBs / 64];
  __shared__ unsigned carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (unsigned b0 = 0; b0 < nb; b0 += kBs) {
    const unsigned b = b0 + threadIdx.x;
    const unsigned v = (b < nb) ? blockcount[b] : 0u;
    const unsigned s = block_incl_scan(v, lds);
    const unsigned c0 = carry;
    if (b < nb) blockcount[b] = c0 + s - v;
    __syncthreads();
    if (threadIdx.x == kBs - 1) carry = c0 + s;
    __syncthreads();
  }
  if (threadIdx.x == 0) blockcount[nb] = carry;
}

// Pass 3: write the entries in particle order, plus before minus.
__global__ void __launch_bounds__(kBs) k_periodic_write(unsigned n, unsigned pini, unsigned nmax,
    double3 mn, double3 mx, double3 inc, const double2* __restrict__ posxy,
    const double* __restrict__ posz, const uint16_t* __restrict__ code,
    const unsigned* __restrict__ blockoff, unsigned* __restrict__ listp) {
  __shared__ unsigned lds[kBs / 64];
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  const unsigned p2 = i + pini;
  const unsigned f = (i < n) ? periodic_flags(p2, mn, mx, inc, posxy, posz, code) : 0u;
  const unsigned cnt = (f & 1) + (f >> 1);
  unsigned o = blockoff[blockIdx.x] + block_incl_scan(cnt, lds) - cnt;
  if (f & 1) { if (o < nmax) listp[o] = p2; o++; }
  if (f & 2) { if (o < nmax) listp[o] = p2 | 0x80000000u; }
}

// KerPeriodicDuplicatePos + KerPeriodicDuplicateSymplectic / Verlet / Normals
// (JSphGpu_ker.cu:2229-2400).
__global__ void __launch_bounds__(kBs) k_periodic_duplicate(unsigned n, unsigned pini,
    uint3 cellmax, double3 inc, double3 dmin, double scell, unsigned cellcode,
    const unsigned* __restrict__ listp, unsigned* idp, uint16_t* code, unsigned* dcell,
    double2* posxy, double* posz, float4* velrho, float* spstau, double2* posxypre,
    double* poszpre, float4* velrhopre, float4* velrhom1, float* boundnor) {
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const unsigned pnew = i + pini;
  const unsigned rp = listp[i];
  const unsigned pc = rp & 0x7fffffffu;
  const bool inverse = rp >= 0x80000000u;
  double2 rxy = posxy[pc];
  double rz = posz[pc];
  rxy.x += (inverse ? -inc.x : inc.x);
  rxy.y += (inverse ? -inc.y : inc.y);
  rz += (inverse ? -inc.z : inc.z);
  unsigned cx = unsigned((rxy.x - dmin.x) / scell);
  unsigned cy = unsigned((rxy.y - dmin.y) / scell);
  unsigned cz = unsigned((rz - dmin.z) / scell);
  cx = (cx <= cellmax.x ? cx : cellmax.x);
  cy = (cy <= cellmax.y ? cy : cellmax.y);
  cz = (cz <= cellmax.z ? cz : cellmax.z);
  posxy[pnew] = rxy;
  posz[pnew] = rz;
  dcell[pnew] = dcel_cell(cellcode, cx, cy, cz);
  idp[pnew] = idp[pc];
  code[pnew] = uint16_t((code[pc] & ~kCodeMaskSpecial) | kCodePeriodic);
  velrho[pnew] = velrho[pc];
  if (posxypre) { posxypre[pnew] = posxypre[pc]; poszpre[pnew] = poszpre[pc]; }
  if (velrhopre) velrhopre[pnew] = velrhopre[pc];
  if (velrhom1) velrhom1[pnew] = velrhom1[pc];
  if (spstau)
    for (int k = 0; k < 6; k++) spstau[6 * size_t(pnew) + k] = spstau[6 * size_t(pc) + k];
  if (boundnor)
    for (int k = 0; k < 3; k++) boundnor[3 * size_t(pnew) + k] = boundnor[3 * size_t(pc) + k];
}
}  // namespace


namespace dsph {

int ens_presort(dsph_ctx* ctx, const EnsCase* cases, const EnsBlock* pmap, unsigned npblocks,
                const unsigned* dcell, const uint16_t* code, unsigned* cellpart,
                unsigned* sortpart) {
  if (!npblocks) return DSPH_OK;
  KTimer kt(ctx, "k_presort");
  k_ens_presort<<<npblocks, kBs, 0, ctx->stream>>>(cases, pmap, dcell, code, cellpart, sortpart);
  DSPH_LAUNCHED(ctx, "k_ens_presort");
  return DSPH_OK;
}

int ens_counts(dsph_ctx* ctx, const EnsCase* cases, EnsState* st, unsigned ncases,
               const int2* begincell) {
  if (!ncases) return DSPH_OK;
  k_ens_counts<<<grid_for(ncases, 64), 64, 0, ctx->stream>>>(cases, st, ncases, begincell);
  DSPH_LAUNCHED(ctx, "k_ens_counts");
  return DSPH_OK;
}

int ens_sort_basic_poscell(dsph_ctx* ctx, const EnsCase* cases, const EnsBlock* pmap,
                           unsigned npblocks, const unsigned* sortpart, const unsigned* idp,
                           const uint16_t* code, const unsigned* dcell, const double2* posxy,
                           const double* posz, const float4* velrho, unsigned* idp2,
                           uint16_t* code2, unsigned* dcell2, double2* posxy2, double* posz2,
                           float4* velrho2, float4* poscell) {
  if (!npblocks) return DSPH_OK;
  KTimer kt(ctx, "k_sort_basic_poscell");
  k_ens_sort_basic<<<npblocks, kBs, 0, ctx->stream>>>(cases, pmap, sortpart, idp, code, dcell,
                                                      posxy, posz, velrho, idp2, code2, dcell2,
                                                      posxy2, posz2, velrho2, poscell);
  DSPH_LAUNCHED(ctx, "k_ens_sort_basic");
  return DSPH_OK;
}

// Radix sort of (keys, values) in place with the scratch `buf` laid out for
// `cap` keys (sort_scratch_words): two alternate key/value buffers A and B
// and the digit histograms. The passes alternate between the caller's arrays
// and A (two or four passes) or go d -> A -> B -> d (three passes), so the
// result always ends in the caller's arrays without a copy; a key range of
// one digit is sorted twice (the second pass leaves the order unchanged).
// Nothing in the scratch needs clearing: every pass writes the histograms
// of its tiles before reading them.
int radix_sort(dsph_ctx* ctx, uint32_t* d_keys, uint32_t* d_values, uint32_t n,
               uint32_t key_limit, uint32_t* buf, size_t cap) {
  unsigned bits_total = 0;
  while (bits_total < 32 && (key_limit - 1u) >> bits_total) bits_total++;
  if (bits_total == 0) return DSPH_OK;  // all keys equal (key_limit<=1)
  unsigned *ka_alt = buf, *va_alt = buf + cap, *kb_alt = buf + 2 * cap, *vb_alt = buf + 3 * cap;
  unsigned* hist = buf + 4 * cap;
  const unsigned ntiles = (n + kTile - 1) / kTile;
  unsigned* totals = hist + size_t(256) * ntiles;
  unsigned npass = (bits_total + 7) / 8;
  // (shift, bits) of every pass.
  unsigned sh[4], bw[4];
  for (unsigned p = 0; p < npass; p++) {
    sh[p] = 8 * p;
    bw[p] = min(8u, bits_total - 8 * p);
  }
  if (npass == 1) {
    sh[1] = sh[0];
    bw[1] = bw[0];
    npass = 2;
  }
  unsigned* kbuf[5];
  unsigned* vbuf[5];
  kbuf[0] = d_keys; vbuf[0] = d_values;
  for (unsigned p = 1; p <= npass; p++) {
    const bool to_caller = (p == npass) || (npass != 3 && p % 2 == 0);
    const bool to_b = npass == 3 && p == 2;
    kbuf[p] = to_caller ? d_keys : (to_b ? kb_alt : ka_alt);
    vbuf[p] = to_caller ? d_values : (to_b ? vb_alt : va_alt);
  }
  for (unsigned p = 0; p < npass; p++) {
    const unsigned mask = (1u << bw[p]) - 1u;
    k_radix_hist<<<ntiles, kBs, 0, ctx->stream>>>(kbuf[p], n, sh[p], mask, ntiles, hist);
    k_radix_scan_digit<<<mask + 1, kBs, 0, ctx->stream>>>(hist, ntiles, totals);
    k_radix_scatter<<<ntiles, kBs, 0, ctx->stream>>>(kbuf[p], vbuf[p], kbuf[p + 1], vbuf[p + 1],
                                                     n, sh[p], bw[p], ntiles, hist, totals);
    DSPH_LAUNCHED(ctx, "radix pass");
  }
  return DSPH_OK;
}

}  // namespace dsph

extern "C" {

int dsph_cell_limits_sync(dsph_ctx* ctx, uint32_t np, uint32_t pini, uint32_t cellcode,
                          const uint32_t* d_dcell, const uint16_t* d_code,
                          uint32_t h_celmin[3], uint32_t h_celmax[3]) {
  if (!ctx || !h_celmin || !h_celmax) return DSPH_E_ARG;
  if (!np) {
    for (int k = 0; k < 3; k++) { h_celmin[k] = UINT_MAX; h_celmax[k] = 0; }
    return DSPH_OK;
  }
  DSPH_REQUIRE(ctx, d_dcell && d_code, "dsph_cell_limits_sync: null array");
  // Block partials in the reduction scratch (>= 256 words after dsph_ctx_reserve).
  if (ctx->redu_cap < 6)
    return set_error(ctx, DSPH_E_NOMEM, "dsph_cell_limits_sync: call dsph_ctx_reserve first");
  unsigned* res = ctx->small_dev;
  unsigned* part = reinterpret_cast<unsigned*>(ctx->redu);
  const unsigned nparts = std::min<unsigned>(
      std::min(grid_for(np, kBs), 1024u), unsigned(std::min<size_t>(ctx->redu_cap / 6, 1024)));
  {
    KTimer kt(ctx, "k_limits");
    k_limits<<<nparts, kBs, 0, ctx->stream>>>(np, pini, cellcode, d_dcell, d_code, part);
    k_limits_final<<<1, kBs, 0, ctx->stream>>>(nparts, part, res);
    DSPH_LAUNCHED(ctx, "k_limits");
  }
  unsigned* h = static_cast<unsigned*>(ctx->pinned);
  DSPH_HIP(ctx, hipMemcpyAsync(h, res, 6 * sizeof(unsigned), hipMemcpyDeviceToHost, ctx->stream));
  DSPH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (int k = 0; k < 3; k++) { h_celmin[k] = h[k]; h_celmax[k] = h[3 + k]; }
  return DSPH_OK;
}

int dsph_periodic_ignore(dsph_ctx* ctx, uint32_t n, uint16_t* d_code) {
  if (!ctx) return DSPH_E_ARG;
  if (!n) return DSPH_OK;
  DSPH_REQUIRE(ctx, d_code, "dsph_periodic_ignore: null argument");
  k_periodic_ignore<<<grid_for(n, kBs), kBs, 0, ctx->stream>>>(n, d_code);
  DSPH_LAUNCHED(ctx, "k_periodic_ignore");
  return DSPH_OK;
}

int dsph_periodic_make_list(dsph_ctx* ctx, uint32_t n, uint32_t pini, uint32_t nmax,
                            const double mapposmin[3], const double mapposmax[3],
                            const double perinc[3], const double* d_posxy, const double* d_posz,
                            const uint16_t* d_code, uint32_t* d_work, uint32_t* d_listp,
                            uint32_t* h_count) {
  if (!ctx || !h_count) return DSPH_E_ARG;
  *h_count = 0;
  if (!n) return DSPH_OK;
  DSPH_REQUIRE(ctx, mapposmin && mapposmax && perinc && d_posxy && d_posz && d_code && d_work &&
                        d_listp, "dsph_periodic_make_list: null argument");
  const double3 mn = make_double3(mapposmin[0], mapposmin[1], mapposmin[2]);
  const double3 mx = make_double3(mapposmax[0], mapposmax[1], mapposmax[2]);
  const double3 inc = make_double3(perinc[0], perinc[1], perinc[2]);
  const unsigned nb = grid_for(n, kBs);
  const double2* pxy = reinterpret_cast<const double2*>(d_posxy);
  k_periodic_count<<<nb, kBs, 0, ctx->stream>>>(n, pini, mn, mx, inc, pxy, d_posz, d_code, d_work);
  k_periodic_scan<<<1, kBs, 0, ctx->stream>>>(nb, d_work);
  k_periodic_write<<<nb, kBs, 0, ctx->stream>>>(n, pini, nmax, mn, mx, inc, pxy, d_posz, d_code,
                                                d_work, d_listp);
  DSPH_LAUNCHED(ctx, "k_periodic_list");
  DSPH_HIP(ctx, hipMemcpyAsync(ctx->pinned, d_work + nb, sizeof(uint32_t), hipMemcpyDeviceToHost,
                               ctx->stream));
  DSPH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *h_count = *static_cast<uint32_t*>(ctx->pinned);
  return DSPH_OK;
}

int dsph_periodic_duplicate(dsph_ctx* ctx, uint32_t n, uint32_t pini, const uint32_t domcells[3],
                            const double perinc[3], const uint32_t* d_listp, uint32_t* d_idp,
                            uint16_t* d_code, uint32_t* d_dcell, double* d_posxy, double* d_posz,
                            float* d_velrho, float* d_spstau, double* d_posxypre,
                            double* d_poszpre, float* d_velrhopre, float* d_velrhom1,
                            float* d_boundnor) {
  if (!ctx) return DSPH_E_ARG;
  if (!n) return DSPH_OK;
  DSPH_REQUIRE(ctx, ctx->cte_set, "dsph_periodic_duplicate: constants not set");
  DSPH_REQUIRE(ctx, domcells && perinc && d_listp && d_idp && d_code && d_dcell && d_posxy &&
                        d_posz && d_velrho && (!d_posxypre == !d_poszpre),
               "dsph_periodic_duplicate: null argument");
  const dsph_cte& c = ctx->cte;
  k_periodic_duplicate<<<grid_for(n, kBs), kBs, 0, ctx->stream>>>(
      n, pini, make_uint3(domcells[0] - 1, domcells[1] - 1, domcells[2] - 1),
      make_double3(perinc[0], perinc[1], perinc[2]),
      make_double3(c.domposminx, c.domposminy, c.domposminz), double(c.scell), c.cellcode,
      d_listp, d_idp, d_code, d_dcell, reinterpret_cast<double2*>(d_posxy), d_posz,
      reinterpret_cast<float4*>(d_velrho), d_spstau, reinterpret_cast<double2*>(d_posxypre),
      d_poszpre, reinterpret_cast<float4*>(d_velrhopre), reinterpret_cast<float4*>(d_velrhom1),
      d_boundnor);
  DSPH_LAUNCHED(ctx, "k_periodic_duplicate");
  return DSPH_OK;
}

int dsph_presort_full(dsph_ctx* ctx, uint32_t np, uint32_t cellcode, const uint32_t* d_dcell,
                      const uint16_t* d_code, const uint32_t cellmin[3],
                      const uint32_t ncells[3], uint32_t* d_cellpart, uint32_t* d_sortpart) {
  if (!ctx) return DSPH_E_ARG;
  if (!np) return DSPH_OK;
  DSPH_REQUIRE(ctx, d_dcell && d_code && d_cellpart && d_sortpart && cellmin && ncells,
               "dsph_presort_full: null argument");
  KTimer kt(ctx, "k_presort");
  k_presort_full<<<grid_for(np, kBs), kBs, 0, ctx->stream>>>(
      np, cellcode, d_dcell, d_code, make_uint3(cellmin[0], cellmin[1], cellmin[2]),
      make_uint3(ncells[0], ncells[1], ncells[2]), d_cellpart, d_sortpart);
  DSPH_LAUNCHED(ctx, "k_presort_full");
  return DSPH_OK;
}

int dsph_presort_fluid(dsph_ctx* ctx, uint32_t npf, uint32_t pini, uint32_t cellcode,
                       const uint32_t* d_dcell, const uint16_t* d_code,
                       const uint32_t cellmin[3], const uint32_t ncells[3],
                       uint32_t* d_cellpart, uint32_t* d_sortpart) {
  if (!ctx) return DSPH_E_ARG;
  if (!npf) return DSPH_OK;
  DSPH_REQUIRE(ctx, d_dcell && d_code && d_cellpart && d_sortpart && cellmin && ncells,
               "dsph_presort_fluid: null argument");
  KTimer kt(ctx, "k_presort");
  k_presort_fluid<<<grid_for(npf, kBs), kBs, 0, ctx->stream>>>(
      npf, pini, cellcode, d_dcell, d_code, make_uint3(cellmin[0], cellmin[1], cellmin[2]),
      make_uint3(ncells[0], ncells[1], ncells[2]), d_cellpart, d_sortpart);
  DSPH_LAUNCHED(ctx, "k_presort_fluid");
  return DSPH_OK;
}

int dsph_sort_by_cell(dsph_ctx* ctx, uint32_t* d_keys, uint32_t* d_values, uint32_t n,
                      uint32_t key_limit) {
  if (!ctx) return DSPH_E_ARG;
  if (n <= 1) return DSPH_OK;
  DSPH_REQUIRE(ctx, d_keys && d_values, "dsph_sort_by_cell: null array");
  int rc = ensure_sort_scratch(ctx, n);
  if (rc) return rc;
  KTimer kt(ctx, "radix_sort");
  return radix_sort(ctx, d_keys, d_values, n, key_limit, ctx->sort_buf, ctx->sort_cap);
}

int dsph_begin_end_cell(dsph_ctx* ctx, int full, uint32_t np, uint32_t npb,
                        uint32_t sizebegcell, uint32_t cellfluid, const uint32_t* d_cellpart,
                        int32_t* d_begcell) {
  if (!ctx) return DSPH_E_ARG;
  DSPH_REQUIRE(ctx, d_cellpart && d_begcell, "dsph_begin_end_cell: null array");
  KTimer kt(ctx, "k_begin_end");
  int2* bc = reinterpret_cast<int2*>(d_begcell);
  const unsigned c0 = full ? 0u : cellfluid;
  if (sizebegcell > c0) {
    const unsigned nc = sizebegcell - c0;
    k_fill_int2<<<grid_for(nc, kBs), kBs, 0, ctx->stream>>>(bc + c0, nc, make_int2(0, 0));
  }
  const unsigned pini = full ? 0u : npb;
  if (np > pini) {
    const unsigned n = np - pini;
    k_begin_end<<<grid_for(n, kBs), kBs, 0, ctx->stream>>>(n, pini, d_cellpart, bc);
  }
  DSPH_LAUNCHED(ctx, "k_begin_end");
  return DSPH_OK;
}

int dsph_cell_begin_end_sync(dsph_ctx* ctx, const int32_t* d_begcell, uint32_t cellini,
                             uint32_t ncells, int32_t* h_out) {
  if (!ctx || !h_out) return DSPH_E_ARG;
  if (!ncells) return DSPH_OK;
  DSPH_REQUIRE(ctx, ncells * 8 <= 4096, "dsph_cell_begin_end_sync: too many cells");
  DSPH_HIP(ctx, hipMemcpyAsync(ctx->pinned, d_begcell + 2 * size_t(cellini), 8 * size_t(ncells),
                               hipMemcpyDeviceToHost, ctx->stream));
  DSPH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  std::memcpy(h_out, ctx->pinned, 8 * size_t(ncells));
  return DSPH_OK;
}

int dsph_sort_data_basic(dsph_ctx* ctx, uint32_t np, uint32_t pini, const uint32_t* d_sortpart,
                         const uint32_t* d_idp, const uint16_t* d_code, const uint32_t* d_dcell,
                         const double* d_posxy, const double* d_posz, const float* d_velrho,
                         uint32_t* d_idp2, uint16_t* d_code2, uint32_t* d_dcell2,
                         double* d_posxy2, double* d_posz2, float* d_velrho2) {
  if (!ctx) return DSPH_E_ARG;
  if (!np) return DSPH_OK;
  k_sort_basic<false><<<grid_for(np, kBs), kBs, 0, ctx->stream>>>(
      np, pini, d_sortpart, d_idp, d_code, d_dcell, (const double2*)d_posxy, d_posz,
      (const float4*)d_velrho, d_idp2, d_code2, d_dcell2, (double2*)d_posxy2, d_posz2,
      (float4*)d_velrho2, make_double3(0, 0, 0), 1.f, nullptr);
  DSPH_LAUNCHED(ctx, "k_sort_basic");
  return DSPH_OK;
}

int dsph_sort_data_basic_poscell(dsph_ctx* ctx, uint32_t np, uint32_t pini,
                                 const uint32_t* d_sortpart, const uint32_t* d_idp,
                                 const uint16_t* d_code, const uint32_t* d_dcell,
                                 const double* d_posxy, const double* d_posz,
                                 const float* d_velrho, uint32_t* d_idp2, uint16_t* d_code2,
                                 uint32_t* d_dcell2, double* d_posxy2, double* d_posz2,
                                 float* d_velrho2, const double posmin[3], float poscellsize,
                                 float* d_poscell) {
  if (!ctx || !posmin) return DSPH_E_ARG;
  if (!np) return DSPH_OK;
  KTimer kt(ctx, "k_sort_basic_poscell");
  k_sort_basic<true><<<grid_for(np, kBs), kBs, 0, ctx->stream>>>(
      np, pini, d_sortpart, d_idp, d_code, d_dcell, (const double2*)d_posxy, d_posz,
      (const float4*)d_velrho, d_idp2, d_code2, d_dcell2, (double2*)d_posxy2, d_posz2,
      (float4*)d_velrho2, make_double3(posmin[0], posmin[1], posmin[2]), poscellsize,
      (float4*)d_poscell);
  DSPH_LAUNCHED(ctx, "k_sort_basic_poscell");
  return DSPH_OK;
}

int dsph_sort_data_f4(dsph_ctx* ctx, uint32_t np, uint32_t pini, const uint32_t* d_sortpart,
                      const float* d_a, float* d_a2) {
  if (!ctx) return DSPH_E_ARG;
  if (!np) return DSPH_OK;
  k_sort_one<float4><<<grid_for(np, kBs), kBs, 0, ctx->stream>>>(
      np, pini, d_sortpart, (const float4*)d_a, (float4*)d_a2);
  DSPH_LAUNCHED(ctx, "k_sort_f4");
  return DSPH_OK;
}

int dsph_sort_data_f3(dsph_ctx* ctx, uint32_t np, uint32_t pini, const uint32_t* d_sortpart,
                      const float* d_a, float* d_a2) {
  if (!ctx) return DSPH_E_ARG;
  if (!np) return DSPH_OK;
  k_sort_one<F3><<<grid_for(np, kBs), kBs, 0, ctx->stream>>>(np, pini, d_sortpart,
                                                              (const F3*)d_a, (F3*)d_a2);
  DSPH_LAUNCHED(ctx, "k_sort_f3");
  return DSPH_OK;
}

int dsph_sort_data_sym6(dsph_ctx* ctx, uint32_t np, uint32_t pini, const uint32_t* d_sortpart,
                        const float* d_a, float* d_a2) {
  if (!ctx) return DSPH_E_ARG;
  if (!np) return DSPH_OK;
  k_sort_one<Sym6><<<grid_for(np, kBs), kBs, 0, ctx->stream>>>(np, pini, d_sortpart,
                                                                (const Sym6*)d_a, (Sym6*)d_a2);
  DSPH_LAUNCHED(ctx, "k_sort_sym6");
  return DSPH_OK;
}

int dsph_sort_data_d2_d_f4(dsph_ctx* ctx, uint32_t np, uint32_t pini,
                           const uint32_t* d_sortpart, const double* d_a, const double* d_b,
                           const float* d_c, double* d_a2, double* d_b2, float* d_c2) {
  if (!ctx) return DSPH_E_ARG;
  if (!np) return DSPH_OK;
  k_sort_d2_d_f4<<<grid_for(np, kBs), kBs, 0, ctx->stream>>>(
      np, pini, d_sortpart, (const double2*)d_a, d_b, (const float4*)d_c, (double2*)d_a2, d_b2,
      (float4*)d_c2);
  DSPH_LAUNCHED(ctx, "k_sort_d2_d_f4");
  return DSPH_OK;
}

int dsph_update_poscell(dsph_ctx* ctx, uint32_t np, const double posmin[3], float poscellsize,
                        const double* d_posxy, const double* d_posz, float* d_poscell) {
  if (!ctx || !posmin) return DSPH_E_ARG;
  if (!np) return DSPH_OK;
  k_update_poscell<<<grid_for(np, kBs), kBs, 0, ctx->stream>>>(
      np, make_double3(posmin[0], posmin[1], posmin[2]), poscellsize, (const double2*)d_posxy,
      d_posz, (float4*)d_poscell);
  DSPH_LAUNCHED(ctx, "k_update_poscell");
  return DSPH_OK;
}

}  // extern "C"
