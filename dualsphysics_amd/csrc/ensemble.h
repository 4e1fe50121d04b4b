// Batched ensembles: many independent cases advanced together, one launch per
// stage for all of them.
//
// The particle arrays of the cases are concatenated (case k owns the index
// range [p0, p0+np0): boundary first, then fluid, as in a single case) and
// every case keeps its own constants, cell grid (fixed domain: the whole map
// of the case), time and adaptive dt. The sort key of a particle is its
// case's key base plus its box key inside the case, so one stable radix sort
// orders all cases at once and never moves a particle out of its case range;
// the begincell array is the concatenation of the cases' arrays. Kernels find
// their case through block maps (one entry per workgroup: case index and the
// first case-local index of the block), so the case record is wave-uniform
// and read with scalar loads. Particles excluded by a case (out of the map,
// bad density) sort to the end of its range and the case's current count
// shrinks, all on the device: the step has no host round trip and can be
// captured in a HIP graph.
#pragma once
#include "sph_math.h"

namespace dsph {

// Limits of KerUpdatePos (JSphGpu_ker.cu:334-392): map of the case and the
// displacement limit.
struct PosCte {
  float movlimit;
  double mapx, mapy, mapz, sizex, sizey, sizez;
  float scell;
  unsigned cellcode;
  unsigned periactive;              // periodic axes, bits 1 X, 2 Y, 4 Z (0 in ensembles)
  double perinc[3][3];              // PeriXinc, PeriYinc, PeriZinc
  double domx, domy, domz;          // DomPosMin (cells of periodic runs)
  unsigned sim2d;                   // 2-D: matrix motions keep y (0 in ensembles)
};

// Constant record of one case of an ensemble.
struct EnsCase {
  Cte c;
  DivData dvd;          // cellzero 0; begincell points at the case's segment
  PosCte pc;
  double posminx, posminy, posminz;  // PosCell map origin (Map_PosMin)
  float gx, gy, gz;     // gravity
  float rhopoutmin, rhopoutmax, viscob, viscof, ks2scan;
  double kernelh, cs0, cfl, dtmin, tend;
  unsigned p0, np0, npb;  // first particle, particles, boundary particles
  unsigned key0, nct;     // first sort key (begincell offset), cells of the grid
  unsigned ncx, ncy, ncz;
};

// Device-updated state of one case.
struct EnsState {
  double dt[2];          // dt of the predictor / corrector interaction
  double dtpre;          // SymplecticDtPre: the dt of the next step
  double time, dtlast;
  unsigned np, npbok;    // current particles (excluded ones at the end), NpbOk
  unsigned active;       // 0 once time >= tend or after an error
  unsigned nstep;
  unsigned flags;        // bit 0 dt NaN/inf, bit 1 dt clamped to DtMin, bit 2 boundary excluded
  unsigned red[3];       // FP32 bits of max |v|^2 (fluid), max viscdt, max |ace|^2 (fluid)
  unsigned pad;
};

// Block map entry: .x case, .y first case-local index of the block.
typedef uint2 EnsBlock;

constexpr unsigned kEnsBs = 256;      // particle kernels: blocks of 256 over [0, np0)
constexpr unsigned kEnsBsBound = 128; // boundary interaction: default blocks of 128 over [0, npb)
constexpr unsigned kEnsBsBoundMax = 256;  // largest boundary block the autotuner tries

// Fluid interaction kernel variants the autotuner chooses from.
enum EnsFluidVariant {
  kEnsTileRing8 = 0,
  kEnsTileRing12 = 1,
  kEnsTileRing16 = 2,
  kEnsRows64 = 3,
  kEnsNumVariants = 4
};

// Persistent launch of the fluid tile kernel: `grid` resident waves pull
// `chunk` work items at a time from ctr[0..8) (one counter per XCD segment of
// the work list; the launcher zeroes them with a fill kernel before the
// kernel, so the pair can sit in a captured graph).
struct EnsPersist {
  unsigned grid, chunk;
  unsigned* ctr;
};

// ---- launchers (defined in forces.hip, step.hip, celldiv.hip); all
// asynchronous on ctx's stream, no host synchronisation.
int ens_prep(dsph_ctx* ctx, const EnsCase* cases, const EnsState* st, const EnsBlock* pmap,
             unsigned npblocks, const float4* poscell, const float4* velrho, float2* press,
             float4* posf);
// wmap: one entry per wave of 64 fluid particles (.y = first fluid index of
// the wave, relative to the case's first fluid particle); without `pers`
// every variant runs one 64-lane workgroup per entry, with it the tile
// variants run persistent (see EnsPersist).
// eos7: every case has gamma = 7 (the kernels evaluate the neighbour
// pressure from its density instead of gathering it).
int ens_forces_fluid(dsph_ctx* ctx, int variant, int tvisco, int tddt, bool eos7,
                     const EnsCase* cases, const EnsState* st, const EnsBlock* wmap,
                     unsigned nblocks, const EnsPersist* pers, const unsigned* dcell,
                     const float4* poscell,
                     const float4* posf, const float4* velrho, const float2* press,
                     float* viscdt, float* ar, float* ace);
// Tile form of the boundary interaction (one wave per 64 boundary particles:
// bmap64 has one entry per 64; posf/press from ens_prep).
int ens_forces_bound_tile(dsph_ctx* ctx, const EnsCase* cases, const EnsState* st,
                          const EnsBlock* bmap64, unsigned nblocks, const unsigned* dcell,
                          const float4* poscell, const float4* posf, const float4* velrho,
                          const float2* press, float* viscdt, float* ar);
// bmap: one entry per block of bs boundary particles (bs = 64, 128, 192 or 256).
int ens_forces_bound(dsph_ctx* ctx, const EnsCase* cases, const EnsState* st,
                     const EnsBlock* bmap, unsigned nblocks, unsigned bs, const unsigned* dcell,
                     const float4* poscell, const float4* velrho, float* viscdt, float* ar);

// Radius^2 of the fluid tile kernel's conservative scan test for a case.
float ens_scan_radius2(const dsph_cte& cte);

// Zeroes ace, ar, viscdt of [0, n) (fill kernel, graph safe).
int ens_zero_forces(dsph_ctx* ctx, unsigned n, float* ace, float* ar, float* viscdt);
// Per-case maxima into EnsState::red (atomicMax on FP32 bits, values >= 0).
int ens_reduce(dsph_ctx* ctx, const EnsCase* cases, EnsState* st, const EnsBlock* pmap,
               unsigned npblocks, int which, const float4* velrho, const float* viscdt,
               const float* ace);
// dt of pass `slot` (0 predictor, 1 corrector) per case; clears red.
int ens_dt(dsph_ctx* ctx, const EnsCase* cases, EnsState* st, unsigned ncases, unsigned slot);
int ens_sym_pre(dsph_ctx* ctx, const EnsCase* cases, const EnsState* st, const EnsBlock* pmap,
                unsigned npblocks, const float4* velrhopre, const float* ar, const float* ace,
                uint16_t* code, double2* movxy, double* movz, float4* velrho);
int ens_sym_cor(dsph_ctx* ctx, const EnsCase* cases, const EnsState* st, const EnsBlock* pmap,
                unsigned npblocks, const float4* velrhopre, const float* ar, const float* ace,
                uint16_t* code, double2* movxy, double* movz, float4* velrho);
// Fluid positions = pre + mov (KerUpdatePos); with copy_bound the boundary
// positions are copied from the pre arrays (the predictor). In the
// predictor, stopped cases copy their state from the pre arrays (which the
// host swapped with the state arrays for all cases).
int ens_pos2(dsph_ctx* ctx, const EnsCase* cases, const EnsState* st, const EnsBlock* pmap,
             unsigned npblocks, bool copy_bound, const double2* posxypre, const double* poszpre,
             const double2* movxy, const double* movz, double2* posxy, double* posz,
             unsigned* dcell, uint16_t* code);
// SymplecticDtPre = min(dt_p, dt_c), time += dt; a case stops at tend.
int ens_advance(dsph_ctx* ctx, EnsState* st, const EnsCase* cases, unsigned ncases);

// Cell keys of all cases (box layout of one case, offset by key0).
int ens_presort(dsph_ctx* ctx, const EnsCase* cases, const EnsBlock* pmap, unsigned npblocks,
                const unsigned* dcell, const uint16_t* code, unsigned* cellpart,
                unsigned* sortpart);
// Current counts per case from its out boxes (NpFinal, NpbOk).
int ens_counts(dsph_ctx* ctx, const EnsCase* cases, EnsState* st, unsigned ncases,
               const int2* begincell);
// Gather of idp, code, dcell, posxy, posz, velrho and PosCell of the case.
int ens_sort_basic_poscell(dsph_ctx* ctx, const EnsCase* cases, const EnsBlock* pmap,
                           unsigned npblocks, const unsigned* sortpart, const unsigned* idp,
                           const uint16_t* code, const unsigned* dcell, const double2* posxy,
                           const double* posz, const float4* velrho, unsigned* idp2,
                           uint16_t* code2, unsigned* dcell2, double2* posxy2, double* posz2,
                           float4* velrho2, float4* poscell);

}  // namespace dsph
