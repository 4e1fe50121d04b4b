// Host driver of batched ensembles: many independent cases (each a
// JSphGpuSingle-style Symplectic run with DBC) advanced together, every stage
// one launch for all cases, the whole step captured in a HIP graph.
//
// Per case the constants and the cell map are derived exactly as the single
// driver does (dsph_sim::config); the cell domain is fixed to the case's map,
// time and dt stay on the device (EnsState), and a case stops by itself once
// its time reaches its end time. Results of a case equal, bit for bit, a
// single run of that case with a fixed cell domain and device-side dt: the
// kernels share their bodies with the single-case kernels and visit the same
// neighbours in the same order.
//
// Pointer swaps: one Symplectic step swaps the state arrays so that their
// assignment repeats every 3 steps; a graph therefore holds a multiple of 3
// steps: kLongSteps for the bulk of a run (one graph launch per 24 steps,
// the host only reads the cases' active flags between launches) and 3 for
// the remainder.
//
// Autotuner: per fluid-size bucket it times the fluid kernel's variants (tile
// kernel with a ring of 8, 12 or 16 LDS tiles, one lane per particle rows
// kernel) on a one-workgroup-per-wave grid, then the best tile variant in its
// persistent form over chunks of 1, 2 or 4 waves of particles per counter
// fetch and grids of 10, 20 or 40 waves per CU; and the boundary kernel:
// one lane per particle in blocks of 64, 128 or 256, or the tile form (one
// wave per 64 boundary particles). All shapes give identical results.
#include <hip/hip_runtime.h>
#include <memory>
#include "ensemble.h"
#include "sph_gpu_single.h"

using dsph::EnsBlock;
using dsph::EnsCase;
using dsph::EnsState;

namespace {

constexpr unsigned kGraphSteps = 3;
constexpr unsigned kLongSteps = 8 * kGraphSteps;
constexpr unsigned kNumXcd = 8;
constexpr unsigned kBoundBs[] = {64u, 128u, 256u};
constexpr unsigned kNumBoundBs = 3;
constexpr unsigned kChunks[] = {1u, 2u, 4u};
constexpr unsigned kWavesPerCu[] = {10u, 20u, 40u};
// Fluid-size buckets (fluid particles per case) tuned separately.
constexpr unsigned kBucketEdges[] = {16384u, 131072u};
constexpr unsigned kNumBuckets = 3;

}  // namespace

struct dsph_ens {
  dsph_ctx* ctx = nullptr;
  std::string err;
  unsigned ncases = 0, ntot = 0, nkeys = 0;
  int tvisco = 0, tddt = 0;
  bool eos7 = true;  // every case has gamma = 7
  std::vector<EnsCase> hc;
  EnsCase* d_cases = nullptr;
  EnsState* d_st = nullptr;
  EnsBlock* d_pmap = nullptr;
  unsigned npblocks = 0;
  // Boundary block maps, one per candidate block size.
  EnsBlock* d_bmaps[kNumBoundBs] = {nullptr, nullptr, nullptr};
  unsigned nbblocks[kNumBoundBs] = {0, 0, 0};
  unsigned ibs = 1;  // kBoundBs[ibs] = 128 until tuned; kNumBoundBs = the tile form
  float bound_ms[kNumBoundBs + 1] = {0, 0, 0, 0};
  unsigned ncu = 256;
  unsigned* d_ctr = nullptr;  // kNumXcd counters per bucket (persistent fluid kernel)
  struct Bucket {
    EnsBlock* d_wmap = nullptr;
    unsigned nblocks = 0, ncases = 0;
    int variant = dsph::kEnsTileRing12;
    float tuned_ms[dsph::kEnsNumVariants] = {0, 0, 0, 0};
    // Persistent form (tile variants only): grid of wpc waves per CU, chunk
    // work items per fetch; persist_ms = its tuned time.
    bool persist = false;
    unsigned chunk = 1, wpc = 20;
    float persist_ms = 0;
  };
  Bucket buckets[kNumBuckets];
  unsigned *idp = nullptr, *idp2 = nullptr, *dcell = nullptr, *dcell2 = nullptr;
  unsigned *cellpart = nullptr, *sortpart = nullptr;
  uint16_t *code = nullptr, *code2 = nullptr;
  double *posxy = nullptr, *posxy2 = nullptr, *posz = nullptr, *posz2 = nullptr;
  double *posxypre = nullptr, *posxypre2 = nullptr, *poszpre = nullptr, *poszpre2 = nullptr;
  double *movxy = nullptr, *movz = nullptr;
  float *velrho = nullptr, *velrho2 = nullptr, *velrhopre = nullptr, *velrhopre2 = nullptr;
  float *poscell = nullptr, *ace = nullptr, *ar = nullptr, *viscdt = nullptr;
  float *press = nullptr, *posf = nullptr;
  int* begincell = nullptr;
  bool pre_active = false;
  bool use_graph = true;
  // [0]: kGraphSteps steps, [1]: kLongSteps steps.
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  std::vector<void*> allocs;
  unsigned long long steps_run = 0;

  ~dsph_ens() { release(); }

  void chk(int rc, const char* what) {
    if (rc != DSPH_OK) throw Err(rc, std::string(what) + ": " + dsph_last_error(ctx));
  }
  void hip(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Err(DSPH_E_LAUNCH, std::string(what) + ": " + hipGetErrorString(e));
  }
  template <typename T>
  T* alloc(size_t n) {
    void* p = nullptr;
    chk(dsph_malloc(ctx, sizeof(T) * std::max<size_t>(n, 1), &p), "dsph_malloc");
    allocs.push_back(p);
    return static_cast<T*>(p);
  }
  void drop_graph() {
    for (int g = 0; g < 2; g++) {
      if (gexec[g]) hipGraphExecDestroy(gexec[g]);
      if (graph[g]) hipGraphDestroy(graph[g]);
      gexec[g] = nullptr;
      graph[g] = nullptr;
    }
  }
  void release() {
    if (!ctx) return;
    dsph_synchronize(ctx);
    drop_graph();
    for (void* p : allocs) dsph_free(ctx, p);
    allocs.clear();
    dsph_ctx_destroy(ctx);
    ctx = nullptr;
  }

  // ---- set-up
  void setup(const dsph_case* cases, const uint32_t* np, const uint32_t* h_idp,
             const double* h_pos, const float* h_velrho, const double* tend) {
    hc.resize(ncases);
    std::vector<uint16_t> hcode;
    std::vector<unsigned> hdcell;
    std::vector<double> hxy, hz;
    unsigned p0 = 0, key0 = 0;
    for (unsigned k = 0; k < ncases; k++) {
      const dsph_case& cs = cases[k];
      if (cs.tstep != 2) throw Err(DSPH_E_ARG, "ensembles run the Symplectic integrator");
      if (cs.tboundary != DSPH_BC_DBC) throw Err(DSPH_E_ARG, "ensembles run DBC boundaries");
      if (cs.shiftmode != DSPH_SHIFT_NONE) throw Err(DSPH_E_ARG, "ensembles run without shifting");
      if (cs.periactive) throw Err(DSPH_E_ARG, "ensembles run without periodic boundaries");
      if (cs.simulate2d) throw Err(DSPH_E_ARG, "ensembles run 3-D cases");
      if (cs.casenmoving || cs.casenfloat) throw Err(DSPH_E_ARG, "ensembles have fixed boundaries");
      if (k == 0) { tvisco = cs.tvisco; tddt = cs.tdensity; }
      eos7 = eos7 && float(cs.gamma) == 7.f;
      if (cs.tvisco != tvisco || cs.tdensity != tddt)
        throw Err(DSPH_E_ARG, "all cases of an ensemble share the viscosity and density options");
      if (!(tvisco == DSPH_VISCO_ARTIFICIAL || tvisco == DSPH_VISCO_LAMINAR) ||
          !(tddt == DSPH_DDT_NONE || tddt == DSPH_DDT_DDT2FULL))
        throw Err(DSPH_E_ARG, "ensembles are built for artificial or laminar viscosity with "
                              "DDT2Full or no density diffusion");
      if (np[k] != cs.casenp || np[k] <= cs.casenbound) throw Err(DSPH_E_ARG, "invalid particle count");
      dsph_sim g;  // constants and cell map only (no device state)
      g.config(cs);
      const dsph_cte cte = g.make_cte();
      EnsCase& e = hc[k];
      std::memset(&e, 0, sizeof(e));
      e.c = dsph::make_cte(cte);
      const unsigned ncx = g.DomCells.x, ncy = g.DomCells.y, ncz = g.DomCells.z;
      const unsigned nct = ncx * ncy * ncz;
      e.ncx = ncx; e.ncy = ncy; e.ncz = ncz; e.nct = nct;
      e.key0 = key0;
      e.dvd.scelldiv = g.ScellDiv;
      e.dvd.ncx = int(ncx); e.dvd.ncy = int(ncy); e.dvd.ncz = int(ncz); e.dvd.ncw = int(ncx * ncy);
      e.dvd.cellfluid = nct + 1;
      e.dvd.czx = e.dvd.czy = e.dvd.czz = 0;
      e.dvd.begincell = nullptr;  // set after allocation
      e.dvd.domcellcode = g.DomCellCode;
      e.dvd.kernelsize2 = g.KernelSize2;
      e.dvd.poscellsize = g.PosCellSize;
      e.dvd.domposminx = g.DomPosMin[0]; e.dvd.domposminy = g.DomPosMin[1];
      e.dvd.domposminz = g.DomPosMin[2];
      e.dvd.scell = g.Scell;
      e.pc.movlimit = g.MovLimit;
      e.pc.mapx = g.MapRealPosMin[0]; e.pc.mapy = g.MapRealPosMin[1]; e.pc.mapz = g.MapRealPosMin[2];
      e.pc.sizex = g.MapRealSize[0]; e.pc.sizey = g.MapRealSize[1]; e.pc.sizez = g.MapRealSize[2];
      e.pc.scell = g.Scell;
      e.pc.cellcode = g.DomCellCode;
      e.posminx = g.Map_PosMin[0]; e.posminy = g.Map_PosMin[1]; e.posminz = g.Map_PosMin[2];
      e.gx = g.Gravity[0]; e.gy = g.Gravity[1]; e.gz = g.Gravity[2];
      e.rhopoutmin = g.RhopOutMin; e.rhopoutmax = g.RhopOutMax;
      e.viscob = g.Visco * g.ViscoBoundFactor;
      e.viscof = g.Visco;
      e.ks2scan = dsph::ens_scan_radius2(cte);
      e.kernelh = double(g.KernelH);
      e.cs0 = g.Cs0;
      e.cfl = g.CFLnumber;
      e.dtmin = g.DtMin;
      e.tend = tend[k];
      e.p0 = p0;
      e.np0 = np[k];
      e.npb = g.CaseNpb;
      // LoadCodeParticles / LoadDcellParticles as the single driver.
      for (unsigned i = 0; i < np[k]; i++) {
        const size_t q = size_t(p0) + i;
        const unsigned id = h_idp[q];
        hcode.push_back(uint16_t(id < cs.casenfixed ? 0x0000 : (id < g.CaseNbound ? 0x0801 : 0x1800)));
        const double x = h_pos[3 * q], y = h_pos[3 * q + 1], z = h_pos[3 * q + 2];
        const bool in = x >= g.MapRealPosMin[0] && y >= g.MapRealPosMin[1] &&
                        z >= g.MapRealPosMin[2] && x < g.MapRealPosMax[0] &&
                        y < g.MapRealPosMax[1] && z < g.MapRealPosMax[2];
        if (!in) throw Err(DSPH_E_ARG, "Found new particles out.");
        hxy.push_back(x);
        hxy.push_back(y);
        hz.push_back(z);
        hdcell.push_back(dcel_cell(g.DomCellCode, unsigned((x - g.DomPosMin[0]) / g.Scell),
                                   unsigned((y - g.DomPosMin[1]) / g.Scell),
                                   unsigned((z - g.DomPosMin[2]) / g.Scell)));
      }
      if (size_t(key0) + 2ull * nct + 5 > 0xffffffffull || size_t(p0) + np[k] > 0x7fffffffull)
        throw Err(DSPH_E_ARG, "ensemble too large for 32-bit keys");
      key0 += 2 * nct + 5;
      p0 += np[k];
    }
    ntot = p0;
    nkeys = key0;
    const size_t n = ntot;
    idp = alloc<unsigned>(n); idp2 = alloc<unsigned>(n);
    dcell = alloc<unsigned>(n); dcell2 = alloc<unsigned>(n);
    cellpart = alloc<unsigned>(n); sortpart = alloc<unsigned>(n);
    code = alloc<uint16_t>(n); code2 = alloc<uint16_t>(n);
    posxy = alloc<double>(2 * n); posxy2 = alloc<double>(2 * n);
    posz = alloc<double>(n); posz2 = alloc<double>(n);
    posxypre = alloc<double>(2 * n); posxypre2 = alloc<double>(2 * n);
    poszpre = alloc<double>(n); poszpre2 = alloc<double>(n);
    movxy = alloc<double>(2 * n); movz = alloc<double>(n);
    velrho = alloc<float>(4 * n); velrho2 = alloc<float>(4 * n);
    velrhopre = alloc<float>(4 * n); velrhopre2 = alloc<float>(4 * n);
    poscell = alloc<float>(4 * n);
    ace = alloc<float>(3 * n); ar = alloc<float>(n); viscdt = alloc<float>(n);
    begincell = alloc<int>(2 * size_t(nkeys));
    for (auto& e : hc)
      e.dvd.begincell = reinterpret_cast<const int2*>(begincell) + e.key0;
    // Sort, reduction and interaction scratch of the context ({press, 1/rho}
    // and map-frame positions per particle).
    chk(dsph_ctx_reserve(ctx, ntot, nkeys), "dsph_ctx_reserve");
    press = ctx->prep_press;
    posf = ctx->prep_posf;
    chk(dsph_memcpy_h2d(ctx, idp, h_idp, 4 * n), "upload idp");
    chk(dsph_memcpy_h2d(ctx, code, hcode.data(), 2 * n), "upload code");
    chk(dsph_memcpy_h2d(ctx, dcell, hdcell.data(), 4 * n), "upload dcell");
    chk(dsph_memcpy_h2d(ctx, posxy, hxy.data(), 16 * n), "upload posxy");
    chk(dsph_memcpy_h2d(ctx, posz, hz.data(), 8 * n), "upload posz");
    chk(dsph_memcpy_h2d(ctx, velrho, h_velrho, 16 * n), "upload velrho");
    d_cases = alloc<EnsCase>(ncases);
    chk(dsph_memcpy_h2d(ctx, d_cases, hc.data(), sizeof(EnsCase) * ncases), "upload cases");
    std::vector<EnsState> st(ncases);
    for (unsigned k = 0; k < ncases; k++) {
      std::memset(&st[k], 0, sizeof(EnsState));
      dsph_sim g;
      g.config(cases[k]);
      st[k].dtpre = g.DtIni;
      st[k].np = hc[k].np0;
      st[k].npbok = hc[k].npb;
      st[k].active = tend[k] > 0 ? 1u : 0u;
    }
    d_st = alloc<EnsState>(ncases);
    chk(dsph_memcpy_h2d(ctx, d_st, st.data(), sizeof(EnsState) * ncases), "upload state");
    d_ctr = alloc<unsigned>(kNumBuckets * kNumXcd);
    int cus = 0;
    hip(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device),
        "compute units");
    ncu = unsigned(std::max(cus, 1));
    build_maps();
    // First divide (JSphGpuSingle::ConfigDomain -> RunCellDivide).
    divide();
    chk(dsph_synchronize(ctx), "first divide");
  }

  void build_maps() {
    std::vector<EnsBlock> pm, bm[kNumBoundBs], wm[kNumBuckets];
    for (unsigned k = 0; k < ncases; k++) {
      const EnsCase& e = hc[k];
      for (unsigned i = 0; i < e.np0; i += dsph::kEnsBs) pm.push_back(make_uint2(k, i));
      for (unsigned s = 0; s < kNumBoundBs; s++)
        for (unsigned i = 0; i < e.npb; i += kBoundBs[s]) bm[s].push_back(make_uint2(k, i));
      const unsigned npf = e.np0 - e.npb;
      unsigned b = 0;
      while (b < kNumBuckets - 1 && npf >= kBucketEdges[b]) b++;
      for (unsigned i = 0; i < npf; i += 64) wm[b].push_back(make_uint2(k, i));
      buckets[b].ncases++;
    }
    npblocks = unsigned(pm.size());
    d_pmap = alloc<EnsBlock>(pm.size());
    chk(dsph_memcpy_h2d(ctx, d_pmap, pm.data(), sizeof(EnsBlock) * pm.size()), "pmap");
    for (unsigned s = 0; s < kNumBoundBs; s++) {
      nbblocks[s] = unsigned(bm[s].size());
      if (bm[s].empty()) continue;
      d_bmaps[s] = alloc<EnsBlock>(bm[s].size());
      chk(dsph_memcpy_h2d(ctx, d_bmaps[s], bm[s].data(), sizeof(EnsBlock) * bm[s].size()),
          "bmap");
    }
    for (unsigned b = 0; b < kNumBuckets; b++) {
      buckets[b].nblocks = unsigned(wm[b].size());
      if (wm[b].empty()) continue;
      buckets[b].d_wmap = alloc<EnsBlock>(wm[b].size());
      chk(dsph_memcpy_h2d(ctx, buckets[b].d_wmap, wm[b].data(), sizeof(EnsBlock) * wm[b].size()),
          "wmap");
    }
  }

  template <typename T>
  static void swp(T*& a, T*& b) { std::swap(a, b); }

  // ---- one step (the single driver's order: JSphGpuSingle::ComputeStep_Sym
  // followed by RunCellDivide), no host synchronisation
  void divide() {
    chk(dsph::ens_presort(ctx, d_cases, d_pmap, npblocks, dcell, code, cellpart, sortpart),
        "presort");
    chk(dsph_sort_by_cell(ctx, cellpart, sortpart, ntot, nkeys), "sort");
    chk(dsph_begin_end_cell(ctx, 1, ntot, 0, nkeys, 0, cellpart, begincell), "begin end cell");
    chk(dsph::ens_counts(ctx, d_cases, d_st, ncases, reinterpret_cast<const int2*>(begincell)),
        "counts");
    chk(dsph::ens_sort_basic_poscell(ctx, d_cases, d_pmap, npblocks, sortpart, idp, code, dcell,
                                     (const double2*)posxy, posz, (const float4*)velrho, idp2,
                                     code2, dcell2, (double2*)posxy2, posz2, (float4*)velrho2,
                                     (float4*)poscell),
        "sort basic");
    swp(idp, idp2); swp(code, code2); swp(dcell, dcell2); swp(posxy, posxy2); swp(posz, posz2);
    swp(velrho, velrho2);
    if (pre_active) {
      chk(dsph_sort_data_d2_d_f4(ctx, ntot, 0, sortpart, posxypre, poszpre, velrhopre, posxypre2,
                                 poszpre2, velrhopre2),
          "sort pre");
      swp(posxypre, posxypre2); swp(poszpre, poszpre2); swp(velrhopre, velrhopre2);
    }
  }

  // PreInteraction_Forces: zero the forces, VelMax (with_velmax), EOS prep.
  void prep(bool with_velmax) {
    chk(dsph::ens_zero_forces(ctx, ntot, ace, ar, viscdt), "zero");
    if (with_velmax)
      chk(dsph::ens_reduce(ctx, d_cases, d_st, d_pmap, npblocks, 0, (const float4*)velrho,
                           viscdt, ace),
          "velmax");
    chk(dsph::ens_prep(ctx, d_cases, d_st, d_pmap, npblocks, (const float4*)poscell,
                       (const float4*)velrho, (float2*)press, (float4*)posf),
        "prep");
  }

  // Fluid interaction of bucket b: the variant on a one-workgroup-per-wave
  // grid, or persistent with (chunk, wpc) when chunk != 0.
  void fluid(unsigned b, int variant, unsigned chunk = 0, unsigned wpc = 0) {
    const Bucket& bk = buckets[b];
    dsph::EnsPersist pers{0, chunk, d_ctr + b * kNumXcd};
    if (chunk) {
      const unsigned need = (bk.nblocks + chunk - 1) / chunk;
      pers.grid = std::max(1u, std::min(ncu * wpc, need));
    }
    chk(dsph::ens_forces_fluid(ctx, variant, tvisco, tddt, eos7, d_cases, d_st, bk.d_wmap,
                               bk.nblocks, chunk ? &pers : nullptr,
                               dcell, (const float4*)poscell, (const float4*)posf,
                               (const float4*)velrho, (const float2*)press, viscdt, ar, ace),
        "fluid interaction");
  }
  void fluid_tuned(unsigned b) {
    const Bucket& bk = buckets[b];
    if (bk.persist) fluid(b, bk.variant, bk.chunk, bk.wpc);
    else fluid(b, bk.variant);
  }
  // Boundary interaction: s < kNumBoundBs one lane per particle in blocks
  // of kBoundBs[s]; s == kNumBoundBs the tile form over the 64-blocks.
  void bound(unsigned s) {
    if (s == kNumBoundBs)
      chk(dsph::ens_forces_bound_tile(ctx, d_cases, d_st, d_bmaps[0], nbblocks[0], dcell,
                                      (const float4*)poscell, (const float4*)posf,
                                      (const float4*)velrho, (const float2*)press, viscdt, ar),
          "bound interaction");
    else
      chk(dsph::ens_forces_bound(ctx, d_cases, d_st, d_bmaps[s], nbblocks[s], kBoundBs[s],
                                 dcell, (const float4*)poscell, (const float4*)velrho, viscdt,
                                 ar),
          "bound interaction");
  }

  void interaction() {
    prep(true);
    for (unsigned b = 0; b < kNumBuckets; b++)
      if (buckets[b].nblocks) fluid_tuned(b);
    bound(ibs);
    chk(dsph::ens_reduce(ctx, d_cases, d_st, d_pmap, npblocks, 1, (const float4*)velrho, viscdt,
                         ace),
        "viscdtmax");
    chk(dsph::ens_reduce(ctx, d_cases, d_st, d_pmap, npblocks, 2, (const float4*)velrho, viscdt,
                         ace),
        "acemax");
  }

  void step() {
    interaction();
    chk(dsph::ens_dt(ctx, d_cases, d_st, ncases, 0), "dt predictor");
    swp(posxypre, posxy); swp(poszpre, posz); swp(velrhopre, velrho);
    pre_active = true;
    chk(dsph::ens_sym_pre(ctx, d_cases, d_st, d_pmap, npblocks, (const float4*)velrhopre, ar, ace,
                          code, (double2*)movxy, movz, (float4*)velrho),
        "symplectic pre");
    chk(dsph::ens_pos2(ctx, d_cases, d_st, d_pmap, npblocks, true, (const double2*)posxypre,
                       poszpre, (const double2*)movxy, movz, (double2*)posxy, posz, dcell, code),
        "pos2");
    divide();
    interaction();
    chk(dsph::ens_dt(ctx, d_cases, d_st, ncases, 1), "dt corrector");
    chk(dsph::ens_sym_cor(ctx, d_cases, d_st, d_pmap, npblocks, (const float4*)velrhopre, ar, ace,
                          code, (double2*)movxy, movz, (float4*)velrho),
        "symplectic cor");
    chk(dsph::ens_pos2(ctx, d_cases, d_st, d_pmap, npblocks, false, (const double2*)posxypre,
                       poszpre, (const double2*)movxy, movz, (double2*)posxy, posz, dcell, code),
        "pos2");
    pre_active = false;
    chk(dsph::ens_advance(ctx, d_st, d_cases, ncases), "advance");
    divide();
  }

  // Pointer assignment, to check that kGraphSteps steps restore it.
  std::vector<void*> layout() const {
    return {idp, idp2, dcell, dcell2, code, code2, posxy, posxy2, posz, posz2, posxypre,
            posxypre2, poszpre, poszpre2, velrho, velrho2, velrhopre, velrhopre2};
  }

  // Captures graph g (0: kGraphSteps steps, 1: kLongSteps steps).
  void capture(int g) {
    const unsigned nsteps = g ? kLongSteps : kGraphSteps;
    const bool timing = ctx->timing;
    ctx->timing = false;  // no event records inside the graph
    const std::vector<void*> before = layout();
    hip(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal), "begin capture");
    try {
      for (unsigned s = 0; s < nsteps; s++) step();
    } catch (...) {
      hipGraph_t gr = nullptr;
      hipStreamEndCapture(ctx->stream, &gr);
      if (gr) hipGraphDestroy(gr);
      ctx->timing = timing;
      throw;
    }
    hip(hipStreamEndCapture(ctx->stream, &graph[g]), "end capture");
    ctx->timing = timing;
    if (layout() != before) throw Err(DSPH_E_STATE, "graph steps do not restore the array layout");
    hip(hipGraphInstantiate(&gexec[g], graph[g], nullptr, nullptr, 0), "graph instantiate");
  }

  // Runs steps in graph launches of kLongSteps (kGraphSteps for the rest)
  // until every case has stopped or max_steps steps ran (rounded down to a
  // multiple of kGraphSteps); the active flags are read once at least
  // check * kGraphSteps steps ran since the last reading.
  unsigned run(unsigned max_steps, unsigned check) {
    if (use_graph) {
      // Both graphs are captured on entry (run(0) captures without running).
      if (!gexec[0]) capture(0);
      if (!gexec[1]) capture(1);
    }
    unsigned done = 0, since = 0;
    std::vector<EnsState> st(ncases);
    while (done + kGraphSteps <= max_steps) {
      const int g = done + kLongSteps <= max_steps ? 1 : 0;
      const unsigned n = g ? kLongSteps : kGraphSteps;
      if (use_graph) {
        hip(hipGraphLaunch(gexec[g], ctx->stream), "graph launch");
      } else {
        for (unsigned s = 0; s < n; s++) step();
      }
      done += n;
      steps_run += n;
      since += n;
      if (since >= std::max(check, 1u) * kGraphSteps) {
        since = 0;
        if (nactive(st) == 0) break;
      }
    }
    chk(dsph_synchronize(ctx), "synchronize");
    return done;
  }

  unsigned nactive(std::vector<EnsState>& st) {
    chk(dsph_memcpy_d2h(ctx, st.data(), d_st, sizeof(EnsState) * ncases), "state readback");
    unsigned n = 0;
    for (const auto& s : st) n += s.active ? 1u : 0u;
    return n;
  }

  // Times the shapes of the fluid kernel on each size bucket and of the
  // boundary kernel over the current state and keeps the fastest (every
  // shape gives identical results): the variants on a one-workgroup-per-wave
  // grid, then the best tile variant persistent over chunks x grids.
  void autotune(unsigned reps) {
    drop_graph();
    hipEvent_t e0, e1;
    hip(hipEventCreate(&e0), "event");
    hip(hipEventCreate(&e1), "event");
    auto timed = [&](auto&& launch) {
      launch();  // warm-up
      float tmin = 1e30f;
      for (unsigned r = 0; r < std::max(reps, 1u); r++) {
        hip(hipEventRecord(e0, ctx->stream), "record");
        launch();
        hip(hipEventRecord(e1, ctx->stream), "record");
        hip(hipEventSynchronize(e1), "sync");
        float ms = 0;
        hip(hipEventElapsedTime(&ms, e0, e1), "elapsed");
        tmin = std::min(tmin, ms);
      }
      return tmin;
    };
    try {
      prep(false);
      for (unsigned b = 0; b < kNumBuckets; b++) {
        Bucket& bk = buckets[b];
        if (!bk.nblocks) continue;
        float best = 1e30f, best_tile = 1e30f;
        int tile = dsph::kEnsTileRing12;
        for (int v = 0; v < dsph::kEnsNumVariants; v++) {
          const float ms = timed([&] { fluid(b, v); });
          bk.tuned_ms[v] = ms;
          if (ms < best) {
            best = ms;
            bk.variant = v;
          }
          if (v != dsph::kEnsRows64 && ms < best_tile) {
            best_tile = ms;
            tile = v;
          }
        }
        bk.persist = false;
        bk.persist_ms = 1e30f;
        unsigned pc = 1, pw = 20;
        for (unsigned ch : kChunks)
          for (unsigned w : kWavesPerCu) {
            const float ms = timed([&] { fluid(b, tile, ch, w); });
            if (ms < bk.persist_ms) {
              bk.persist_ms = ms;
              pc = ch;
              pw = w;
            }
          }
        bk.chunk = pc;
        bk.wpc = pw;
        if (bk.persist_ms < best) {
          bk.persist = true;
          bk.variant = tile;
        }
      }
      float bbest = 1e30f;
      for (unsigned s = 0; s <= kNumBoundBs; s++) {
        if (!nbblocks[s < kNumBoundBs ? s : 0]) continue;
        bound_ms[s] = timed([&] { bound(s); });
        if (bound_ms[s] < bbest) {
          bbest = bound_ms[s];
          ibs = s;
        }
      }
    } catch (...) {
      hipEventDestroy(e0);
      hipEventDestroy(e1);
      throw;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    // The forces of the trials are discarded: every step starts by zeroing them.
  }
};

extern "C" {

int dsph_ens_create(int device, uint32_t ncases, const dsph_case* cases, const uint32_t* np,
                    const uint32_t* h_idp, const double* h_pos, const float* h_velrho,
                    const double* tend, dsph_ens** out) {
  if (!out) return DSPH_E_ARG;
  *out = nullptr;
  if (!ncases || !cases || !np || !h_idp || !h_pos || !h_velrho || !tend) return DSPH_E_ARG;
  uint64_t ntot = 0;
  for (uint32_t k = 0; k < ncases; k++) ntot += np[k];
  if (ntot >= DSPH_NP_LIMIT) return DSPH_E_ARG;
  dsph_ens* e = new dsph_ens();
  int rc = dsph_ctx_create(device, &e->ctx);
  if (rc) {
    delete e;
    return rc;
  }
  e->ncases = ncases;
  try {
    e->setup(cases, np, h_idp, h_pos, h_velrho, tend);
  } catch (const Err& x) {
    e->err = x.what();
    *out = e;
    return x.code;
  }
  *out = e;
  return DSPH_OK;
}

int dsph_ens_destroy(dsph_ens* e) {
  delete e;
  return DSPH_OK;
}

const char* dsph_ens_last_error(dsph_ens* e) { return e ? e->err.c_str() : "null ensemble"; }

int dsph_ens_set_graph(dsph_ens* e, int enable) {
  if (!e) return DSPH_E_ARG;
  e->use_graph = enable != 0;
  if (!e->use_graph) e->drop_graph();
  return DSPH_OK;
}

int dsph_ens_run(dsph_ens* e, uint32_t max_steps, uint32_t check_every, uint32_t* h_steps) {
  if (!e) return DSPH_E_ARG;
  try {
    const unsigned n = e->run(max_steps, check_every);
    if (h_steps) *h_steps = n;
  } catch (const Err& x) {
    e->err = x.what();
    return x.code;
  }
  return DSPH_OK;
}

int dsph_ens_case_state(dsph_ens* e, uint32_t icase, dsph_ens_state* out) {
  if (!e || !out || icase >= e->ncases) return DSPH_E_ARG;
  EnsState s;
  int rc = dsph_memcpy_d2h(e->ctx, &s, e->d_st + icase, sizeof(s));
  if (rc) {
    e->err = dsph_last_error(e->ctx);
    return rc;
  }
  out->np = s.np;
  out->npbok = s.npbok;
  out->active = s.active;
  out->nstep = s.nstep;
  out->flags = s.flags;
  out->time = s.time;
  out->dt_last = s.dtlast;
  out->dt_next = s.dtpre;
  out->p0 = e->hc[icase].p0;
  out->np0 = e->hc[icase].np0;
  return DSPH_OK;
}

int dsph_ens_download(dsph_ens* e, uint32_t icase, uint32_t* h_count, uint32_t* h_idp,
                      double* h_pos, float* h_velrho, uint16_t* h_code) {
  if (!e || icase >= e->ncases || !h_count) return DSPH_E_ARG;
  try {
    EnsState s;
    e->chk(dsph_memcpy_d2h(e->ctx, &s, e->d_st + icase, sizeof(s)), "state");
    const size_t p0 = e->hc[icase].p0, n = s.np;
    *h_count = unsigned(n);
    if (h_idp) e->chk(dsph_memcpy_d2h(e->ctx, h_idp, e->idp + p0, 4 * n), "idp");
    if (h_velrho) e->chk(dsph_memcpy_d2h(e->ctx, h_velrho, e->velrho + 4 * p0, 16 * n), "velrho");
    if (h_code) e->chk(dsph_memcpy_d2h(e->ctx, h_code, e->code + p0, 2 * n), "code");
    if (h_pos) {
      std::vector<double> xy(2 * n), z(n);
      e->chk(dsph_memcpy_d2h(e->ctx, xy.data(), e->posxy + 2 * p0, 16 * n), "posxy");
      e->chk(dsph_memcpy_d2h(e->ctx, z.data(), e->posz + p0, 8 * n), "posz");
      for (size_t i = 0; i < n; i++) {
        h_pos[3 * i] = xy[2 * i];
        h_pos[3 * i + 1] = xy[2 * i + 1];
        h_pos[3 * i + 2] = z[i];
      }
    }
  } catch (const Err& x) {
    e->err = x.what();
    return x.code;
  }
  return DSPH_OK;
}

int dsph_ens_autotune(dsph_ens* e, uint32_t reps, int32_t* h_variant, float* h_ms) {
  if (!e) return DSPH_E_ARG;
  try {
    e->autotune(reps);
  } catch (const Err& x) {
    e->err = x.what();
    return x.code;
  }
  for (unsigned b = 0; b < kNumBuckets; b++) {
    if (h_variant) h_variant[b] = e->buckets[b].nblocks ? e->buckets[b].variant : -1;
    if (h_ms)
      for (int v = 0; v < dsph::kEnsNumVariants; v++)
        h_ms[b * dsph::kEnsNumVariants + v] = e->buckets[b].tuned_ms[v];
  }
  return DSPH_OK;
}

int dsph_ens_set_variant(dsph_ens* e, uint32_t bucket, int32_t variant) {
  if (!e || bucket >= kNumBuckets || variant < 0 || variant >= dsph::kEnsNumVariants)
    return DSPH_E_ARG;
  e->buckets[bucket].variant = variant;
  e->buckets[bucket].persist = false;
  e->drop_graph();
  return DSPH_OK;
}

int dsph_ens_get_shape(dsph_ens* e, uint32_t bucket, dsph_ens_shape* out) {
  if (!e || !out || bucket >= kNumBuckets) return DSPH_E_ARG;
  const auto& bk = e->buckets[bucket];
  out->variant = bk.nblocks ? bk.variant : -1;
  out->persistent = bk.persist ? 1u : 0u;
  out->chunk = bk.chunk;
  out->waves_per_cu = bk.wpc;
  out->bound_block = e->ibs < kNumBoundBs ? kBoundBs[e->ibs] : 64;
  out->bound_tile = e->ibs == kNumBoundBs ? 1u : 0u;
  float plain = 0;
  for (int v = 0; v < dsph::kEnsNumVariants; v++)
    if (bk.tuned_ms[v] > 0 && (plain == 0 || bk.tuned_ms[v] < plain)) plain = bk.tuned_ms[v];
  out->ms_plain = plain;
  out->ms_persistent = bk.persist_ms < 1e30f ? bk.persist_ms : 0.f;
  out->ms_bound = e->bound_ms[e->ibs];
  return DSPH_OK;
}

int dsph_ens_set_shape(dsph_ens* e, uint32_t bucket, const dsph_ens_shape* shape) {
  if (!e || !shape || bucket >= kNumBuckets) return DSPH_E_ARG;
  if (shape->variant < 0 || shape->variant >= dsph::kEnsNumVariants) {
    e->err = "dsph_ens_set_shape: invalid variant";
    return DSPH_E_ARG;
  }
  if (shape->persistent &&
      (shape->variant == dsph::kEnsRows64 || shape->chunk < 1 || shape->chunk > 64 ||
       shape->waves_per_cu < 1 || shape->waves_per_cu > 64)) {
    e->err = "dsph_ens_set_shape: persistent launches run a tile variant with chunk and "
             "waves_per_cu in 1..64";
    return DSPH_E_ARG;
  }
  unsigned s = kNumBoundBs + 1;
  for (unsigned k = 0; k < kNumBoundBs; k++)
    if (kBoundBs[k] == shape->bound_block) s = k;
  if (shape->bound_tile) s = (shape->bound_block == 64 ? kNumBoundBs : kNumBoundBs + 1);
  if (s > kNumBoundBs) {
    e->err = "dsph_ens_set_shape: bound_block must be 64, 128 or 256 (64 with bound_tile)";
    return DSPH_E_ARG;
  }
  auto& bk = e->buckets[bucket];
  bk.variant = shape->variant;
  bk.persist = shape->persistent != 0;
  if (bk.persist) {
    bk.chunk = shape->chunk;
    bk.wpc = shape->waves_per_cu;
  }
  e->ibs = s;
  e->drop_graph();
  return DSPH_OK;
}

int dsph_ens_info(dsph_ens* e, uint32_t* h_info) {
  if (!e || !h_info) return DSPH_E_ARG;
  h_info[0] = e->ncases;
  h_info[1] = e->ntot;
  h_info[2] = e->nkeys;
  for (unsigned b = 0; b < kNumBuckets; b++) h_info[3 + b] = e->buckets[b].ncases;
  h_info[6] = kLongSteps;
  h_info[7] = unsigned(e->steps_run);
  return DSPH_OK;
}

int dsph_ens_set_timing(dsph_ens* e, int enable) {
  if (!e) return DSPH_E_ARG;
  return dsph_ctx_set_timing(e->ctx, enable);
}

dsph_ctx* dsph_ens_context(dsph_ens* e) { return e ? e->ctx : nullptr; }

}  // extern "C"
