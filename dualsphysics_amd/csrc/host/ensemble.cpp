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
// assignment repeats every 3 steps; a graph therefore holds 3 steps.
#include <hip/hip_runtime.h>
#include <memory>
#include "ensemble.h"
#include "sph_gpu_single.h"

using dsph::EnsBlock;
using dsph::EnsCase;
using dsph::EnsState;

namespace {

constexpr unsigned kGraphSteps = 3;
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
  EnsBlock *d_pmap = nullptr, *d_bmap = nullptr;
  unsigned npblocks = 0, nbblocks = 0;
  struct Bucket {
    EnsBlock* d_wmap = nullptr;
    unsigned nblocks = 0, ncases = 0;
    int variant = dsph::kEnsTileRing12;
    float tuned_ms[dsph::kEnsNumVariants] = {0, 0, 0, 0};
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
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
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
    if (gexec) hipGraphExecDestroy(gexec);
    if (graph) hipGraphDestroy(graph);
    gexec = nullptr;
    graph = nullptr;
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
    build_maps();
    // First divide (JSphGpuSingle::ConfigDomain -> RunCellDivide).
    divide();
    chk(dsph_synchronize(ctx), "first divide");
  }

  void build_maps() {
    std::vector<EnsBlock> pm, bm, wm[kNumBuckets];
    for (unsigned k = 0; k < ncases; k++) {
      const EnsCase& e = hc[k];
      for (unsigned i = 0; i < e.np0; i += dsph::kEnsBs) pm.push_back(make_uint2(k, i));
      for (unsigned i = 0; i < e.npb; i += dsph::kEnsBsBound) bm.push_back(make_uint2(k, i));
      const unsigned npf = e.np0 - e.npb;
      unsigned b = 0;
      while (b < kNumBuckets - 1 && npf >= kBucketEdges[b]) b++;
      for (unsigned i = 0; i < npf; i += 64) wm[b].push_back(make_uint2(k, i));
      buckets[b].ncases++;
    }
    npblocks = unsigned(pm.size());
    nbblocks = unsigned(bm.size());
    d_pmap = alloc<EnsBlock>(pm.size());
    chk(dsph_memcpy_h2d(ctx, d_pmap, pm.data(), sizeof(EnsBlock) * pm.size()), "pmap");
    if (!bm.empty()) {
      d_bmap = alloc<EnsBlock>(bm.size());
      chk(dsph_memcpy_h2d(ctx, d_bmap, bm.data(), sizeof(EnsBlock) * bm.size()), "bmap");
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

  void fluid(unsigned b, int variant) {
    const Bucket& bk = buckets[b];
    chk(dsph::ens_forces_fluid(ctx, variant, tvisco, tddt, eos7, d_cases, d_st, bk.d_wmap,
                               bk.nblocks,
                               dcell, (const float4*)poscell, (const float4*)posf,
                               (const float4*)velrho, (const float2*)press, viscdt, ar, ace),
        "fluid interaction");
  }

  void interaction() {
    prep(true);
    for (unsigned b = 0; b < kNumBuckets; b++)
      if (buckets[b].nblocks) fluid(b, buckets[b].variant);
    chk(dsph::ens_forces_bound(ctx, d_cases, d_st, d_bmap, nbblocks, dcell,
                               (const float4*)poscell, (const float4*)velrho, viscdt, ar),
        "bound interaction");
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

  void capture() {
    drop_graph();
    const bool timing = ctx->timing;
    ctx->timing = false;  // no event records inside the graph
    const std::vector<void*> before = layout();
    hip(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal), "begin capture");
    try {
      for (unsigned s = 0; s < kGraphSteps; s++) step();
    } catch (...) {
      hipGraph_t g = nullptr;
      hipStreamEndCapture(ctx->stream, &g);
      if (g) hipGraphDestroy(g);
      ctx->timing = timing;
      throw;
    }
    hip(hipStreamEndCapture(ctx->stream, &graph), "end capture");
    ctx->timing = timing;
    if (layout() != before) throw Err(DSPH_E_STATE, "graph steps do not restore the array layout");
    hip(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0), "graph instantiate");
  }

  // Runs groups of kGraphSteps steps until every case has stopped or
  // max_steps steps ran; the active flags are read every `check` groups.
  unsigned run(unsigned max_steps, unsigned check) {
    if (use_graph && !gexec) capture();
    unsigned done = 0, groups = 0;
    std::vector<EnsState> st(ncases);
    while (done + kGraphSteps <= max_steps) {
      if (use_graph) {
        hip(hipGraphLaunch(gexec, ctx->stream), "graph launch");
      } else {
        for (unsigned s = 0; s < kGraphSteps; s++) step();
      }
      done += kGraphSteps;
      steps_run += kGraphSteps;
      if (++groups % std::max(check, 1u) == 0 && nactive(st) == 0) break;
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

  // Times every fluid-kernel variant on each size bucket over the current
  // state and keeps the fastest (all variants give identical results).
  void autotune(unsigned reps) {
    drop_graph();
    hipEvent_t e0, e1;
    hip(hipEventCreate(&e0), "event");
    hip(hipEventCreate(&e1), "event");
    prep(false);
    for (unsigned b = 0; b < kNumBuckets; b++) {
      Bucket& bk = buckets[b];
      if (!bk.nblocks) continue;
      float best = 1e30f;
      for (int v = 0; v < dsph::kEnsNumVariants; v++) {
        fluid(b, v);  // warm-up
        float tmin = 1e30f;
        for (unsigned r = 0; r < std::max(reps, 1u); r++) {
          hip(hipEventRecord(e0, ctx->stream), "record");
          fluid(b, v);
          hip(hipEventRecord(e1, ctx->stream), "record");
          hip(hipEventSynchronize(e1), "sync");
          float ms = 0;
          hip(hipEventElapsedTime(&ms, e0, e1), "elapsed");
          tmin = std::min(tmin, ms);
        }
        bk.tuned_ms[v] = tmin;
        if (tmin < best) {
          best = tmin;
          bk.variant = v;
        }
      }
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
  e->drop_graph();
  return DSPH_OK;
}

int dsph_ens_info(dsph_ens* e, uint32_t* h_info) {
  if (!e || !h_info) return DSPH_E_ARG;
  h_info[0] = e->ncases;
  h_info[1] = e->ntot;
  h_info[2] = e->nkeys;
  for (unsigned b = 0; b < kNumBuckets; b++) h_info[3 + b] = e->buckets[b].ncases;
  h_info[6] = kGraphSteps;
  h_info[7] = unsigned(e->steps_run);
  return DSPH_OK;
}

int dsph_ens_set_timing(dsph_ens* e, int enable) {
  if (!e) return DSPH_E_ARG;
  return dsph_ctx_set_timing(e->ctx, enable);
}

dsph_ctx* dsph_ens_context(dsph_ens* e) { return e ? e->ctx : nullptr; }

}  // extern "C"
