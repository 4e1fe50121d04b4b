// Internal context of the C ABI: device, stream, constants and scratch.
#pragma once
#include <hip/hip_runtime.h>
#include <map>
#include <string>
#include <utility>
#include <vector>
#include "dsph_common.h"

struct dsph_ktime {
  unsigned long long count = 0;
  double ms = 0;
};

struct dsph_ctx {
  // Per-kernel timing with HIP events on the context stream (off by default).
  bool timing = false;
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> ev_pending;
  std::map<std::string, dsph_ktime> ktimes;
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t own_stream = nullptr;
  std::string err;
  dsph_cte cte{};
  bool cte_set = false;
  // Radix-sort scratch for up to sort_cap keys (layout: sort_scratch_words):
  // two alternate key/value buffers and the digit histograms.
  uint32_t* sort_buf = nullptr;
  size_t sort_cap = 0;
  // Reduction partials (float) and a small device result area.
  float* redu = nullptr;
  size_t redu_cap = 0;
  uint32_t* small_dev = nullptr;   // 64 words: limits, pips counters
  void* pinned = nullptr;          // 4 KiB pinned host buffer for readbacks
  // Per-interaction preparation: pressure and map-frame float positions.
  float* prep_press = nullptr;     // float2 {press, 1/rho} per particle
  float* prep_posf = nullptr;      // float4 per particle
  size_t prep_cap = 0;
  int forces_kernel = 0;           // DSPH_OPT_FORCES_KERNEL: 0 queue, 1 rows
  // DSPH_OPT_BOUND_KERNEL: 0 auto, 1 tile, 2 one lane per particle. Auto
  // times both forms (identical bits) on interactions 1 and 3 of every
  // kBoundRetrial (0, 2 are warm-ups) and keeps the faster until the next
  // trial; bound_choice is the form in use.
  int bound_kernel = 0;
  int bound_choice = 1;
  unsigned long long bound_count = 0;
  hipEvent_t bound_ev[4] = {nullptr, nullptr, nullptr, nullptr};
  // Device address of the static scratch of dsph_interaction_forces_noalloc
  // on this context's device (set by dsph_ctx_create).
  void* noalloc_scratch = nullptr;
};

namespace dsph {

int set_error(dsph_ctx* ctx, int code, const std::string& msg);

// Checks a HIP call; on failure records the error and returns DSPH_E_LAUNCH.
#define DSPH_HIP(ctx, call)                                                  \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess)                                                    \
      return dsph::set_error((ctx), DSPH_E_LAUNCH,                           \
                             std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Checks the launch of the last kernel.
#define DSPH_LAUNCHED(ctx, what)                                             \
  do {                                                                       \
    hipError_t e_ = hipGetLastError();                                       \
    if (e_ != hipSuccess)                                                    \
      return dsph::set_error((ctx), DSPH_E_LAUNCH,                           \
                             std::string(what) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define DSPH_REQUIRE(ctx, cond, msg)                                         \
  do {                                                                       \
    if (!(cond)) return dsph::set_error((ctx), DSPH_E_ARG, (msg));           \
  } while (0)

inline unsigned grid_for(size_t n, unsigned bs) { return unsigned((n + bs - 1) / bs); }

// Records start/stop events around the launches in its scope when the
// context's timing is on; resolved by dsph_ctx_timing_collect.
struct KTimer {
  dsph_ctx* ctx;
  const char* name;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  static hipEvent_t take(dsph_ctx* c) {
    hipEvent_t e = nullptr;
    if (!c->ev_pool.empty()) {
      e = c->ev_pool.back();
      c->ev_pool.pop_back();
    } else if (hipEventCreate(&e) != hipSuccess) {
      e = nullptr;
    }
    return e;
  }
  KTimer(dsph_ctx* c, const char* n) : ctx(c), name(n) {
    if (!ctx->timing) return;
    e0 = take(ctx);
    e1 = take(ctx);
    if (e0) hipEventRecord(e0, ctx->stream);
  }
  ~KTimer() {
    if (!ctx->timing || !e0 || !e1) return;
    hipEventRecord(e1, ctx->stream);
    ctx->ev_pending.push_back({name, {e0, e1}});
  }
};

// Words of radix-sort scratch for up to cap keys: alternate buffers A and B
// (keys, values: 4 cap words) and the digit histograms of 4096-key tiles
// (256 per tile, plus the digit totals).
constexpr size_t sort_scratch_words(size_t cap) {
  return 4 * cap + 256 * ((cap + 4095) / 4096 + 1) + 512;
}
// Ensures ctx scratch can hold a sort of n keys with key_limit.
int ensure_sort_scratch(dsph_ctx* ctx, uint32_t n);
// Stable LSD radix sort of (keys, values) in place with scratch buf laid out
// for cap >= n keys (celldiv.hip).
int radix_sort(dsph_ctx* ctx, uint32_t* d_keys, uint32_t* d_values, uint32_t n,
               uint32_t key_limit, uint32_t* buf, size_t cap);
int ensure_redu_scratch(dsph_ctx* ctx, uint32_t n);
int ensure_prep_scratch(dsph_ctx* ctx, uint32_t n);
// Looks up the device address of the no-alloc scratch on the context's
// device (current at the call) and keeps it in the context (forces.hip).
int noalloc_scratch_init(dsph_ctx* ctx);

}  // namespace dsph
