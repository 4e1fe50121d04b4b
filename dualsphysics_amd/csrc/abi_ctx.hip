// Context, memory and constants entry points of the C ABI.
#include <cstring>
#include "dsph_ctx.h"

namespace dsph {

int set_error(dsph_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  return code;
}

int ensure_sort_scratch(dsph_ctx* ctx, uint32_t n) {
  if (size_t(n) > ctx->sort_cap)
    return set_error(ctx, DSPH_E_NOMEM,
                     "sort scratch too small: call dsph_ctx_reserve with np_max >= " +
                         std::to_string(n));
  return DSPH_OK;
}

int ensure_prep_scratch(dsph_ctx* ctx, uint32_t n) {
  if (size_t(n) > ctx->prep_cap)
    return set_error(ctx, DSPH_E_NOMEM,
                     "interaction scratch too small: call dsph_ctx_reserve with np_max >= " +
                         std::to_string(n));
  return DSPH_OK;
}

int ensure_redu_scratch(dsph_ctx* ctx, uint32_t n) {
  const size_t need = (size_t(n) + 255) / 256 + 256;
  if (need > ctx->redu_cap)
    return set_error(ctx, DSPH_E_NOMEM,
                     "reduction scratch too small: call dsph_ctx_reserve with np_max >= " +
                         std::to_string(n));
  return DSPH_OK;
}

}  // namespace dsph

using namespace dsph;

extern "C" {

int dsph_ctx_create(int device, dsph_ctx** out) {
  if (!out) return DSPH_E_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return DSPH_E_STATE;
  if (device < 0 || device >= ndev) return DSPH_E_ARG;
  if (hipSetDevice(device) != hipSuccess) return DSPH_E_LAUNCH;
  dsph_ctx* c = new dsph_ctx();
  c->device = device;
  if (hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return DSPH_E_LAUNCH;
  }
  c->stream = c->own_stream;
  if (hipMalloc(&c->small_dev, 64 * sizeof(uint32_t)) != hipSuccess ||
      hipHostMalloc(&c->pinned, 4096, hipHostMallocDefault) != hipSuccess) {
    dsph_ctx_destroy(c);
    return DSPH_E_NOMEM;
  }
  if (noalloc_scratch_init(c) != DSPH_OK) {
    dsph_ctx_destroy(c);
    return DSPH_E_LAUNCH;
  }
  *out = c;
  return DSPH_OK;
}

int dsph_ctx_destroy(dsph_ctx* ctx) {
  if (!ctx) return DSPH_OK;
  hipSetDevice(ctx->device);
  if (ctx->stream) hipStreamSynchronize(ctx->stream);
  for (auto& it : ctx->ev_pending) {
    hipEventDestroy(it.second.first);
    hipEventDestroy(it.second.second);
  }
  for (hipEvent_t e : ctx->ev_pool) hipEventDestroy(e);
  for (hipEvent_t e : ctx->bound_ev)
    if (e) hipEventDestroy(e);
  hipFree(ctx->sort_buf);
  hipFree(ctx->redu);
  hipFree(ctx->prep_press);
  hipFree(ctx->prep_posf);
  hipFree(ctx->small_dev);
  if (ctx->pinned) hipHostFree(ctx->pinned);
  if (ctx->own_stream) hipStreamDestroy(ctx->own_stream);
  delete ctx;
  return DSPH_OK;
}

int dsph_ctx_reserve(dsph_ctx* ctx, uint32_t np_max, uint32_t nctt_max) {
  if (!ctx) return DSPH_E_ARG;
  (void)nctt_max;
  DSPH_HIP(ctx, hipSetDevice(ctx->device));
  DSPH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (np_max > ctx->sort_cap) {
    hipFree(ctx->sort_buf); ctx->sort_buf = nullptr;
    ctx->sort_cap = 0;
    DSPH_HIP(ctx, hipMalloc(&ctx->sort_buf, sizeof(uint32_t) * sort_scratch_words(np_max)));
    ctx->sort_cap = np_max;
  }
  if (np_max > ctx->prep_cap) {
    hipFree(ctx->prep_press); ctx->prep_press = nullptr;
    hipFree(ctx->prep_posf); ctx->prep_posf = nullptr;
    ctx->prep_cap = 0;
    DSPH_HIP(ctx, hipMalloc(&ctx->prep_press, 2 * sizeof(float) * size_t(np_max)));
    DSPH_HIP(ctx, hipMalloc(&ctx->prep_posf, 4 * sizeof(float) * size_t(np_max)));
    ctx->prep_cap = np_max;
  }
  const size_t rw = (size_t(np_max) + 255) / 256 + 256;
  if (rw > ctx->redu_cap) {
    hipFree(ctx->redu); ctx->redu = nullptr; ctx->redu_cap = 0;
    DSPH_HIP(ctx, hipMalloc(&ctx->redu, sizeof(float) * rw));
    ctx->redu_cap = rw;
  }
  return DSPH_OK;
}

int dsph_ctx_set_stream(dsph_ctx* ctx, void* stream) {
  if (!ctx) return DSPH_E_ARG;
  ctx->stream = stream ? reinterpret_cast<hipStream_t>(stream) : ctx->own_stream;
  return DSPH_OK;
}

void* dsph_ctx_get_stream(dsph_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int dsph_synchronize(dsph_ctx* ctx) {
  if (!ctx) return DSPH_E_ARG;
  DSPH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return DSPH_OK;
}

const char* dsph_last_error(dsph_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int dsph_malloc(dsph_ctx* ctx, size_t bytes, void** d_out) {
  if (!ctx || !d_out) return DSPH_E_ARG;
  *d_out = nullptr;
  if (!bytes) return DSPH_OK;
  DSPH_HIP(ctx, hipSetDevice(ctx->device));
  hipError_t e = hipMalloc(d_out, bytes);
  if (e != hipSuccess)
    return set_error(ctx, DSPH_E_NOMEM, "hipMalloc(" + std::to_string(bytes) + "): " +
                                            hipGetErrorString(e));
  return DSPH_OK;
}

int dsph_free(dsph_ctx* ctx, void* d_ptr) {
  if (!ctx) return DSPH_E_ARG;
  if (d_ptr) DSPH_HIP(ctx, hipFree(d_ptr));
  return DSPH_OK;
}

int dsph_memset_async(dsph_ctx* ctx, void* d_ptr, int value, size_t bytes) {
  if (!ctx) return DSPH_E_ARG;
  if (bytes) DSPH_HIP(ctx, hipMemsetAsync(d_ptr, value, bytes, ctx->stream));
  return DSPH_OK;
}

int dsph_memcpy_h2d(dsph_ctx* ctx, void* d_dst, const void* h_src, size_t bytes) {
  if (!ctx) return DSPH_E_ARG;
  if (bytes) {
    DSPH_HIP(ctx, hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, ctx->stream));
    DSPH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  return DSPH_OK;
}

int dsph_memcpy_d2h(dsph_ctx* ctx, void* h_dst, const void* d_src, size_t bytes) {
  if (!ctx) return DSPH_E_ARG;
  if (bytes) {
    DSPH_HIP(ctx, hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    DSPH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  return DSPH_OK;
}

int dsph_memcpy_d2d_async(dsph_ctx* ctx, void* d_dst, const void* d_src, size_t bytes) {
  if (!ctx) return DSPH_E_ARG;
  if (bytes)
    DSPH_HIP(ctx, hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  return DSPH_OK;
}

int dsph_ctx_set_timing(dsph_ctx* ctx, int enabled) {
  if (!ctx) return DSPH_E_ARG;
  ctx->timing = enabled != 0;
  return DSPH_OK;
}

int dsph_ctx_timing_collect(dsph_ctx* ctx) {
  if (!ctx) return DSPH_E_ARG;
  DSPH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (auto& it : ctx->ev_pending) {
    float ms = 0;
    DSPH_HIP(ctx, hipEventElapsedTime(&ms, it.second.first, it.second.second));
    dsph_ktime& k = ctx->ktimes[it.first];
    k.count++;
    k.ms += ms;
    ctx->ev_pool.push_back(it.second.first);
    ctx->ev_pool.push_back(it.second.second);
  }
  ctx->ev_pending.clear();
  return DSPH_OK;
}

int dsph_ctx_timing_get(dsph_ctx* ctx, uint32_t idx, char* name, size_t name_len,
                        uint64_t* count, double* total_ms) {
  if (!ctx) return DSPH_E_ARG;
  if (idx >= ctx->ktimes.size()) return DSPH_E_ARG;
  auto it = ctx->ktimes.begin();
  std::advance(it, idx);
  if (name && name_len) {
    std::strncpy(name, it->first.c_str(), name_len - 1);
    name[name_len - 1] = 0;
  }
  if (count) *count = it->second.count;
  if (total_ms) *total_ms = it->second.ms;
  return DSPH_OK;
}

int dsph_ctx_timing_reset(dsph_ctx* ctx) {
  if (!ctx) return DSPH_E_ARG;
  int rc = dsph_ctx_timing_collect(ctx);
  ctx->ktimes.clear();
  return rc;
}

int dsph_set_constants(dsph_ctx* ctx, const dsph_cte* cte) {
  if (!ctx || !cte) return DSPH_E_ARG;
  ctx->cte = *cte;
  ctx->cte_set = true;
  return DSPH_OK;
}

}  // extern "C"
