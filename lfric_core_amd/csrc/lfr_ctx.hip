// Context, memory and function-space handles of liblfr.so.
#include <algorithm>
#include <atomic>
#include <cstdarg>
#include <cstring>
#include <mutex>
#include <unordered_set>

#include <cstdlib>

#include "lfr_internal.h"

int lfr_set_error(lfr_ctx* ctx, int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (ctx) ctx->err = buf;
  else fprintf(stderr, "lfr: %s\n", buf);
  return code;
}

int lfr_work_reserve(lfr_ctx* ctx, size_t bytes) {
  if (bytes <= ctx->work_bytes) return LFR_OK;
  if (ctx->work) {
    LFR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    LFR_HIP(ctx, hipFree(ctx->work));
    ctx->work = nullptr;
    ctx->work_bytes = 0;
  }
  size_t want = bytes < (1u << 20) ? (1u << 20) : bytes;
  if (hipMalloc(&ctx->work, want) != hipSuccess)
    return lfr_set_error(ctx, LFR_ERR_NOMEM, "workspace allocation of %zu bytes failed", want);
  ctx->work_bytes = want;
  return LFR_OK;
}

uint64_t lfr_next_uid() {
  static std::atomic<uint64_t> next{0};
  return ++next;
}

// Contexts not yet finalised: a handle may outlive its context (a caller that
// finalises first), and then has no graph cache to evict from.
static std::mutex g_live_mu;
static std::unordered_set<const lfr_ctx*> g_live;

void lfr_graphs_evict(lfr_ctx* ctx, uint64_t uid) {
  if (!ctx || !uid) return;
  {
    std::lock_guard<std::mutex> lk(g_live_mu);
    if (!g_live.count(ctx)) return;
  }
  for (size_t i = 0; i < ctx->graphs.size();) {
    const auto& d = ctx->graphs[i].deps;
    bool hit = false;
    for (uint64_t u : d) hit = hit || u == uid;
    if (hit) {
      (void)hipGraphExecDestroy(ctx->graphs[i].exec);
      ctx->graphs.erase(ctx->graphs.begin() + i);
    } else {
      ++i;
    }
  }
}

extern "C" {

int lfr_init(int device, lfr_ctx** out) {
  if (!out) return LFR_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return lfr_set_error(nullptr, LFR_ERR_HIP, "no HIP device available");
  if (device < 0 || device >= ndev)
    return lfr_set_error(nullptr, LFR_ERR_ARG, "device %d out of range (%d devices)", device, ndev);
  lfr_ctx* ctx = new lfr_ctx();
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess) {
    delete ctx;
    return lfr_set_error(nullptr, LFR_ERR_HIP, "hipSetDevice(%d) failed", device);
  }
  // Default priority stream (no priority changes anywhere in this library).
  if (hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return lfr_set_error(nullptr, LFR_ERR_HIP, "stream creation failed");
  }
  ctx->stream = ctx->own_stream;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) ctx->cu_count = prop.multiProcessorCount;
  if (hipHostMalloc((void**)&ctx->host_scalars, 64 * sizeof(double), hipHostMallocDefault) !=
      hipSuccess) {
    (void)hipStreamDestroy(ctx->own_stream);
    delete ctx;
    return lfr_set_error(nullptr, LFR_ERR_NOMEM, "pinned scalar buffer allocation failed");
  }
  auto drop_events = [ctx] {
    for (auto& e : ctx->ev)
      if (e) (void)hipEventDestroy(e);
  };
  for (auto& e : ctx->ev)
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
      e = nullptr;
      drop_events();
      (void)hipHostFree(ctx->host_scalars);
      (void)hipStreamDestroy(ctx->own_stream);
      delete ctx;
      return lfr_set_error(nullptr, LFR_ERR_HIP, "event creation failed");
    }
  int rc = lfr_work_reserve(ctx, 1u << 20);
  if (rc == LFR_OK) rc = lfr_noalloc_setup(ctx);
  if (rc != LFR_OK) {
    if (ctx->work) (void)hipFree(ctx->work);
    fprintf(stderr, "lfr_init: %s\n", ctx->err.c_str());
    drop_events();
    (void)hipHostFree(ctx->host_scalars);
    (void)hipStreamDestroy(ctx->own_stream);
    delete ctx;
    return rc;
  }
  {
    std::lock_guard<std::mutex> lk(g_live_mu);
    g_live.insert(ctx);
  }
  // launch-shape choices of an earlier run (a missing file is not an error)
  if (const char* tf = getenv("LFR_TUNE_FILE")) {
    if (FILE* f = fopen(tf, "r")) {
      fclose(f);
      (void)lfr_tune_load_file(ctx, tf);
    }
  }
  *out = ctx;
  return LFR_OK;
}

int lfr_finalise(lfr_ctx* ctx) {
  if (!ctx) return LFR_OK;
  {
    std::lock_guard<std::mutex> lk(g_live_mu);
    g_live.erase(ctx);
  }
  (void)hipStreamSynchronize(ctx->stream);
  if (const char* tf = getenv("LFR_TUNE_FILE"))
    if (!ctx->tune.empty()) (void)lfr_tune_save_file(ctx, tf);
  for (auto& g : ctx->graphs) (void)hipGraphExecDestroy(g.exec);
  for (auto& e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : ctx->tev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->work) (void)hipFree(ctx->work);
  if (ctx->kwork) (void)hipFree(ctx->kwork);
  if (ctx->host_scalars) (void)hipHostFree(ctx->host_scalars);
  if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
  delete ctx;
  return LFR_OK;
}

const char* lfr_last_error(const lfr_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int lfr_set_stream(lfr_ctx* ctx, void* s) {
  if (!ctx) return LFR_ERR_ARG;
  ctx->stream = (hipStream_t)s;  // NULL selects HIP's default (null) stream
  return LFR_OK;
}

void* lfr_get_stream(const lfr_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int lfr_synchronize(lfr_ctx* ctx) {
  if (!ctx) return LFR_ERR_ARG;
  LFR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return LFR_OK;
}

int lfr_malloc(lfr_ctx* ctx, size_t bytes, void** p) {
  if (!ctx || !p) return LFR_ERR_ARG;
  if (hipMalloc(p, bytes ? bytes : 1) != hipSuccess)
    return lfr_set_error(ctx, LFR_ERR_NOMEM, "hipMalloc(%zu) failed", bytes);
  return LFR_OK;
}

int lfr_free(lfr_ctx* ctx, void* p) {
  if (!ctx) return LFR_ERR_ARG;
  if (p) LFR_HIP(ctx, hipFree(p));
  return LFR_OK;
}

int lfr_memcpy_h2d(lfr_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx) return LFR_ERR_ARG;
  LFR_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  LFR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return LFR_OK;
}

int lfr_memcpy_d2h(lfr_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx) return LFR_ERR_ARG;
  LFR_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  LFR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return LFR_OK;
}

int lfr_space_create(lfr_ctx* ctx, const int32_t* map, int ndf, int ncell, int nlayers, int undf,
                     lfr_space** out) {
  if (!ctx || !out) return LFR_ERR_ARG;
  *out = nullptr;
  LFR_REQUIRE(ctx, map && ndf > 0 && ncell > 0 && nlayers > 0 && undf > 0,
              "lfr_space_create: bad sizes ndf=%d ncell=%d nlayers=%d undf=%d", ndf, ncell,
              nlayers, undf);
  const int64_t n = (int64_t)ndf * ncell;
  std::vector<int32_t> m0(n);
  for (int64_t q = 0; q < n; ++q) {
    int32_t v = map[q];
    // every dof the column touches, v .. v+nlayers-1 (1-based), must exist
    if (v < 1 || (int64_t)v - 1 + nlayers > undf)
      return lfr_set_error(ctx, LFR_ERR_ARG,
                           "lfr_space_create: map(%d,%d)=%d addresses dofs outside 1..%d",
                           (int)(q % ndf) + 1, (int)(q / ndf) + 1, v, undf);
    m0[q] = v - 1;
  }
  lfr_space* s = new lfr_space();
  s->ctx = ctx;
  s->uid = lfr_next_uid();
  s->ndf = ndf;
  s->ncell = ncell;
  s->nlayers = nlayers;
  s->undf = undf;
  s->map_h.assign(map, map + n);
  for (int c = 0; c < ncell; ++c) s->max_first = std::max(s->max_first, (int)map[(int64_t)c * ndf]);
  {
    // columns at a constant stride (W3, Wtheta, ...): kernels then form a
    // column's base instead of reading it first
    const int64_t st = ncell > 1 ? (int64_t)m0[ndf] - m0[0] : 0;
    bool aff = m0[0] == 0 && st >= 0;
    for (int c = 1; c < ncell && aff; ++c) aff = (int64_t)m0[(int64_t)c * ndf] == c * st;
    s->col_stride = aff ? (int)st : -1;
  }
  if (hipMalloc(&s->map0_d, n * sizeof(int32_t)) != hipSuccess) {
    delete s;
    return lfr_set_error(ctx, LFR_ERR_NOMEM, "dofmap allocation failed");
  }
  if (hipMemcpy(s->map0_d, m0.data(), n * sizeof(int32_t), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(s->map0_d);
    delete s;
    return lfr_set_error(ctx, LFR_ERR_HIP, "dofmap upload failed");
  }
  *out = s;
  return LFR_OK;
}

int lfr_space_destroy(lfr_space* s) {
  if (!s) return LFR_OK;
  lfr_graphs_evict(s->ctx, s->uid);
  if (s->map0_d) (void)hipFree(s->map0_d);
  delete s;
  return LFR_OK;
}

}  // extern "C"
