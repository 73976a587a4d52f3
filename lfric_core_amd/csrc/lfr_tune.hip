// Autotuner: launch-shape choices measured on the device and remembered.
//
// Several kernels of the path come in variants that give identical results
// and differ only in how the work is mapped (columns per wave and stage
// height of the column solves, output staging and cache policy of the packed
// apply, workgroup size of the multigrid smoother).  The fastest variant
// depends on the problem size and the part, so the first call of each shape
// times every variant on the context's stream (a warm-up call, then three
// timed calls, HIP events) and keeps the fastest under a key naming the
// kernel family and the shape.  Later calls, and every call once a choice is
// stored, launch the kept variant directly.  No tuning happens during stream
// capture (the default is used and not stored) or with LFR_TUNE=0.
//
// The choices can be listed, set, saved to a text file and loaded back
// (lfr_tune_*; LFR_TUNE_FILE names a file loaded by lfr_init and written by
// lfr_finalise), so a production run reuses the choices of a tuning run.
// File format: one "choice key" pair per line.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "lfr_internal.h"

bool lfr_tune_enabled() {
  const char* e = getenv("LFR_TUNE");
  return !(e && e[0] == '0');
}

int lfr_tune_pick(lfr_ctx* ctx, const std::string& key, int ncand, int dflt, const std::function<int(int)>& run) {
  auto it = ctx->tune.find(key);
  if (it != ctx->tune.end() && it->second >= 0 && it->second < ncand) return it->second;
  if (!lfr_tune_enabled()) return dflt;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(ctx->stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) {
    (void)hipGetLastError();
    return dflt;
  }
  for (auto& e : ctx->tev)
    if (!e && hipEventCreate(&e) != hipSuccess) {
      e = nullptr;
      return dflt;
    }
  constexpr int kReps = 3;
  int best = -1;
  float best_ms = INFINITY;
  const std::string saved_err = ctx->err;
  for (int c = 0; c < ncand; ++c) {
    if (run(c) != LFR_OK) {   // a variant that cannot run this shape is skipped
      (void)hipGetLastError();
      continue;
    }
    if (hipEventRecord(ctx->tev[0], ctx->stream) != hipSuccess) break;
    int rc = LFR_OK;
    for (int r = 0; r < kReps && rc == LFR_OK; ++r) rc = run(c);
    if (rc != LFR_OK || hipEventRecord(ctx->tev[1], ctx->stream) != hipSuccess ||
        hipEventSynchronize(ctx->tev[1]) != hipSuccess) {
      (void)hipGetLastError();
      continue;
    }
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ctx->tev[0], ctx->tev[1]) != hipSuccess) continue;
    if (ms < best_ms) {
      best_ms = ms;
      best = c;
    }
  }
  ctx->err = saved_err;
  if (best < 0) return dflt;
  ctx->tune[key] = best;
  return best;
}

int lfr_tune_load_file(lfr_ctx* ctx, const char* path) {
  FILE* f = fopen(path, "r");
  if (!f) return lfr_set_error(ctx, LFR_ERR_ARG, "lfr_tune_load: cannot open %s", path);
  char line[1024];
  while (fgets(line, sizeof line, f)) {
    char* end = nullptr;
    const long choice = strtol(line, &end, 10);
    if (end == line || choice < 0 || choice > 1000) continue;
    while (*end == ' ' || *end == '\t') ++end;
    std::string key(end);
    while (!key.empty() && (key.back() == '\n' || key.back() == '\r' || key.back() == ' ')) key.pop_back();
    if (key.empty()) continue;
    ctx->tune[key] = (int)choice;
  }
  fclose(f);
  return LFR_OK;
}

int lfr_tune_save_file(const lfr_ctx* ctx, const char* path) {
  FILE* f = fopen(path, "w");
  if (!f) return LFR_ERR_ARG;
  for (const auto& kv : ctx->tune) fprintf(f, "%d %s\n", kv.second, kv.first.c_str());
  return fclose(f) == 0 ? LFR_OK : LFR_ERR_ARG;
}

extern "C" {

int lfr_tune_set(lfr_ctx* ctx, const char* key, int choice) {
  if (!ctx) return LFR_ERR_ARG;
  LFR_REQUIRE(ctx, key && key[0] && choice >= 0, "lfr_tune_set: bad key or choice");
  ctx->tune[key] = choice;
  return LFR_OK;
}

int lfr_tune_get(lfr_ctx* ctx, const char* key, int* choice) {
  if (!ctx) return LFR_ERR_ARG;
  LFR_REQUIRE(ctx, key && choice, "lfr_tune_get: null argument");
  auto it = ctx->tune.find(key);
  LFR_REQUIRE(ctx, it != ctx->tune.end(), "lfr_tune_get: no choice stored for %s", key);
  *choice = it->second;
  return LFR_OK;
}

int lfr_tune_count(const lfr_ctx* ctx, int* n) {
  if (!ctx || !n) return LFR_ERR_ARG;
  *n = (int)ctx->tune.size();
  return LFR_OK;
}

int lfr_tune_entry(lfr_ctx* ctx, int i, char* key, int keylen, int* choice) {
  if (!ctx) return LFR_ERR_ARG;
  LFR_REQUIRE(ctx, key && keylen > 0 && choice && i >= 0 && i < (int)ctx->tune.size(),
              "lfr_tune_entry: bad argument");
  auto it = ctx->tune.begin();
  std::advance(it, i);
  snprintf(key, (size_t)keylen, "%s", it->first.c_str());
  *choice = it->second;
  return LFR_OK;
}

int lfr_tune_clear(lfr_ctx* ctx) {
  if (!ctx) return LFR_ERR_ARG;
  ctx->tune.clear();
  return LFR_OK;
}

int lfr_tune_save(lfr_ctx* ctx, const char* path) {
  if (!ctx) return LFR_ERR_ARG;
  LFR_REQUIRE(ctx, path, "lfr_tune_save: null path");
  const int rc = lfr_tune_save_file(ctx, path);
  if (rc) return lfr_set_error(ctx, rc, "lfr_tune_save: cannot write %s", path);
  return LFR_OK;
}

int lfr_tune_load(lfr_ctx* ctx, const char* path) {
  if (!ctx) return LFR_ERR_ARG;
  LFR_REQUIRE(ctx, path, "lfr_tune_load: null path");
  return lfr_tune_load_file(ctx, path);
}

}  // extern "C"
