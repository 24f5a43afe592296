// Tile-shape autotuner of the level kernels.
//
// The reference fixes its CUTLASS tile shapes at compile time
// (renderer/cuda/layer_gemms.cu:1612-1856).  Here the point-column blocks per
// wave (the knob of mp_stack_set_tiles) are measured on the device for one
// (width, depth, precision): a stack with seeded weights evaluates a grid of
// 2^22 samples once per candidate, timed with HIP events after a warm-up, and
// the fastest candidate is returned.  Results are cached in a text file (one
// "width n_hidden precision column_blocks" line per entry, guarded by a
// process-wide mutex); a cached entry is returned without measuring.
#include <cmath>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

#include "eval_dispatch.h"

namespace mp {

static std::mutex g_tune_mutex;

static bool cache_lookup(const char* path, int width, int n_hidden, int precision, int* out) {
  FILE* f = path ? std::fopen(path, "r") : nullptr;
  if (!f) return false;
  int w, h, p, c;
  bool found = false;
  while (std::fscanf(f, "%d %d %d %d", &w, &h, &p, &c) == 4) {
    if (w == width && h == n_hidden && p == precision) {
      *out = c;
      found = true;  // the last entry wins
    }
  }
  std::fclose(f);
  return found;
}

static void cache_store(const char* path, int width, int n_hidden, int precision, int cols) {
  if (!path) return;
  FILE* f = std::fopen(path, "a");
  if (!f) return;
  std::fprintf(f, "%d %d %d %d\n", width, n_hidden, precision, cols);
  std::fclose(f);
}

// Deterministic weights in the SIREN init range (layer 0: U(-1/3, 1/3) * 30,
// others U(-sqrt(6/W), sqrt(6/W))), generated on the host.
static std::vector<float> seeded(size_t n, float scale, uint32_t seed) {
  std::vector<float> v(n);
  uint32_t s = seed * 2654435761u + 12345u;
  for (size_t i = 0; i < n; ++i) {
    s ^= s << 13; s ^= s >> 17; s ^= s << 5;
    v[i] = scale * (2.0f * (float)(s & 0xffffff) / 16777216.0f - 1.0f);
  }
  return v;
}

}  // namespace mp

using namespace mp;

extern "C" {

int mp_autotune(const char* cache_path, int width, int n_hidden, int precision, int* column_blocks) {
  MP_REQUIRE(column_blocks != nullptr, MP_EINVAL, "mp_autotune: column_blocks is NULL");
  MP_REQUIRE(width >= 1 && width <= 512 && n_hidden >= 1 && n_hidden <= 16, MP_EINVAL,
             "mp_autotune: width must be in [1, 512] and n_hidden in [1, 16]");
  MP_REQUIRE(precision == MP_FP32 || precision == MP_BF16 || precision == MP_FP16, MP_EINVAL,
             "mp_autotune: unknown precision %d", precision);
  std::lock_guard<std::mutex> lock(g_tune_mutex);
  if (cache_lookup(cache_path, width, n_hidden, precision, column_blocks)) return MP_OK;

  // seeded level 3 -> width x n_hidden -> 1
  std::vector<int> dims(n_hidden + 2, width);
  dims[0] = 3;
  dims[n_hidden + 1] = 1;
  size_t nw = 0, nb = 0;
  for (int k = 0; k <= n_hidden; ++k) {
    nw += (size_t)dims[k] * dims[k + 1];
    nb += (size_t)dims[k + 1];
  }
  std::vector<float> hw, hb;
  for (int k = 0; k <= n_hidden; ++k) {
    const float sc = k == 0 ? 10.0f : std::sqrt(6.0f / (float)dims[k]);
    std::vector<float> w = seeded((size_t)dims[k] * dims[k + 1], sc, 7u + k);
    std::vector<float> b = seeded((size_t)dims[k + 1], 0.1f, 101u + k);
    hw.insert(hw.end(), w.begin(), w.end());
    hb.insert(hb.end(), b.begin(), b.end());
  }
  const int64_t N = 160;  // 4,096,000 samples
  float *dw = nullptr, *db = nullptr, *vol = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  mp_stack* st = nullptr;
  int rc = MP_OK, best = 2;
  float best_ms = 1e30f;
  auto fail = [&](hipError_t e, const char* what) {
    set_error("mp_autotune: %s: %s", what, hipGetErrorString(e));
    rc = MP_EHIP;
  };
  hipError_t e = hipMalloc(&dw, nw * 4);
  if (e == hipSuccess) e = hipMalloc(&db, nb * 4);
  if (e == hipSuccess) e = hipMalloc(&vol, (size_t)N * N * N * 4);
  if (e == hipSuccess) e = hipMemcpy(dw, hw.data(), nw * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(db, hb.data(), nb * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e != hipSuccess) fail(e, "setup");
  if (rc == MP_OK) {
    std::vector<int> hidden(n_hidden, width);
    mp_level_desc d{3, 1, n_hidden, hidden.data(), dw, db};
    rc = mp_stack_create(&d, 1, precision, nullptr, &st);
  }
  for (int cols = 1; rc == MP_OK && cols <= 2; ++cols) {
    rc = mp_stack_set_tiles(st, cols);
    for (int rep = 0; rc == MP_OK && rep < 3; ++rep) {
      if (rep == 1) e = hipEventRecord(e0, nullptr);
      if (rc == MP_OK) rc = mp_extract_volume(st, (int)N, MP_GRID_LATTICE, 32768, nullptr, 0, vol, nullptr, nullptr);
    }
    if (rc != MP_OK) break;
    e = hipEventRecord(e1, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e != hipSuccess) {
      fail(e, "timing");
      break;
    }
    if (ms < best_ms) {
      best_ms = ms;
      best = cols;
    }
  }
  if (st) mp_stack_destroy(st);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (vol) (void)hipFree(vol);
  if (db) (void)hipFree(db);
  if (dw) (void)hipFree(dw);
  if (rc != MP_OK) return rc;
  cache_store(cache_path, width, n_hidden, precision, best);
  *column_blocks = best;
  return MP_OK;
}

}  // extern "C"
