// Level-stack handle shared by the translation units of libmplicits_hip.so.
#pragma once

#include <vector>

#include "common.h"
#include "mlp_core.h"

struct mp_stack {
  struct Level {
    int n_in = 3, n_out = 1, n_hidden = 0;
    std::vector<int> hidden;      // true widths
    int width = 0;                // padded width
    const float* weights = nullptr;  // borrowed, renderer .bin order
    const float* biases = nullptr;   // borrowed
    void* buf = nullptr;          // packed arrays (one allocation)
    size_t buf_bytes = 0;
    int shape = mp::SH_16x16x4_F32;  // fragment shape of this level's hidden GEMMs
    mp::LevelDev dev{};
  };
  int n_levels = 0;
  int precision = MP_FP32;
  int shape = mp::SH_16x16x4_F32;  // of every level, except MP_BF16X3 stacks (per level)
  int nc = 1;  // point column blocks per wave
  // tile shapes of the other kernel families (mp_stack_set_kernel_tiles):
  // batched extraction (1: each width's default column blocks, 2: half of
  // them) and tracing (0: follow nc, 1 or 2: column blocks per wave)
  int batch_tiles = 1;
  int trace_tiles = 0;
  Level lv[MP_MAX_LEVELS];
  void* scratch = nullptr;  // extraction / tracing work lists, grown on demand
  size_t scratch_bytes = 0;
  uint64_t version = 0;     // mp_stack_update calls so far
  // training workspace (mp_train_fwd / mp_train_bwd), grown on demand, and the
  // activations the last mp_train_fwd kept in it
  void* train_ws = nullptr;
  size_t train_ws_bytes = 0;
  int kept_level = -1;
  int64_t kept_n = 0;
  const float* kept_x = nullptr;
  uint64_t kept_version = 0;
};

namespace mp {

// Padded width used by the kernels for a level whose widest layer is `w`.
inline int padded_width(int w) {
  if (w <= 64) return 64;
  if (w <= 128) return 128;
  if (w <= 256) return 256;
  if (w <= 512) return 512;
  return -1;
}

// Bytes of readable padding after the packed hidden-layer A fragments: the
// level kernels prefetch a fixed number of fragments ahead of their use.
constexpr size_t APK_PAD = 32 * 64 * 16;

// Launch-shape selection shared by every kernel family (a tile-shape override
// from the autotuner is applied at stack creation).
int default_nc(int shape, int width);

// Ensure st->scratch holds at least `bytes`; returns MP_OK or MP_ENOMEM.
int ensure_scratch(mp_stack* st, size_t bytes);

}  // namespace mp
