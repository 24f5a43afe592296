// One-level evaluation kernels over explicit points, grid index ranges and
// compacted grid-index lists, with the in-band compaction of the extraction
// pipeline fused into the store epilogue.
#include "eval_dispatch.h"

namespace mp {

template <class S, int W, int NC>
__global__ __launch_bounds__(256) void k_level(LevelDev L, EvalTarget t) {
  typedef Mlp<S, W, NC> M;
  const int lane = threadIdx.x & 63;
  const int64_t wave = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int64_t base = wave * M::TILE;
  const int64_t n = (t.src == SRC_LIST) ? (int64_t)*t.list_count : t.n;
  if (base >= n) return;  // wave-uniform

  float x[NC][3];
  int64_t key[NC];
  bool valid[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const int64_t i = base + c * S::ROWS + (lane % S::ROWS);
    valid[c] = i < n;
    const int64_t ii = valid[c] ? i : n - 1;
    if (t.src == SRC_POINTS) {
      key[c] = ii;
      x[c][0] = t.xyz[3 * ii + 0];
      x[c][1] = t.xyz[3 * ii + 1];
      x[c][2] = t.xyz[3 * ii + 2];
    } else {
      key[c] = (t.src == SRC_GRID) ? t.first + ii : (int64_t)t.list[ii];
      grid_coord(key[c], t.N, t.grid_mode, t.h, x[c]);
    }
  }
  float f[NC];
  M::eval(L, x, f);

  // Lanes of group 0 own the store (every group holds the same value).
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    bool flag = false;
    if (lane < S::ROWS && valid[c]) {
      float v = f[c];
      if (t.accumulate) v = t.out[key[c]] + v;
      t.out[key[c]] = v;
      flag = t.band > 0.f && fabsf(v) < t.band;
    }
    if (t.band > 0.f) {
      const uint64_t m = __ballot(flag);
      if (m) {
        const int cnt = __popcll(m);
        uint32_t pos = 0;
        if (lane == 0) pos = atomicAdd(t.cand_count, (uint32_t)cnt);
        pos = __shfl(pos, 0);
        if (flag) {
          const uint64_t below = (lane == 0) ? 0ull : (m & ((~0ull) >> (64 - lane)));
          t.cand[pos + __popcll(below)] = (uint32_t)key[c];
        }
      }
    }
  }
}

template <class S, int W, int NC>
static int launch_typed(const LevelDev& L, const EvalTarget& t, hipStream_t stream) {
  const int64_t tile = (int64_t)NC * S::ROWS * 4;
  const int64_t blocks = (t.n + tile - 1) / tile;
  if (blocks <= 0) return MP_OK;
  MP_REQUIRE(blocks < (1ll << 31), MP_EINVAL, "level launch: %lld points is too many", (long long)t.n);
  hipLaunchKernelGGL((k_level<S, W, NC>), dim3((unsigned)blocks), dim3(256), 0, stream, L, t);
  MP_LAUNCHED("k_level");
  return MP_OK;
}

template <class S, int NC>
static int launch_width(const LevelDev& L, const EvalTarget& t, hipStream_t stream) {
  switch (L.width) {
    case 64: return launch_typed<S, 64, NC>(L, t, stream);
    case 128: return launch_typed<S, 128, NC>(L, t, stream);
    case 256: return launch_typed<S, 256, NC>(L, t, stream);
    case 512: return launch_typed<S, 512, 1>(L, t, stream);  // two 512-wide tiles do not fit the VGPRs
  }
  set_error("level launch: unsupported padded width %d", L.width);
  return MP_EUNSUPPORTED;
}

static int level_nc(const mp_stack* st, int level) {
  return (st->nc == 2 && st->lv[level].width <= 256) ? 2 : 1;
}

int level_tile(const mp_stack* st, int level) {
  const ShapeInfo s = shape_info(st->shape);
  return s.rows * level_nc(st, level);
}

int launch_level(const mp_stack* st, int level, const EvalTarget& t, hipStream_t stream) {
  const LevelDev& L = st->lv[level].dev;
  const int nc = level_nc(st, level);
  switch (st->shape) {
    case SH_32x32x16_BF16:
    case SH_32x32x16_F16:
      return launch_level16(st, level, t, stream);
    case SH_16x16x4_F32:
      return launch_width<S16x16x4_f32, 1>(L, t, stream);
  }
  set_error("level launch: unsupported fragment shape %d", st->shape);
  return MP_EUNSUPPORTED;
}

}  // namespace mp

