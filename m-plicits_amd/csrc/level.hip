// One-level evaluation kernel over explicit points, grid index ranges and
// compacted grid-index lists (every MFMA shape), with the in-band compaction of
// the extraction pipeline fused into the store.  The level itself is
// LevelCore (level_core.h).
#include "eval_dispatch.h"
#include "level_core.h"

namespace mp {

template <class S, int W, int NC>
__global__ __launch_bounds__(256) void k_level(LevelDev L, EvalTarget t) {
  typedef LevelCore<S, W, NC, true> M;
  typedef typename S::Frag Frag;
  extern __shared__ uint4 lds_raw[];
  const int lane = threadIdx.x & 63;
  const int64_t n = (t.src == SRC_LIST) ? (int64_t)*t.list_count : t.n;
  // Workgroup-uniform exit: the waves of a workgroup may share barriers, so a
  // wave whose tile lies past n still runs (on clamped points) and stores nothing.
  if ((int64_t)blockIdx.x * 4 * M::TILE >= n) return;
  const int64_t base = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * M::TILE;

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
  Frag* lds = reinterpret_cast<Frag*>(lds_raw);
  M::eval(L, x, f, lds, lds + M::RING_FRAGS + (threadIdx.x >> 6) * M::HANDOFF_FRAGS);

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
  typedef LevelCore<S, W, NC, true> M;
  const int64_t per_block = (int64_t)M::TILE * 4;
  const int64_t blocks = (t.n + per_block - 1) / per_block;
  if (blocks <= 0) return MP_OK;
  MP_REQUIRE(blocks < (1ll << 31), MP_EINVAL, "level launch: %lld points is too many", (long long)t.n);
  static_assert(M::LDS_BYTES <= 160 * 1024, "LDS budget of one workgroup");
  static bool attr_set = false;
  if (!attr_set) {
    MP_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_level<S, W, NC>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)M::LDS_BYTES));
    attr_set = true;
  }
  hipLaunchKernelGGL((k_level<S, W, NC>), dim3((unsigned)blocks), dim3(256), M::LDS_BYTES, stream, L, t);
  MP_LAUNCHED("k_level");
  return MP_OK;
}

// Column blocks per wave: the B operand set takes NC * W / 4 VGPRs (16-bit) or
// NC * W / 4 VGPRs of FP32 at 16 rows per block; the defaults fill 128 VGPRs.
// nc == 1 asks for single-block tiles.
template <class S>
static int launch_width(const LevelDev& L, int nc, const EvalTarget& t, hipStream_t stream) {
  constexpr bool wide = S::ROWS == 32;
  if (nc == 1) {
    switch (L.width) {
      case 64: return launch_typed<S, 64, 1>(L, t, stream);
      case 128: return launch_typed<S, 128, 1>(L, t, stream);
      case 256: return launch_typed<S, 256, 1>(L, t, stream);
      case 512: return launch_typed<S, 512, 1>(L, t, stream);
    }
  } else {
    switch (L.width) {
      case 64: return launch_typed<S, 64, wide ? 4 : 8>(L, t, stream);
      case 128: return launch_typed<S, 128, wide ? 2 : 4>(L, t, stream);
      case 256: return launch_typed<S, 256, 2>(L, t, stream);
      case 512: return launch_typed<S, 512, 1>(L, t, stream);
    }
  }
  set_error("level launch: unsupported padded width %d", L.width);
  return MP_EUNSUPPORTED;
}

int launch_level(const mp_stack* st, int level, const EvalTarget& t, hipStream_t stream) {
  const LevelDev& L = st->lv[level].dev;
  switch (st->shape) {
    case SH_32x32x16_BF16: return launch_width<S32x32x16_bf16>(L, st->nc, t, stream);
    case SH_32x32x16_F16: return launch_width<S32x32x16_f16>(L, st->nc, t, stream);
    case SH_16x16x4_F32: return launch_width<S16x16x4_f32>(L, st->nc, t, stream);
  }
  set_error("level launch: unsupported fragment shape %d", st->shape);
  return MP_EUNSUPPORTED;
}

}  // namespace mp
