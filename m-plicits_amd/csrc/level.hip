// One-level evaluation kernel over explicit points, grid index ranges and
// compacted grid-index lists (every MFMA shape), with the in-band compaction of
// the extraction pipeline fused into the store.  The level itself is
// LevelCore (level_core.h).
#include "eval_dispatch.h"
#include "level_core.h"

namespace mp {

template <class S, int W, int NC, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_level(LevelDev L, EvalTarget t) {
  extern __shared__ uint4 lds_raw[];
  const int64_t n = (t.src == SRC_LIST) ? (int64_t)*t.list_count : t.n;
  // Workgroup-uniform exit: the waves of a workgroup may share barriers.
  if ((int64_t)blockIdx.x * WAVES * LevelCore<S, W, NC, true, WAVES>::TILE >= n) return;
  eval_block<S, W, NC, WAVES>(L, t, n, blockIdx.x, reinterpret_cast<typename S::Frag*>(lds_raw));
}

template <class S, int W, int NC, int WAVES = 4>
static int launch_typed(const LevelDev& L, const EvalTarget& t, hipStream_t stream) {
  typedef LevelCore<S, W, NC, true, WAVES> M;
  const int64_t per_block = (int64_t)M::TILE * WAVES;
  const int64_t blocks = (t.n + per_block - 1) / per_block;
  if (blocks <= 0) return MP_OK;
  MP_REQUIRE(blocks < (1ll << 31), MP_EINVAL, "level launch: %lld points is too many", (long long)t.n);
  static_assert(M::LDS_BYTES <= 160 * 1024, "LDS budget of one workgroup");
  static std::atomic<uint64_t> attr_devs{0};
  MP_HIP(set_lds_once(attr_devs, reinterpret_cast<const void*>(&k_level<S, W, NC, WAVES>), (int)M::LDS_BYTES));
  hipLaunchKernelGGL((k_level<S, W, NC, WAVES>), dim3((unsigned)blocks), dim3(64 * WAVES), M::LDS_BYTES, stream, L,
                     t);
  MP_LAUNCHED("k_level");
  return MP_OK;
}

// Column blocks per wave: default_cols (the B operand set fills 128 VGPRs), or
// one when the stack asks for single-block tiles.
template <class S>
static int launch_width(const LevelDev& L, int nc, const EvalTarget& t, hipStream_t stream) {
  if (nc == 1) {
    // 16-bit single-column tiles up to width 256: two waves per SIMD (their
    // registers and hand-off regions fit twice)
    constexpr int W8 = S::ROWS == 32 ? 8 : 4;
    switch (L.width) {
      case 64: return launch_typed<S, 64, 1, W8>(L, t, stream);
      case 128: return launch_typed<S, 128, 1, W8>(L, t, stream);
      case 256: return launch_typed<S, 256, 1, W8>(L, t, stream);
      case 512: return launch_typed<S, 512, 1>(L, t, stream);
    }
  } else {
    switch (L.width) {
      case 64: return launch_typed<S, 64, default_cols<S>(64)>(L, t, stream);
      case 128: return launch_typed<S, 128, default_cols<S>(128)>(L, t, stream);
      case 256: return launch_typed<S, 256, default_cols<S>(256)>(L, t, stream);
      case 512: return launch_typed<S, 512, default_cols<S>(512)>(L, t, stream);
    }
  }
  set_error("level launch: unsupported padded width %d", L.width);
  return MP_EUNSUPPORTED;
}

// Three-term BF16: one column block per wave (the hi and lo operand sets of a
// 256-wide layer fill the registers and, with the staging ring, the 160 KB of
// LDS of four waves).
static int launch_split(const LevelDev& L, const EvalTarget& t, hipStream_t stream) {
  switch (L.width) {
    case 64: return launch_typed<S32x32x16_bf16x3, 64, 1>(L, t, stream);
    case 128: return launch_typed<S32x32x16_bf16x3, 128, 1>(L, t, stream);
    case 256: return launch_typed<S32x32x16_bf16x3, 256, 1>(L, t, stream);
  }
  set_error("level launch: three-term BF16 supports padded widths up to 256, got %d", L.width);
  return MP_EUNSUPPORTED;
}

// Three-term BF16 with the last hidden product in FP16 (levels with two
// hidden products, one column block per wave as above).
static int launch_split_f(const LevelDev& L, const EvalTarget& t, hipStream_t stream) {
  MP_REQUIRE(L.n_hidden == 3, MP_EUNSUPPORTED, "level launch: the mixed three-term shape needs 3 hidden layers, got %d",
             L.n_hidden);
  switch (L.width) {
    case 64: return launch_typed<S32x32x16_bf16x3f, 64, 1>(L, t, stream);
    case 128: return launch_typed<S32x32x16_bf16x3f, 128, 1>(L, t, stream);
    case 256: return launch_typed<S32x32x16_bf16x3f, 256, 1>(L, t, stream);
  }
  set_error("level launch: the mixed three-term shape supports padded widths up to 256, got %d", L.width);
  return MP_EUNSUPPORTED;
}

// f and the input gradient at n points, forward mode (LevelCore's tangent
// columns: four columns per point), for the widths where the reverse-mode
// gradient kernel does not fit its registers.  Lane group 0 stores: the value
// column f, tangent column t component t - 1 of the gradient.
template <class S, int W, int NC>
__global__ __launch_bounds__(256) void k_level_tangent(LevelDev L, const float* __restrict__ xyz, int64_t n,
                                                       float* f_out, float* g_out, int accumulate,
                                                       const unsigned long long* n_dev) {
  typedef LevelCore<S, W, NC, true, 4, true> M;
  constexpr int PTS = M::TILE / 4;  // points per wave
  extern __shared__ uint4 lds_raw[];
  if (n_dev) n = (int64_t)*n_dev < n ? (int64_t)*n_dev : n;  // device count, n = upper bound
  // workgroup-uniform exit: the waves of a workgroup share the stream's barriers
  if ((int64_t)blockIdx.x * 4 * PTS >= n) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col = lane % S::ROWS, tl = col & 3;
  const int64_t base = ((int64_t)blockIdx.x * 4 + wave) * PTS;
  float x[NC][3];
  int64_t idx[NC];
  bool valid[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const int64_t i = base + c * (S::ROWS / 4) + col / 4;
    valid[c] = i < n;
    idx[c] = valid[c] ? i : n - 1;
    x[c][0] = xyz[3 * idx[c]];
    x[c][1] = xyz[3 * idx[c] + 1];
    x[c][2] = xyz[3 * idx[c] + 2];
  }
  typename S::Frag* lds = reinterpret_cast<typename S::Frag*>(lds_raw);
  float f[NC];
  M::eval(L, x, f, lds, lds + M::RING_FRAGS + wave * M::HANDOFF_FRAGS);
  if (lane < S::ROWS) {
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      if (!valid[c]) continue;
      float* o = tl == 0 ? (f_out ? f_out + idx[c] : nullptr) : g_out + 3 * idx[c] + (tl - 1);
      if (o) *o = accumulate ? *o + f[c] : f[c];
    }
  }
}

template <class S, int W>
static int launch_tangent_typed(const LevelDev& L, const float* xyz, int64_t n, float* f, float* g, int acc,
                                const unsigned long long* n_dev, hipStream_t stream) {
  // three-term BF16: one column block (the operand pairs fill the registers)
  constexpr int NC = S::SPLIT ? 1 : default_cols<S>(W);
  typedef LevelCore<S, W, NC, true, 4, true> M;
  static_assert(M::LDS_BYTES <= 160 * 1024, "LDS budget of one workgroup");
  const int64_t per_block = (int64_t)M::TILE;  // four waves x TILE / 4 points
  const int64_t blocks = (n + per_block - 1) / per_block;
  if (blocks <= 0) return MP_OK;
  MP_REQUIRE(blocks < (1ll << 31), MP_EINVAL, "gradient launch: %lld points is too many", (long long)n);
  static std::atomic<uint64_t> attr_devs{0};
  MP_HIP(set_lds_once(attr_devs, reinterpret_cast<const void*>(&k_level_tangent<S, W, NC>), (int)M::LDS_BYTES));
  hipLaunchKernelGGL((k_level_tangent<S, W, NC>), dim3((unsigned)blocks), dim3(256), M::LDS_BYTES, stream, L, xyz, n,
                     f, g, acc, n_dev);
  MP_LAUNCHED("k_level_tangent");
  return MP_OK;
}

template <class S>
static int launch_tangent_width(const LevelDev& L, const float* xyz, int64_t n, float* f, float* g, int acc,
                                const unsigned long long* n_dev, hipStream_t stream) {
  switch (L.width) {
    case 64: return launch_tangent_typed<S, 64>(L, xyz, n, f, g, acc, n_dev, stream);
    case 128: return launch_tangent_typed<S, 128>(L, xyz, n, f, g, acc, n_dev, stream);
    case 256: return launch_tangent_typed<S, 256>(L, xyz, n, f, g, acc, n_dev, stream);
    case 512: return launch_tangent_typed<S, 512>(L, xyz, n, f, g, acc, n_dev, stream);
  }
  set_error("gradient launch: unsupported padded width %d", L.width);
  return MP_EUNSUPPORTED;
}

int launch_level_tangent(int shape, const LevelDev& L, const float* xyz, int64_t n, float* f, float* g,
                         int accumulate, hipStream_t stream, const unsigned long long* n_dev) {
  switch (shape) {
    case SH_32x32x16_BF16: return launch_tangent_width<S32x32x16_bf16>(L, xyz, n, f, g, accumulate, n_dev, stream);
    case SH_32x32x16_F16: return launch_tangent_width<S32x32x16_f16>(L, xyz, n, f, g, accumulate, n_dev, stream);
    case SH_16x16x4_F32: return launch_tangent_width<S16x16x4_f32>(L, xyz, n, f, g, accumulate, n_dev, stream);
    case SH_32x32x16_BF16X3:
      switch (L.width) {
        case 64: return launch_tangent_typed<S32x32x16_bf16x3, 64>(L, xyz, n, f, g, accumulate, n_dev, stream);
        case 128: return launch_tangent_typed<S32x32x16_bf16x3, 128>(L, xyz, n, f, g, accumulate, n_dev, stream);
        case 256: return launch_tangent_typed<S32x32x16_bf16x3, 256>(L, xyz, n, f, g, accumulate, n_dev, stream);
      }
      set_error("gradient launch: three-term BF16 supports padded widths up to 256, got %d", L.width);
      return MP_EUNSUPPORTED;
    case SH_32x32x16_BF16X3F:
      MP_REQUIRE(L.n_hidden == 3, MP_EUNSUPPORTED, "gradient launch: the mixed three-term shape needs 3 hidden layers");
      switch (L.width) {
        case 64: return launch_tangent_typed<S32x32x16_bf16x3f, 64>(L, xyz, n, f, g, accumulate, n_dev, stream);
        case 128: return launch_tangent_typed<S32x32x16_bf16x3f, 128>(L, xyz, n, f, g, accumulate, n_dev, stream);
        case 256: return launch_tangent_typed<S32x32x16_bf16x3f, 256>(L, xyz, n, f, g, accumulate, n_dev, stream);
      }
      set_error("gradient launch: the mixed three-term shape supports padded widths up to 256, got %d", L.width);
      return MP_EUNSUPPORTED;
  }
  set_error("gradient launch: unsupported fragment shape %d", shape);
  return MP_EUNSUPPORTED;
}

int launch_level(const mp_stack* st, int level, const EvalTarget& t, hipStream_t stream) {
  const LevelDev& L = st->lv[level].dev;
  switch (st->lv[level].shape) {
    case SH_32x32x16_BF16: return launch_width<S32x32x16_bf16>(L, st->nc, t, stream);
    case SH_32x32x16_F16: return launch_width<S32x32x16_f16>(L, st->nc, t, stream);
    case SH_16x16x4_F32: return launch_width<S16x16x4_f32>(L, st->nc, t, stream);
    case SH_32x32x16_BF16X3: return launch_split(L, t, stream);
    case SH_32x32x16_BF16X3F: return launch_split_f(L, t, stream);
  }
  set_error("level launch: unsupported fragment shape %d", st->lv[level].shape);
  return MP_EUNSUPPORTED;
}

}  // namespace mp
