// Value and input gradient of one SIREN level in one pass: the chain rule
//     grad f = W_0^T diag(cos a_0) W_1^T diag(cos a_1) ... W_G^T diag(cos a_G) w_L
// evaluated backward on the matrix cores (transposed GEMMs with the gradient
// vectors as the B operand), after a forward pass that keeps cos(a_j) per
// hidden unit.  Replaces the renderer's Normals::inference (three forward
// tangent passes per level, renderer/cuda/gemms.cu:263-337 with
// CalculateG0 / HadamardGi, kernels.cu:277-313) and the autograd gradient of
// i3d/diff_operators.py:259-265.
//
// The cosines live in LDS (one [G][KBS][64] fragment array per wave), stored
// in the register order in which both the forward accumulators and the
// backward accumulators hold the units, so the backward epilogue reads them
// lane-local with one 16-byte LDS access per fragment.
#include "eval_dispatch.h"

namespace mp {

template <class S, int W, int G>
struct MlpGrad {
  static_assert(S::CREGS % S::EPL == 0, "accumulator rows must fill whole fragments");
  typedef Mlp<S, W, 1> M;
  static constexpr int KBS = M::KBS, RBS = M::RBS;
  typedef typename S::Frag Frag;
  typedef typename S::Acc Acc;
  static constexpr size_t LDS_PER_WAVE = (size_t)G * KBS * 64 * sizeof(Frag);

  static __device__ inline float elem(const Frag& f, int e) {
    if constexpr (S::EPL == 1) return (float)f;
    else return (float)f[e];
  }

  // f and grad f at the lane's point; cs = this wave's LDS cosine store.
  static __device__ void eval(const LevelDev& L, const float (&x)[1][3], Frag* __restrict__ cs,
                              float& f_out, float (&gr)[3]) {
    const int lane = threadIdx.x & 63;
    const int g = lane / S::ROWS;
    Frag b0[1][KBS], b1[1][KBS];
    // layer 0: sines to b0, cosines to cs[0]
#pragma unroll
    for (int kb = 0; kb < KBS; ++kb) {
      Frag cf;
#pragma unroll
      for (int e = 0; e < S::EPL; ++e) {
        const float4 w = L.w0v[kb * S::KB + g * S::EPL + e];
        const float a = fmaf(w.z, x[0][2], fmaf(w.y, x[0][1], fmaf(w.x, x[0][0], w.w)));
        S::set(b0[0][kb], e, S::FIRST_IN_REVS ? sin_rev(a) : sin_reduced(a));
        S::set(cf, e, S::FIRST_IN_REVS ? cos_rev(a) : cos_reduced(a));
      }
      cs[kb * 64 + lane] = cf;
    }
    const Frag* A = reinterpret_cast<const Frag*>(L.apk);
    const Frag* At = reinterpret_cast<const Frag*>(L.apkt);
    const size_t a_step = (size_t)RBS * KBS * 64;
    float f = 0.f;
    // forward through the hidden GEMMs; the input of GEMM j is b0 for odd j
    auto fwd_hidden = [&](int j, const Frag(&bin)[1][KBS], Frag(&bout)[1][KBS]) {
      Frag* cj = cs + (size_t)j * KBS * 64;
      M::gemm(A + (j - 1) * a_step, L.biasc + (size_t)(j - 1) * W, bin, lane, g,
              [&](int rb, Acc(&acc)[1]) {
#pragma unroll
                for (int q = 0; q < S::CREGS; ++q) {
                  const int t = rb * S::CREGS + q;
                  S::set(bout[0][t / S::EPL], t % S::EPL, sin_rev(acc[0][q]));
                }
#pragma unroll
                for (int k = 0; k < S::CREGS / S::EPL; ++k) {
                  Frag cf;
#pragma unroll
                  for (int e = 0; e < S::EPL; ++e) S::set(cf, e, cos_rev(acc[0][k * S::EPL + e]));
                  cj[(rb * S::CREGS / S::EPL + k) * 64 + lane] = cf;
                }
              });
    };
    // last hidden GEMM G: f = w_L . sin(acc) + b_L, delta = w_L * cos(acc)
    auto fwd_last = [&](const Frag(&bin)[1][KBS], Frag(&dout)[1][KBS]) {
      M::gemm(A + (G - 1) * a_step, L.biasc + (size_t)(G - 1) * W, bin, lane, g,
              [&](int rb, Acc(&acc)[1]) {
                const float* wrow = L.wlc + (rb * S::GROUPS + g) * S::CREGS;
#pragma unroll
                for (int q = 0; q < S::CREGS; ++q) {
                  const int t = rb * S::CREGS + q;
                  const float wv = wrow[q];
                  f = fmaf(wv, sin_rev(acc[0][q]), f);
                  S::set(dout[0][t / S::EPL], t % S::EPL, wv * cos_rev(acc[0][q]));
                }
              });
    };
    // backward GEMM j: din (gradient w.r.t. a_j) -> dout (w.r.t. a_{j-1});
    // for j == 1 the epilogue folds W_0^T into the input gradient.
    float gx = 0.f, gy = 0.f, gz = 0.f;
    auto bwd = [&](int j, const Frag(&din)[1][KBS], Frag(&dout)[1][KBS]) {
      const Frag* cprev = cs + (size_t)(j - 1) * KBS * 64;
      M::gemm(At + (j - 1) * a_step, nullptr, din, lane, g, [&](int rb, Acc(&acc)[1]) {
#pragma unroll
        for (int k = 0; k < S::CREGS / S::EPL; ++k) {
          const int kb = rb * S::CREGS / S::EPL + k;
          const Frag cf = cprev[kb * 64 + lane];
#pragma unroll
          for (int e = 0; e < S::EPL; ++e) {
            const float ab = elem(cf, e) * acc[0][k * S::EPL + e];
            if (j == 1) {
              const float4 w = L.w0v[kb * S::KB + g * S::EPL + e];
              gx = fmaf(ab, w.x, gx);
              gy = fmaf(ab, w.y, gy);
              gz = fmaf(ab, w.z, gz);
            } else {
              S::set(dout[0][kb], e, ab);
            }
          }
        }
      });
    };
    if constexpr (G == 1) {
      fwd_last(b0, b1);
      bwd(1, b1, b0);
    } else if constexpr (G == 2) {
      fwd_hidden(1, b0, b1);
      fwd_last(b1, b0);
      bwd(2, b0, b1);
      bwd(1, b1, b0);
    } else {
      fwd_hidden(1, b0, b1);
      fwd_hidden(2, b1, b0);
      fwd_last(b0, b1);
      bwd(3, b1, b0);
      bwd(2, b0, b1);
      bwd(1, b1, b0);
    }
    const float scale = S::FIRST_IN_REVS ? 6.28318530717958648f : 1.0f;  // w0v is in revolutions
    f_out = M::group_sum(f) + *L.blp;
    gr[0] = M::group_sum(gx) * scale;
    gr[1] = M::group_sum(gy) * scale;
    gr[2] = M::group_sum(gz) * scale;
  }
};

// One point per lane column; `waves` waves per block share the dynamic LDS.
template <class S, int W, int G>
__global__ __launch_bounds__(256) void k_level_grad(LevelDev L, const float* __restrict__ xyz, int64_t n, float* f_out,
                             float* g_out, int accumulate, const unsigned long long* n_dev) {
  typedef MlpGrad<S, W, G> MG;
  if (n_dev) n = (int64_t)*n_dev < n ? (int64_t)*n_dev : n;  // device count, n = upper bound
  extern __shared__ __align__(16) unsigned char lds[];
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int waves = blockDim.x >> 6;
  const int64_t base = ((int64_t)blockIdx.x * waves + wv) * S::ROWS;
  if (base >= n) return;
  const int64_t i = base + lane % S::ROWS;
  const bool valid = i < n;
  const int64_t ii = valid ? i : n - 1;
  float x[1][3] = {{xyz[3 * ii], xyz[3 * ii + 1], xyz[3 * ii + 2]}};
  typename MG::Frag* cs = reinterpret_cast<typename MG::Frag*>(lds + wv * MG::LDS_PER_WAVE);
  float f, gr[3];
  MG::eval(L, x, cs, f, gr);
  if (lane < S::ROWS && valid) {
    if (accumulate) {
      if (f_out) f_out[ii] += f;
      g_out[3 * ii + 0] += gr[0];
      g_out[3 * ii + 1] += gr[1];
      g_out[3 * ii + 2] += gr[2];
    } else {
      if (f_out) f_out[ii] = f;
      g_out[3 * ii + 0] = gr[0];
      g_out[3 * ii + 1] = gr[1];
      g_out[3 * ii + 2] = gr[2];
    }
  }
}

// 3 -> W -> 1 (no hidden GEMM): grad = sum_u w_L[u] cos(a0[u]) W0[u].
template <class S, int W>
__global__ __launch_bounds__(256) void k_level_grad1(LevelDev L, const float* __restrict__ xyz, int64_t n, float* f_out,
                              float* g_out, int accumulate, const unsigned long long* n_dev) {
  typedef Mlp<S, W, 1> M;
  if (n_dev) n = (int64_t)*n_dev < n ? (int64_t)*n_dev : n;
  const int lane = threadIdx.x & 63;
  const int64_t base = ((int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * S::ROWS;
  if (base >= n) return;
  const int64_t i = base + lane % S::ROWS;
  const bool valid = i < n;
  const int64_t ii = valid ? i : n - 1;
  const float x0 = xyz[3 * ii], x1 = xyz[3 * ii + 1], x2 = xyz[3 * ii + 2];
  const int g = lane / S::ROWS;
  float f = 0.f, gx = 0.f, gy = 0.f, gz = 0.f;
  for (int v0 = 0; v0 < W; v0 += S::KB)
    for (int e = 0; e < S::EPL; ++e) {
      const int v = v0 + g * S::EPL + e;
      const float4 w = L.w0v[v];
      const float wl = L.wlv[v];
      const float a = fmaf(w.z, x2, fmaf(w.y, x1, fmaf(w.x, x0, w.w)));
      f = fmaf(wl, S::FIRST_IN_REVS ? sin_rev(a) : sin_reduced(a), f);
      const float d = wl * (S::FIRST_IN_REVS ? cos_rev(a) : cos_reduced(a));
      gx = fmaf(d, w.x, gx);
      gy = fmaf(d, w.y, gy);
      gz = fmaf(d, w.z, gz);
    }
  const float scale = S::FIRST_IN_REVS ? 6.28318530717958648f : 1.0f;
  f = M::group_sum(f) + *L.blp;
  gx = M::group_sum(gx) * scale;
  gy = M::group_sum(gy) * scale;
  gz = M::group_sum(gz) * scale;
  if (lane < S::ROWS && valid) {
    if (accumulate) {
      if (f_out) f_out[ii] += f;
      g_out[3 * ii] += gx; g_out[3 * ii + 1] += gy; g_out[3 * ii + 2] += gz;
    } else {
      if (f_out) f_out[ii] = f;
      g_out[3 * ii] = gx; g_out[3 * ii + 1] = gy; g_out[3 * ii + 2] = gz;
    }
  }
}

static constexpr size_t LDS_BUDGET = 160 * 1024;

template <class S, int W, int G>
static int launch_grad_typed(const LevelDev& L, const float* xyz, int64_t n, float* f, float* g,
                             int acc, const unsigned long long* n_dev, hipStream_t stream) {
  typedef MlpGrad<S, W, G> MG;
  int waves = (int)(LDS_BUDGET / MG::LDS_PER_WAVE);
  waves = waves > 4 ? 4 : waves;
  MP_REQUIRE(waves >= 1, MP_EUNSUPPORTED, "gradient: width %d with %d hidden GEMMs needs %zu B of LDS per wave",
             W, G, MG::LDS_PER_WAVE);
  const int64_t per_block = (int64_t)waves * S::ROWS;
  const int64_t blocks = (n + per_block - 1) / per_block;
  const size_t lds = (size_t)waves * MG::LDS_PER_WAVE;
  static std::atomic<uint64_t> attr_devs{0};
  if (lds > 64 * 1024) MP_HIP(set_lds_once(attr_devs, reinterpret_cast<const void*>(&k_level_grad<S, W, G>), (int)lds));
  hipLaunchKernelGGL((k_level_grad<S, W, G>), dim3((unsigned)blocks), dim3(64 * waves), lds, stream, L,
                     xyz, n, f, g, acc, n_dev);
  MP_LAUNCHED("k_level_grad");
  return MP_OK;
}

template <class S, int W>
static int launch_grad_depth(const LevelDev& L, const float* xyz, int64_t n, float* f, float* g, int acc,
                             const unsigned long long* n_dev, hipStream_t stream) {
  switch (L.n_hidden - 1) {
    case 0: {
      const int64_t blocks = (n + 4 * S::ROWS - 1) / (4 * S::ROWS);
      hipLaunchKernelGGL((k_level_grad1<S, W>), dim3((unsigned)blocks), dim3(256), 0, stream, L, xyz, n, f,
                         g, acc, n_dev);
      MP_LAUNCHED("k_level_grad1");
      return MP_OK;
    }
    case 1: return launch_grad_typed<S, W, 1>(L, xyz, n, f, g, acc, n_dev, stream);
    case 2: return launch_grad_typed<S, W, 2>(L, xyz, n, f, g, acc, n_dev, stream);
    case 3: return launch_grad_typed<S, W, 3>(L, xyz, n, f, g, acc, n_dev, stream);
  }
  set_error("gradient: levels with more than 4 sine layers are not supported (got %d)", L.n_hidden);
  return MP_EUNSUPPORTED;
}

template <class S>
static int launch_grad_width(const LevelDev& L, const float* xyz, int64_t n, float* f, float* g, int acc,
                             const unsigned long long* n_dev, hipStream_t stream) {
  switch (L.width) {
    case 64: return launch_grad_depth<S, 64>(L, xyz, n, f, g, acc, n_dev, stream);
    case 128: return launch_grad_depth<S, 128>(L, xyz, n, f, g, acc, n_dev, stream);
    case 256: return launch_grad_depth<S, 256>(L, xyz, n, f, g, acc, n_dev, stream);
    case 512: return launch_grad_depth<S, 512>(L, xyz, n, f, g, acc, n_dev, stream);
  }
  set_error("gradient: unsupported padded width %d", L.width);
  return MP_EUNSUPPORTED;
}

int launch_level_grad(const mp_stack* st, int level, const float* xyz, int64_t n, float* f, float* g,
                      int accumulate, hipStream_t stream, const unsigned long long* n_dev) {
  const LevelDev& L = st->lv[level].dev;
  // the level's own shape: a three-term BF16 stack runs its 512-wide levels in FP16
  const int shape = st->lv[level].shape;
  // forward mode where the reverse-mode kernel spills: 16-bit width 512 (and
  // 256 with one hidden GEMM), FP32 widths >= 256, and every three-term BF16
  // level (its operand pairs fill the registers the reverse mode would need)
  const int gemms = L.n_hidden - 1;
  const bool fp32 = shape == SH_16x16x4_F32;
  if (shape == SH_32x32x16_BF16X3 || shape == SH_32x32x16_BF16X3F ||
      (gemms >= 1 && (fp32 ? L.width >= 256 : (L.width == 512 || (L.width == 256 && gemms == 1)))))
    return launch_level_tangent(shape, L, xyz, n, f, g, accumulate, stream, n_dev);
  switch (shape) {
    case SH_32x32x16_BF16: return launch_grad_width<S32x32x16_bf16>(L, xyz, n, f, g, accumulate, n_dev, stream);
    case SH_32x32x16_F16: return launch_grad_width<S32x32x16_f16>(L, xyz, n, f, g, accumulate, n_dev, stream);
    case SH_16x16x4_F32: return launch_grad_width<S16x16x4_f32>(L, xyz, n, f, g, accumulate, n_dev, stream);
  }
  set_error("gradient: unsupported fragment shape %d", shape);
  return MP_EUNSUPPORTED;
}

int launch_grad(const mp_stack* st, int lvl0, int lvl1, const float* xyz, int64_t n, float* sdf, float* grad,
                hipStream_t stream) {
  for (int l = lvl0; l < lvl1; ++l) {
    int rc = launch_level_grad(st, l, xyz, n, sdf, grad, l > lvl0, stream);
    if (rc != MP_OK) return rc;
  }
  return MP_OK;
}

}  // namespace mp
