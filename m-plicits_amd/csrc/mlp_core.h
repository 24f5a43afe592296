// Register-resident SIREN evaluation on gfx950 matrix cores.
//
// One wavefront (64 lanes) evaluates one SIREN level (i3d/model.py:87-129, a
// chain of Linear + sin with the frequency baked into the weights) for a tile
// of NC*ROWS points.  The hidden GEMMs are computed transposed,
//     A_i^T [width x pts] = W_i [width x width] . P_{i-1}^T [width x pts],
// with W_i as the MFMA A operand (streamed from L2 in pre-packed fragment
// order) and the activations as the B operand, which never leave the VGPRs:
// the accumulator registers of layer i, after bias + sin + conversion, are
// reused in place as the B registers of layer i+1.  That works because the
// reduction index of a GEMM may be permuted freely as long as the A operand is
// permuted the same way, so the packer stores W_{i+1} with its columns in the
// order in which layer i's accumulator registers hold the units ("virtual"
// order, see Shape::perm below).
//
// First layer (3 -> width): FP32 FMA + accurately range-reduced sine, computed
// directly in the B-operand order of layer 1.  Output layer (width -> 1): FP32
// dot product on the FP32 accumulators of the last hidden layer, reduced across
// the lane groups that hold the same point.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#ifndef MP_PF
#define MP_PF 4
#endif

namespace mp {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// MFMA shapes.  For each: ROWS x ROWS output block per instruction, KB = K per
// instruction, EPL = operand elements per lane, CREGS = accumulator registers
// per lane, GROUPS = 64 / ROWS lane groups.  Lane l holds output column
// l % ROWS and rows rowmap(l / ROWS, q) in accumulator register q; operand
// element e of lane l covers reduction index KB*kb + EPL*(l / ROWS) + e.
// ---------------------------------------------------------------------------
// SPLIT shapes carry every operand as a BF16 pair v = hi + lo (hi = bf16(v),
// lo = bf16(v - hi)) and form each product as hi.hi + hi.lo + lo.hi: three
// MFMAs per step, about 16 significant bits per operand.
enum ShapeId { SH_32x32x16_BF16 = 0, SH_32x32x16_F16 = 1, SH_16x16x32_BF16 = 2,
               SH_16x16x32_F16 = 3, SH_16x16x4_F32 = 4, SH_32x32x2_F32 = 5, SH_32x32x16_BF16X3 = 6,
               SH_32x32x16_BF16X3F = 7 };

struct S32x32x16_bf16 {
  static constexpr bool FIRST_IN_REVS = true;
  static constexpr bool SPLIT = false, LAST_F16 = false;
  static constexpr int ROWS = 32, KB = 16, EPL = 8, CREGS = 16, GROUPS = 2;
  typedef bf16x8 Frag;
  typedef f32x16 Acc;
  static __device__ inline Acc mma(Frag a, Frag b, Acc c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
  }
  static __device__ inline void set(Frag& f, int e, float v) { f[e] = (__bf16)v; }
  static __host__ __device__ constexpr int rowmap(int g, int q) { return 8 * (q >> 2) + 4 * g + (q & 3); }
};

// Three-term BF16 (see SPLIT above): the fragments and products of the BF16
// shape, each operand held as a hi/lo pair.
struct S32x32x16_bf16x3 : S32x32x16_bf16 {
  static constexpr bool SPLIT = true;
};

// Three-term BF16 for every hidden product but the last, which runs on the
// FP16 matrix cores with single FP16 weights and the activations as an FP16
// pair: a.(bh + bl), two MFMAs per step (levels with exactly two hidden
// products).  The last product's error reaches the output through one sine
// and the output weights only; the error of the level stays near 5e-5.
struct S32x32x16_bf16x3f : S32x32x16_bf16x3 {
  static constexpr bool LAST_F16 = true;
};

struct S32x32x16_f16 {
  static constexpr bool FIRST_IN_REVS = true;
  static constexpr bool SPLIT = false, LAST_F16 = false;
  static constexpr int ROWS = 32, KB = 16, EPL = 8, CREGS = 16, GROUPS = 2;
  typedef f16x8 Frag;
  typedef f32x16 Acc;
  static __device__ inline Acc mma(Frag a, Frag b, Acc c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
  }
  static __device__ inline void set(Frag& f, int e, float v) { f[e] = (_Float16)v; }
  static __host__ __device__ constexpr int rowmap(int g, int q) { return 8 * (q >> 2) + 4 * g + (q & 3); }
};

struct S16x16x32_bf16 {
  static constexpr bool FIRST_IN_REVS = true;
  static constexpr bool SPLIT = false, LAST_F16 = false;
  static constexpr int ROWS = 16, KB = 32, EPL = 8, CREGS = 4, GROUPS = 4;
  typedef bf16x8 Frag;
  typedef f32x4 Acc;
  static __device__ inline Acc mma(Frag a, Frag b, Acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
  static __device__ inline void set(Frag& f, int e, float v) { f[e] = (__bf16)v; }
  static __host__ __device__ constexpr int rowmap(int g, int q) { return 4 * g + q; }
};

struct S16x16x32_f16 {
  static constexpr bool FIRST_IN_REVS = true;
  static constexpr bool SPLIT = false, LAST_F16 = false;
  static constexpr int ROWS = 16, KB = 32, EPL = 8, CREGS = 4, GROUPS = 4;
  typedef f16x8 Frag;
  typedef f32x4 Acc;
  static __device__ inline Acc mma(Frag a, Frag b, Acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
  }
  static __device__ inline void set(Frag& f, int e, float v) { f[e] = (_Float16)v; }
  static __host__ __device__ constexpr int rowmap(int g, int q) { return 4 * g + q; }
};

struct S16x16x4_f32 {
  static constexpr bool FIRST_IN_REVS = false;
  static constexpr bool SPLIT = false, LAST_F16 = false;
  static constexpr int ROWS = 16, KB = 4, EPL = 1, CREGS = 4, GROUPS = 4;
  typedef float Frag;
  typedef f32x4 Acc;
  static __device__ inline Acc mma(Frag a, Frag b, Acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ inline void set(Frag& f, int, float v) { f = v; }
  static __host__ __device__ constexpr int rowmap(int g, int q) { return 4 * g + q; }
};

struct S32x32x2_f32 {
  static constexpr bool FIRST_IN_REVS = false;
  static constexpr bool SPLIT = false, LAST_F16 = false;
  static constexpr int ROWS = 32, KB = 2, EPL = 1, CREGS = 16, GROUPS = 2;
  typedef float Frag;
  typedef f32x16 Acc;
  static __device__ inline Acc mma(Frag a, Frag b, Acc c) {
    return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
  }
  static __device__ inline void set(Frag& f, int, float v) { f = v; }
  static __host__ __device__ constexpr int rowmap(int g, int q) { return 8 * (q >> 2) + 4 * g + (q & 3); }
};

// Shape parameters at run time (packing on the host side of the kernels).
struct ShapeInfo {
  int rows, kb, epl, cregs, groups, elem_bytes;
};
__host__ __device__ inline ShapeInfo shape_info(int id) {
  switch (id) {
    case SH_32x32x16_BF16: case SH_32x32x16_F16: case SH_32x32x16_BF16X3: case SH_32x32x16_BF16X3F:
      return {32, 16, 8, 16, 2, 2};
    case SH_16x16x32_BF16: case SH_16x16x32_F16: return {16, 32, 8, 4, 4, 2};
    case SH_16x16x4_F32: return {16, 4, 1, 4, 4, 4};
    default: return {32, 2, 1, 16, 2, 4};
  }
}
__host__ __device__ inline int shape_rowmap(const ShapeInfo& s, int g, int q) {
  return s.rows == 32 ? 8 * (q >> 2) + 4 * g + (q & 3) : 4 * g + q;
}
// Virtual reduction index v -> hidden unit u of the previous layer.
__host__ __device__ inline int shape_perm(const ShapeInfo& s, int v) {
  const int kb = v / s.kb, rem = v % s.kb, g = rem / s.epl, e = rem % s.epl;
  const int t = kb * s.epl + e;
  const int rb = t / s.cregs, q = t % s.cregs;
  return rb * s.rows + shape_rowmap(s, g, q);
}

// ---------------------------------------------------------------------------
// Sine.  v_sin_f32 takes its argument in revolutions.  The packer scales the
// hidden layers' weights and biases by 1/(2 pi), so their accumulators are
// already in revolutions (sin_rev).  The first layer sees arguments of tens of
// radians (omega_0 = 30 baked in): in the FP32 path it stays in radians and is
// reduced with a two-constant Cody-Waite step, r = a - k*2pi with
// 2pi = HI + LO, HI = fl32(2pi), LO = 2pi - HI; in the 16-bit paths its
// weights are scaled to revolutions too (accuracy of the FP32 reference's own
// rounding of a at |a| ~ 50).
// ---------------------------------------------------------------------------
__device__ inline float sin_rev(float r) { return __builtin_amdgcn_sinf(r); }

__device__ inline float sin_reduced(float a) {
  const float k = rintf(a * 0.159154943091895336f);
  float r = fmaf(k, -6.28318548202514648f, a);
  r = fmaf(k, 1.74845553146951557e-07f, r);
  return __builtin_amdgcn_sinf(r * 0.159154943091895336f);
}

__device__ inline float cos_rev(float r) { return __builtin_amdgcn_cosf(r); }

// sin and cos of a radian argument with one shared reduction.
__device__ inline void sincos_reduced(float a, float& s, float& c) {
  const float k = rintf(a * 0.159154943091895336f);
  float r = fmaf(k, -6.28318548202514648f, a);
  r = fmaf(k, 1.74845553146951557e-07f, r);
  const float rev = r * 0.159154943091895336f;
  s = __builtin_amdgcn_sinf(rev);
  c = __builtin_amdgcn_cosf(rev);
}

__device__ inline float cos_reduced(float a) {
  const float k = rintf(a * 0.159154943091895336f);
  float r = fmaf(k, -6.28318548202514648f, a);
  r = fmaf(k, 1.74845553146951557e-07f, r);
  return __builtin_amdgcn_cosf(r * 0.159154943091895336f);
}

// ---------------------------------------------------------------------------
// Packed level, as the kernels read it.
// ---------------------------------------------------------------------------
struct LevelDev {
  int width;            // padded hidden width (64, 128, 256 or 512)
  int n_hidden;         // sine layers
  const float4* w0v;    // [width]  (w_x, w_y, w_z, b) of layer 0, virtual order
  const float2* w0m;    // [width/32][64] layer 0 as A operands of two 32x32x2 FP32 MFMAs
  const void* apk;      // [n_hidden-1][width/ROWS][width/KB][64] A fragments
  const void* apkt;     // same for the transposed hidden weights (backward pass, radians)
  const float* biasc;   // [n_hidden-1][width/ROWS][GROUPS][CREGS] hidden biases, C order
  const float* wlc;     // [width/ROWS][GROUPS][CREGS] output weights, C order
  const float* wlv;     // [width] output weights, virtual order (n_hidden == 1)
  const float* blp;     // output bias (device scalar, read by the kernels)
};

template <class S, int W, int NC>
struct Mlp {
  static_assert(W % S::ROWS == 0 && W % S::KB == 0, "width must tile the MFMA shape");
  static constexpr int RBS = W / S::ROWS;
  static constexpr int KBS = W / S::KB;
  static constexpr int TILE = NC * S::ROWS;  // points per wave
  typedef typename S::Frag Frag;
  typedef typename S::Acc Acc;

  // Layer 0: a = W0 x + b0, p = sin(a), written in layer-1 B order.
  static __device__ inline void first_layer(const LevelDev& L, const float (&x)[NC][3],
                                            Frag (&b)[NC][KBS], int g) {
#pragma unroll
    for (int kb = 0; kb < KBS; ++kb) {
#pragma unroll
      for (int e = 0; e < S::EPL; ++e) {
        const float4 w = L.w0v[kb * S::KB + g * S::EPL + e];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          const float a = fmaf(w.z, x[c][2], fmaf(w.y, x[c][1], fmaf(w.x, x[c][0], w.w)));
          S::set(b[c][kb], e, S::FIRST_IN_REVS ? sin_rev(a) : sin_reduced(a));
        }
      }
    }
  }

  // Prefetch depth of the A-fragment stream.  The loads are issued PF
  // fragments ahead of their MFMA; the empty asm with a memory clobber stops
  // the compiler from hoisting the whole row block's loads (which would cost
  // KBS * 4 VGPRs and spill the activations).
  static constexpr int PF = MP_PF;

  // One weight layer: for every row block rb, acc = b + W . bin (b = 0 when
  // bc is null: the transposed GEMMs of the backward pass), then epi(rb, acc)
  // consumes the FP32 accumulators.
  template <class Epi>
  static __device__ inline void gemm(const Frag* __restrict__ A, const float* __restrict__ bc,
                                     const Frag (&bin)[NC][KBS], int lane, int g, Epi epi) {
    Frag pf[PF];
#pragma unroll
    for (int d = 0; d < PF; ++d) pf[d] = A[d * 64 + lane];
#pragma unroll
    for (int rb = 0; rb < RBS; ++rb) {
      // the bias enters as the C operand of the first MFMA of the row block
      Acc bias;
      if (bc) {
        const float* brow = bc + (rb * S::GROUPS + g) * S::CREGS;
#pragma unroll
        for (int q = 0; q < S::CREGS; ++q) bias[q] = brow[q];
      } else {
#pragma unroll
        for (int q = 0; q < S::CREGS; ++q) bias[q] = 0.f;
      }
      Acc acc[NC];
#pragma unroll
      for (int kb = 0; kb < KBS; ++kb) {
        const int idx = rb * KBS + kb;
        const Frag a = pf[idx % PF];
        if (idx + PF < RBS * KBS) pf[idx % PF] = A[(idx + PF) * 64 + lane];
#pragma unroll
        for (int c = 0; c < NC; ++c) acc[c] = S::mma(a, bin[c][kb], kb == 0 ? bias : acc[c]);
        asm volatile("" ::: "memory");
      }
      epi(rb, acc);
    }
  }

  // One hidden layer: bout = sin(W . bin + b), in the next layer's B order.
  static __device__ inline void hidden(const Frag* __restrict__ A, const float* __restrict__ bc,
                                       const Frag (&bin)[NC][KBS], Frag (&bout)[NC][KBS],
                                       int lane, int g) {
    gemm(A, bc, bin, lane, g, [&](int rb, Acc (&acc)[NC]) {
#pragma unroll
      for (int q = 0; q < S::CREGS; ++q) {
        const int t = rb * S::CREGS + q;
#pragma unroll
        for (int c = 0; c < NC; ++c) S::set(bout[c][t / S::EPL], t % S::EPL, sin_rev(acc[c][q]));
      }
    });
  }

  // Last hidden layer fused with the output layer: f = w_L . sin(W . bin + b).
  // Returns this lane group's partial sum (FP32).
  static __device__ inline void last(const Frag* __restrict__ A, const float* __restrict__ bc,
                                     const float* __restrict__ wlc, const Frag (&bin)[NC][KBS],
                                     float (&f)[NC], int lane, int g) {
#pragma unroll
    for (int c = 0; c < NC; ++c) f[c] = 0.f;
    gemm(A, bc, bin, lane, g, [&](int rb, Acc (&acc)[NC]) {
      const float* wrow = wlc + (rb * S::GROUPS + g) * S::CREGS;
#pragma unroll
      for (int q = 0; q < S::CREGS; ++q) {
        const float wv = wrow[q];
#pragma unroll
        for (int c = 0; c < NC; ++c) f[c] = fmaf(wv, sin_rev(acc[c][q]), f[c]);
      }
    });
  }

  // Sum the partial outputs of the lane groups holding the same point.
  static __device__ inline float group_sum(float v) {
    if (S::GROUPS == 4) v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 32);
    return v;
  }

  // f_level(x) for the NC points of this lane (every lane group returns the
  // same value for its column).
  static __device__ inline void eval(const LevelDev& L, const float (&x)[NC][3], float (&f)[NC]) {
    const int lane = threadIdx.x & 63;
    const int g = lane / S::ROWS;
    const float bl = *L.blp;
    if (L.n_hidden == 1) {
      // 3 -> width -> 1: the dot product runs on the FP32 sines.
#pragma unroll
      for (int c = 0; c < NC; ++c) f[c] = 0.f;
      for (int v0 = 0; v0 < W; v0 += S::KB) {
#pragma unroll
        for (int e = 0; e < S::EPL; ++e) {
          const int v = v0 + g * S::EPL + e;
          const float4 w = L.w0v[v];
          const float wl = L.wlv[v];
#pragma unroll
          for (int c = 0; c < NC; ++c) {
            const float a = fmaf(w.z, x[c][2], fmaf(w.y, x[c][1], fmaf(w.x, x[c][0], w.w)));
            f[c] = fmaf(wl, S::FIRST_IN_REVS ? sin_rev(a) : sin_reduced(a), f[c]);
          }
        }
      }
#pragma unroll
      for (int c = 0; c < NC; ++c) f[c] = group_sum(f[c]) + bl;
      return;
    }
    Frag b0[NC][KBS], b1[NC][KBS];
    first_layer(L, x, b0, g);
    const Frag* A = reinterpret_cast<const Frag*>(L.apk);
    const size_t a_step = (size_t)RBS * KBS * 64;
    const float* bc = L.biasc;
    const int n_gemm = L.n_hidden - 1;
    // ping-pong between b0 and b1 so no register copies are needed
    int j = 0;
    for (; j + 2 < n_gemm; j += 2) {
      hidden(A + j * a_step, bc + (size_t)j * W, b0, b1, lane, g);
      hidden(A + (j + 1) * a_step, bc + (size_t)(j + 1) * W, b1, b0, lane, g);
    }
    if (n_gemm - j == 2) {
      hidden(A + j * a_step, bc + (size_t)j * W, b0, b1, lane, g);
      last(A + (j + 1) * a_step, bc + (size_t)(j + 1) * W, L.wlc, b1, f, lane, g);
    } else {
      last(A + j * a_step, bc + (size_t)j * W, L.wlc, b0, f, lane, g);
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) f[c] = group_sum(f[c]) + bl;
  }
};

// ---------------------------------------------------------------------------
// Extraction grid coordinates.  MP_GRID_REFERENCE reproduces
// gen_mc_coordinate_grid (i3d/meshing.py:55-63) as torch evaluates it: the
// int64 index is divided by N in float32 true division (`/` on a LongTensor),
// reduced with fmod-based remainder, then scaled and shifted as two separately
// rounded float32 operations.  Column 0 takes the slowest index component.
// ---------------------------------------------------------------------------
__device__ inline void grid_coord(int64_t idx, int N, int mode, float h, float out[3]) {
  float vx, vy, vz;
  if ((N & (N - 1)) == 0 && N <= 1024) {
    // N = 2^k: dividing by N is an exact scaling and, for 0 <= a < N^2 (so
    // ulp(a) < N), fmod(a, N) = a - floor(a / N) N is exact as one fma: the
    // same values as the general expressions below, without 64-bit divisions
    // or fmodf
    const int k = __ffs(N) - 1;
    const float fN = (float)N, inv = 1.0f / fN;
    vz = (float)(int)(idx & (N - 1));
    if (mode == 0) {
      const float t = (float)idx * inv;
      vy = fmaf(-floorf(t * inv), fN, t);
      const float u = t * inv;
      vx = fmaf(-floorf(u * inv), fN, u);
    } else {
      vy = (float)(int)((idx >> k) & (N - 1));
      vx = (float)(idx >> (2 * k));
    }
  } else if (mode == 0) {
    const float fN = (float)N;
    vz = (float)(idx % N);
    const float t = __fdiv_rn((float)idx, fN);
    vy = fmodf(t, fN);
    vx = fmodf(__fdiv_rn(t, fN), fN);
  } else {
    const int64_t nn = (int64_t)N * N;
    vx = (float)(idx / nn);
    vy = (float)((idx / N) % N);
    vz = (float)(idx % N);
  }
  out[0] = __fadd_rn(__fmul_rn(vx, h), -1.0f);
  out[1] = __fadd_rn(__fmul_rn(vy, h), -1.0f);
  out[2] = __fadd_rn(__fmul_rn(vz, h), -1.0f);
}

}  // namespace mp
