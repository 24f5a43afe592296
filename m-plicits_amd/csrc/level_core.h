// One SIREN level evaluated by one wave on the matrix cores, laid out for one
// wave per SIMD (i3d/model.py:87-129 with the frequency baked in).
//
// Chained-layer scheme (shared with mlp_core.h): the hidden GEMMs run
// transposed, A = W_i (pre-packed fragments, streamed), B = activations of the
// tile's NC*ROWS points held in VGPRs; accumulator register q of row block rb
// is reduction index t = CREGS*rb + q of the next layer, i.e. element t % EPL
// of its B fragment t / EPL (the packer permutes the next layer's columns to
// match).
//
// Differences from the Mlp chain:
//  - 16-bit shapes: layer 0 (3 -> width) runs on the FP32 matrix cores, two
//    32x32x2 MFMAs per 32-row block over (x, y, z, 1) (weights pre-scaled to
//    revolutions); FP32 shape: range-reduced FMA + sine on the VALU, written
//    straight into the B registers (its accuracy is the FP32 path's);
//  - a layer that feeds another hidden layer hands its output over through the
//    wave's own LDS region (B registers -> LDS -> B registers) instead of a
//    second register set, leaving VGPRs for a deep A look-ahead and two
//    accumulator sets;
//  - the A fragments of all hidden layers form one continuous stream.  16-bit
//    shapes at width >= 128 (when the caller allows workgroup barriers): the
//    four waves of a workgroup copy it, 16 KB stage by 16 KB stage, into a
//    two-slot LDS ring with direct global->LDS loads a whole stage ahead and
//    read it back a few steps ahead.  Otherwise: a register ring of PF
//    fragments loaded from global memory, continuous across layers;
//  - the epilogue of row block rb (bias is the MFMA C operand; sine, packing,
//    LDS store or output dot product) is issued in slices between the MFMAs
//    of row block rb + 1, so the transcendental unit and the matrix core
//    overlap as far as one wave's in-order issue allows.
#pragma once

#include "eval_dispatch.h"

namespace mp {

typedef __attribute__((address_space(3))) void lds_void;

// WAVES: waves of the workgroup that share the staged weight stream (four: one
// per SIMD; eight: two per SIMD, for single-column tiles whose registers and
// LDS hand-off regions fit twice).
// TANGENT: forward-mode input gradient.  Columns 4 p .. 4 p + 3 of a tile
// carry point p's value and its derivatives along x, y, z: the first layer
// sees (x, y, z, 1) in the value column and (e_t, 0) in tangent column t, the
// biases reach the value column only, and tangent column t activates to
// cos(a_0) * a_t with a_0 taken from the value column of its quad (the
// accumulators are in revolutions: 2 pi cos_rev(r_0) r_t).  The output of
// tangent column t is df/dx_t.
// S::SPLIT (three-term BF16): every B fragment has a lo partner (registers
// b[c][LO + kb], hand-off region after the hi fragments) and every A step two
// fragments in the stream (hi, then lo); a step issues hi.hi, hi.lo and lo.hi.
// S::LAST_F16 (levels with two hidden products): the first product hands over
// its activations as FP16 pairs, and the last product's stream holds one FP16
// fragment per step; a step there issues a.hi and a.lo on the FP16 cores.
template <class S, int W, int NC, bool ALLOW_STAGED, int WAVES = 4, bool TANGENT = false>
struct LevelCore {
  typedef typename S::Frag Frag;
  typedef typename S::Acc Acc;
  static constexpr int ROWS = S::ROWS, KB = S::KB, CREGS = S::CREGS, EPL = S::EPL, GROUPS = S::GROUPS;
  static constexpr bool MFMA_FIRST = S::FIRST_IN_REVS;  // 16-bit shapes (32-row C layout)
  static_assert(!MFMA_FIRST || ROWS == 32, "the MFMA first layer uses the 32x32 C layout");
  static constexpr int RBS = W / ROWS;           // row blocks per layer
  static constexpr int KBS = W / KB;             // K blocks per layer
  static constexpr int STEPS = RBS * KBS;        // MFMA steps per layer
  static constexpr bool SPLIT = S::SPLIT;
  static constexpr bool LF = S::LAST_F16;
  static constexpr int BP = SPLIT ? 2 : 1;       // fragments per operand (hi, lo)
  static constexpr int LO = SPLIT ? KBS : 0;     // offset of the lo B fragments
  static constexpr int LSTEPS = STEPS * BP;      // A fragments per layer
  static constexpr int TILE = NC * ROWS;         // points per wave
  static constexpr int HANDOFF_FRAGS = BP * NC * KBS * 64;  // per wave, in fragments

  static constexpr bool STAGED = ALLOW_STAGED && sizeof(Frag) == 16 && LSTEPS % 32 == 0;
  static constexpr int SF = 16;  // fragments per stage
  // LDS read-ahead in fragments (three-term: four steps of hi and lo fragments)
  static constexpr int D = SPLIT ? 8 : 4;
  static constexpr int RING_FRAGS = STAGED ? 2 * SF * 64 : 0;
  static constexpr int PF_REG = sizeof(Frag) == 16 ? 8 : 32;
  static constexpr int PF = STAGED ? D : (LSTEPS < PF_REG ? LSTEPS : PF_REG);
  static_assert(LSTEPS % PF == 0, "the look-ahead ring must tile a layer");
  static_assert((size_t)(STAGED ? 2 * SF : PF) * 64 * sizeof(Frag) <= APK_PAD,
                "the stream reads past the padding");
  static_assert(SF % WAVES == 0, "every wave copies the same share of a stage");
  // LDS of a workgroup: ring, then one handoff region per wave
  static constexpr size_t LDS_BYTES = (size_t)(RING_FRAGS + WAVES * HANDOFF_FRAGS) * sizeof(Frag);

  // epilogue slices: accumulator registers retired per K step
  static constexpr int QPK = KBS >= CREGS ? 1 : CREGS / KBS;
  static constexpr int KPQ = KBS >= CREGS ? KBS / CREGS : 1;

  struct Stream {
    // A is deliberately not __restrict__: the empty asm with a memory clobber
    // after every step pins each look-ahead load to its step (through a
    // restrict pointer the loads may sink to their uses and expose their
    // latency).
    const Frag* A;  // hidden-layer fragments, layer-major, APK_PAD readable bytes after the end
    Frag* ring;     // LDS ring (staged)
    int lane, wave;
    Frag pf[PF];

    // Stage `stage` into ring slot `slot`: each of the WAVES waves copies
    // SF / WAVES of the stage's 16 fragments (one 16-byte direct load per lane
    // each).
    __device__ __forceinline__ void copy_stage(int64_t stage, int slot) {
#pragma unroll
      for (int i = 0; i < SF / WAVES; ++i) {
        const int f = (SF / WAVES) * wave + i;
        __builtin_amdgcn_global_load_lds((const void*)(A + (stage * SF + f) * 64 + lane),
                                         (lds_void*)(ring + (slot * SF + f) * 64), 16, 0, 0);
      }
    }
    __device__ __forceinline__ void barrier() {
      asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
      __syncthreads();
    }
    __device__ __forceinline__ void start() {
      if constexpr (STAGED) {
        copy_stage(0, 0);
        copy_stage(1, 1);
      } else {
#pragma unroll
        for (int d = 0; d < PF; ++d) pf[d] = A[d * 64 + lane];
      }
    }
    __device__ __forceinline__ void begin() {
      if constexpr (STAGED) {
        barrier();
#pragma unroll
        for (int d = 0; d < D; ++d) pf[d] = ring[d * 64 + lane];
      }
    }
    // Fragment s of hidden layer j's stream; issues the look-ahead.
    __device__ __forceinline__ Frag next(int j, int s) {
      if constexpr (STAGED) {
        const int t = s + D;  // step whose fragment is read now
        if (t % SF == 0) {
          // every wave is past its last read of the slot about to be refilled
          barrier();
          copy_stage((int64_t)j * (LSTEPS / SF) + t / SF + 1, (t / SF + 1) % 2);
        }
        const Frag a = pf[s % D];
        pf[s % D] = ring[((t / SF) % 2 * SF + t % SF) * 64 + lane];
        return a;
      } else {
        const Frag a = pf[s % PF];
        pf[s % PF] = A[((size_t)j * LSTEPS + s + PF) * 64 + lane];
        return a;
      }
    }
    // Outstanding look-ahead copies land in the ring; let them finish before
    // the workgroup can retire and its LDS be reassigned.
    __device__ __forceinline__ void finish() {
      if constexpr (STAGED) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    }
  };

  static __device__ __forceinline__ void loadc(const float* __restrict__ p, Acc& v) {
    if constexpr (CREGS % 4 == 0) {
      const float4* p4 = reinterpret_cast<const float4*>(p);
#pragma unroll
      for (int i = 0; i < CREGS / 4; ++i) {
        const float4 t = p4[i];
        v[4 * i] = t.x; v[4 * i + 1] = t.y; v[4 * i + 2] = t.z; v[4 * i + 3] = t.w;
      }
    } else {
#pragma unroll
      for (int q = 0; q < CREGS; ++q) v[q] = p[q];
    }
  }

  static __device__ __forceinline__ float group_sum(float v) {
    if (GROUPS == 4) v += __shfl_xor(v, 16);
    return v + __shfl_xor(v, 32);
  }

  // Tangent index of this lane's column (0: the value column).
  static __device__ __forceinline__ int tangent_of(int lane) { return TANGENT ? (lane % ROWS) & 3 : 0; }

  // Activation of an accumulator in revolutions: sin, or in a tangent column
  // the derivative cos(r_0) 2 pi r_t (every lane must call it: it shuffles).
  static __device__ __forceinline__ float act_rev(float v, int lane) {
    if constexpr (TANGENT) {
      // value column of the quad (lanes 4p .. 4p + 3): DPP quad broadcast of its first lane
      const float r0 = __builtin_bit_cast(
          float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x00, 0xF, 0xF, false));
      const bool value = tangent_of(lane) == 0;
      if constexpr (MFMA_FIRST) {
        // 16-bit shapes: one transcendental per lane, cos(2 pi r) = sin(2 pi (r + 1/4)) (the
        // rounding of r + 1/4, half an ulp of r, stays far below the rounding of the 16-bit
        // operands every tangent has passed through)
        const float s = sin_rev(value ? v : r0 + 0.25f);
        return value ? s : s * (6.28318530717958648f * v);
      } else {
        return value ? sin_rev(v) : cos_rev(r0) * (6.28318530717958648f * v);
      }
    } else {
      return sin_rev(v);
    }
  }

  // Element e of an operand: hi, and in SPLIT mode the residual into lo.
  static __device__ __forceinline__ void set2(Frag& hi, Frag& lo, int e, float v) {
    S::set(hi, e, v);
    if constexpr (SPLIT) S::set(lo, e, v - (float)hi[e]);
  }
  // The same as an FP16 pair (operands of a LAST_F16 level's last product).
  static __device__ __forceinline__ void set2h(Frag& hi, Frag& lo, int e, float v) {
    f16x8 h = __builtin_bit_cast(f16x8, hi), l = __builtin_bit_cast(f16x8, lo);
    h[e] = (_Float16)v;
    l[e] = (_Float16)(v - (float)h[e]);
    hi = __builtin_bit_cast(Frag, h);
    lo = __builtin_bit_cast(Frag, l);
  }

  // Layer-0 operands of the FP32 matrix cores (16-bit shapes), loaded by the
  // caller ahead of its own set-up (point coordinates) so that their latency
  // overlaps it.
  struct First {
    float2 w[MFMA_FIRST ? RBS : 1];
  };
  static __device__ __forceinline__ First load_first(const LevelDev& L) {
    First p;
    if constexpr (MFMA_FIRST) {
      const int lane = threadIdx.x & 63;
#pragma unroll
      for (int rb = 0; rb < RBS; ++rb) p.w[rb] = L.w0m[rb * 64 + lane];
    }
    return p;
  }

  // Layer 0 into the B registers of layer 1.
  static __device__ __forceinline__ void first(const LevelDev& L, const float (&x)[NC][3], Frag (&b)[NC][BP * KBS],
                                      int lane, int g, const First& w0) {
    if constexpr (MFMA_FIRST) {
      float bxy[NC], bz1[NC];
      const int tl = tangent_of(lane);
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        if (TANGENT && tl != 0) {
          // unit direction e_t, no bias input
          bxy[c] = g == 0 ? (tl == 1 ? 1.f : 0.f) : (tl == 2 ? 1.f : 0.f);
          bz1[c] = g == 0 ? (tl == 3 ? 1.f : 0.f) : 0.f;
        } else {
          bxy[c] = g == 0 ? x[c][0] : x[c][1];
          bz1[c] = g == 0 ? x[c][2] : 1.f;
        }
      }
      Acc zero;
#pragma unroll
      for (int q = 0; q < CREGS; ++q) zero[q] = 0.f;
#pragma unroll
      for (int rb = 0; rb < RBS; ++rb) {
        const float2 a = w0.w[rb];
        Acc acc[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          acc[c] = __builtin_amdgcn_mfma_f32_32x32x2f32(a.x, bxy[c], zero, 0, 0, 0);
          acc[c] = __builtin_amdgcn_mfma_f32_32x32x2f32(a.y, bz1[c], acc[c], 0, 0, 0);
        }
#pragma unroll
        for (int q = 0; q < CREGS; ++q) {
          const int t = rb * CREGS + q;
#pragma unroll
          for (int c = 0; c < NC; ++c) set2(b[c][t / EPL], b[c][LO + t / EPL], t % EPL, act_rev(acc[c][q], lane));
        }
      }
    } else {
      // B element e of K block kb holds virtual unit v = KB kb + EPL g + e
      const int tl = tangent_of(lane);
#pragma unroll
      for (int kb = 0; kb < KBS; ++kb) {
#pragma unroll
        for (int e = 0; e < EPL; ++e) {
          const float4 w = L.w0v[kb * KB + g * EPL + e];
#pragma unroll
          for (int c = 0; c < NC; ++c) {
            const float a = fmaf(w.z, x[c][2], fmaf(w.y, x[c][1], fmaf(w.x, x[c][0], w.w)));
            if (TANGENT && tl != 0) {
              // a_0 from this lane's own point, a_t = w . e_t (radians)
              const float at = tl == 1 ? w.x : (tl == 2 ? w.y : w.z);
              S::set(b[c][kb], e, cos_reduced(a) * at);
            } else {
              S::set(b[c][kb], e, sin_reduced(a));
            }
          }
        }
      }
    }
  }

  // A level with one sine layer: f = w_L . sin(W_0 x + b_0) on the VALU
  // (tangent columns: w_L . (cos(a) w_t), scaled to radians).
  static __device__ __forceinline__ void single(const LevelDev& L, const float (&x)[NC][3], float (&f)[NC], int g) {
    const int tl = tangent_of(threadIdx.x & 63);
#pragma unroll
    for (int c = 0; c < NC; ++c) f[c] = 0.f;
    for (int v0 = 0; v0 < W; v0 += KB) {
#pragma unroll
      for (int e = 0; e < EPL; ++e) {
        const int v = v0 + g * EPL + e;
        const float4 w = L.w0v[v];
        const float wl = L.wlv[v];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          const float a = fmaf(w.z, x[c][2], fmaf(w.y, x[c][1], fmaf(w.x, x[c][0], w.w)));
          if (TANGENT && tl != 0) {
            const float at = tl == 1 ? w.x : (tl == 2 ? w.y : w.z);
            f[c] = fmaf(wl, (MFMA_FIRST ? cos_rev(a) * 6.28318530717958648f : cos_reduced(a)) * at, f[c]);
          } else {
            f[c] = fmaf(wl, MFMA_FIRST ? sin_rev(a) : sin_reduced(a), f[c]);
          }
        }
      }
    }
  }

  // Hidden GEMM j over all row blocks with the epilogue of block rb - 1 issued
  // in slices between the MFMAs of block rb.  epi(rb, q, acc) retires
  // accumulator register q of row block rb.
  template <bool LASTF = false, class Epi>
  static __device__ __forceinline__ void gemm(Stream& A, int j, const float* __restrict__ bc,
                                     const Frag (&b)[NC][BP * KBS], int g, Epi epi) {
    Acc acc[2][NC];
    Acc bias;
    loadc(bc + g * CREGS, bias);
    // tangent columns carry no bias
    const float bmask = (TANGENT && tangent_of(threadIdx.x & 63) != 0) ? 0.f : 1.f;
    if constexpr (TANGENT) {
#pragma unroll
      for (int q = 0; q < CREGS; ++q) bias[q] *= bmask;
    }
#pragma unroll
    for (int rb = 0; rb < RBS; ++rb) {
      Acc bias_next;
#pragma unroll
      for (int kb = 0; kb < KBS; ++kb) {
        const Frag a = A.next(j, (LASTF ? 1 : BP) * (rb * KBS + kb));
        if constexpr (LASTF) {
          // FP16 weights against the activations as an FP16 pair
          const f16x8 af = __builtin_bit_cast(f16x8, a);
#pragma unroll
          for (int c = 0; c < NC; ++c) {
            acc[rb & 1][c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(af, __builtin_bit_cast(f16x8, b[c][kb]),
                                                                    kb == 0 ? bias : acc[rb & 1][c], 0, 0, 0);
            acc[rb & 1][c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(af, __builtin_bit_cast(f16x8, b[c][LO + kb]),
                                                                    acc[rb & 1][c], 0, 0, 0);
          }
        } else if constexpr (SPLIT) {
          const Frag al = A.next(j, BP * (rb * KBS + kb) + 1);
#pragma unroll
          for (int c = 0; c < NC; ++c) {
            acc[rb & 1][c] = S::mma(a, b[c][kb], kb == 0 ? bias : acc[rb & 1][c]);
            acc[rb & 1][c] = S::mma(a, b[c][LO + kb], acc[rb & 1][c]);
            acc[rb & 1][c] = S::mma(al, b[c][kb], acc[rb & 1][c]);
          }
        } else {
#pragma unroll
          for (int c = 0; c < NC; ++c)
            acc[rb & 1][c] = S::mma(a, b[c][kb], kb == 0 ? bias : acc[rb & 1][c]);
        }
        if (kb == 0 && rb + 1 < RBS) loadc(bc + ((rb + 1) * GROUPS + g) * CREGS, bias_next);
        if (rb > 0 && kb % KPQ == 0) {
#pragma unroll
          for (int i = 0; i < QPK; ++i) epi(rb - 1, (kb / KPQ) * QPK + i, acc[(rb - 1) & 1]);
        }
        asm volatile("" ::: "memory");
        // staged stream: the LDS look-ahead reads stay in the step that issues
        // them (the compiler otherwise sinks them to their MFMA and waits on
        // each one; the empty asm does not pin LDS loads)
        if constexpr (STAGED) __builtin_amdgcn_sched_barrier(0);
      }
      if (rb + 1 < RBS) {
        bias = bias_next;
        if constexpr (TANGENT) {
#pragma unroll
          for (int q = 0; q < CREGS; ++q) bias[q] *= bmask;
        }
      }
    }
#pragma unroll
    for (int q = 0; q < CREGS; ++q) epi(RBS - 1, q, acc[(RBS - 1) & 1]);
  }

  // f_level(x) for the NC points of this lane (every lane group returns the
  // same value for its column).  ring: the workgroup's staging ring (staged
  // stream only; then every wave of the workgroup must call it); hand: this
  // wave's handoff region of HANDOFF_FRAGS fragments.
  static __device__ __forceinline__ void eval(const LevelDev& L, const float (&x)[NC][3], float (&f)[NC],
                                     Frag* ring, Frag* hand) {
    eval(L, x, f, ring, hand, load_first(L));
  }
  // The same with the layer-0 operands already loaded (load_first).
  static __device__ __forceinline__ void eval(const LevelDev& L, const float (&x)[NC][3], float (&f)[NC],
                                     Frag* ring, Frag* hand, const First& w0) {
    const int lane = threadIdx.x & 63;
    const int g = lane / ROWS;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const float bl = *L.blp;
    const int n_gemm = L.n_hidden - 1;
    // the output bias reaches the value column only
    const float blv = tangent_of(lane) == 0 ? bl : 0.f;
    if (n_gemm == 0) {
      single(L, x, f, g);
#pragma unroll
      for (int c = 0; c < NC; ++c) f[c] = group_sum(f[c]) + blv;
      return;
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) f[c] = 0.f;
    Frag b[NC][BP * KBS];
    Stream A{reinterpret_cast<const Frag*>(L.apk), ring, lane, wave, {}};
    A.start();
    first(L, x, b, lane, g, w0);
    A.begin();

    // hidden layers that feed another hidden layer: B registers -> LDS -> B registers
    if constexpr (LF) {
      // the one hidden-to-hidden product (the stack gives LAST_F16 only to
      // levels with n_gemm == 2) hands over FP16 pairs
      Frag st[NC][BP];
      gemm(A, 0, L.biasc, b, g, [&](int rb, int q, const Acc (&acc)[NC]) {
        const int t = rb * CREGS + q;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          set2h(st[c][0], st[c][1], t % EPL, act_rev(acc[c][q], lane));
          if (t % EPL == EPL - 1) {
            hand[(c * KBS + t / EPL) * 64 + lane] = st[c][0];
            hand[((NC + c) * KBS + t / EPL) * 64 + lane] = st[c][1];
          }
        }
      });
#pragma unroll
      for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int kb = 0; kb < KBS; ++kb) {
          b[c][kb] = hand[(c * KBS + kb) * 64 + lane];
          b[c][LO + kb] = hand[((NC + c) * KBS + kb) * 64 + lane];
        }
    } else {
      for (int j = 0; j + 1 < n_gemm; ++j) {
        Frag st[NC][BP];
        gemm(A, j, L.biasc + (size_t)j * W, b, g, [&](int rb, int q, const Acc (&acc)[NC]) {
          const int t = rb * CREGS + q;
#pragma unroll
          for (int c = 0; c < NC; ++c) {
            set2(st[c][0], st[c][BP - 1], t % EPL, act_rev(acc[c][q], lane));
            if (t % EPL == EPL - 1) {
              hand[(c * KBS + t / EPL) * 64 + lane] = st[c][0];
              if constexpr (SPLIT) hand[((NC + c) * KBS + t / EPL) * 64 + lane] = st[c][1];
            }
          }
        });
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
          for (int kb = 0; kb < KBS; ++kb) {
            b[c][kb] = hand[(c * KBS + kb) * 64 + lane];
            if constexpr (SPLIT) b[c][LO + kb] = hand[((NC + c) * KBS + kb) * 64 + lane];
          }
      }
    }

    // last hidden layer fused with the output layer (output weights of row
    // block rb + 1 are fetched while rb retires)
    Acc wl, wl_next;
    loadc(L.wlc + g * CREGS, wl);
    gemm<LF>(A, n_gemm - 1, L.biasc + (size_t)(n_gemm - 1) * W, b, g,
         [&](int rb, int q, const Acc (&acc)[NC]) {
           if (q == 0 && rb + 1 < RBS) loadc(L.wlc + ((rb + 1) * GROUPS + g) * CREGS, wl_next);
#pragma unroll
           for (int c = 0; c < NC; ++c) f[c] = fmaf(wl[q], act_rev(acc[c][q], lane), f[c]);
           if (q == CREGS - 1 && rb + 1 < RBS) wl = wl_next;
         });
    A.finish();
#pragma unroll
    for (int c = 0; c < NC; ++c) f[c] = group_sum(f[c]) + blv;
  }
};

// Default column blocks per wave at width W: the B operand set then takes 128
// VGPRs (W / 4 per 32-point block for 16-bit shapes, per 16-point block for
// FP32).
template <class S>
constexpr int default_cols(int W) {
  return S::ROWS == 32 ? (W == 64 ? 4 : W == 512 ? 1 : 2) : (W == 64 ? 8 : W == 128 ? 4 : W == 256 ? 2 : 1);
}

// Block `block` (WAVES waves x TILE points) of one level launch: evaluate,
// store (accumulate), append in-band keys.  n: points of the launch (device
// count for index lists), with block * WAVES * TILE < n.  Every wave of the
// workgroup calls it (the staged stream has barriers); a wave whose tile lies
// past n runs on clamped points and stores nothing.
template <class S, int W, int NC, int WAVES = 4>
__device__ __forceinline__ void eval_block(const LevelDev& L, const EvalTarget& t, int64_t n, int64_t block,
                                  typename S::Frag* lds) {
  typedef LevelCore<S, W, NC, true, WAVES> M;
  const int lane = threadIdx.x & 63;
  const int64_t base = (block * WAVES + (threadIdx.x >> 6)) * M::TILE;
  // layer-0 operands first: their latency overlaps the coordinates below
  const typename M::First w0 = M::load_first(L);
  __builtin_amdgcn_sched_barrier(0);

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
  M::eval(L, x, f, lds, lds + M::RING_FRAGS + (threadIdx.x >> 6) * M::HANDOFF_FRAGS, w0);

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

}  // namespace mp
