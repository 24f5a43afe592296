// One-level evaluation kernel for the 16-bit (BF16 / FP16) hidden layers.
//
// Same contract as k_level (explicit points, grid ranges, compacted index
// lists; band compaction fused into the store), with a layout built around the
// limits of one wave per SIMD:
//  - layer 0 (3 -> width) runs on the FP32 matrix cores: two 32x32x2 MFMAs per
//    32-row block over the operand (x, y, z, 1); its accumulator registers come
//    out in the row order the chained hidden layers use, so the sines go
//    straight into layer 1's B operand;
//  - the activations of a layer that feeds another hidden layer are handed over
//    through the wave's own LDS region instead of a second register buffer, so
//    the VGPRs hold one B operand set and two accumulator sets;
//  - the A fragments (weights) of all hidden layers form one continuous stream.
//    For widths >= 128 the four waves of a workgroup copy it, 16 KB stage by
//    16 KB stage, into a two-slot LDS ring with direct global->LDS loads issued
//    a whole stage ahead (one barrier per stage), and each wave reads its
//    fragments from LDS a few steps ahead; width 64 streams from global memory
//    through a register ring;
//  - the epilogue of row block rb (bias is the MFMA C operand; sine, packing,
//    LDS store or output dot product) is issued in slices between the MFMAs of
//    row block rb + 1, so the transcendental unit and the matrix core overlap.
#include "eval_dispatch.h"

namespace mp {

typedef __attribute__((address_space(3))) void lds_void;

template <class S, int W, int NC>
struct Lvl16 {
  static_assert(S::ROWS == 32 && S::KB == 16 && S::CREGS == 16, "32x32x16 shapes only");
  typedef typename S::Frag Frag;
  typedef typename S::Acc Acc;
  static constexpr int RBS = W / 32;            // 32-row blocks per layer
  static constexpr int KBS = W / 16;            // 16-deep K blocks per layer
  static constexpr int STEPS = RBS * KBS;       // A fragments per layer
  static constexpr int TILE = NC * 32;          // points per wave
  static constexpr int HANDOFF_FRAGS = NC * KBS * 64;  // per wave, in fragments

  // A stream.  Staged: stages of SF fragments, two LDS slots, read D steps
  // ahead.  The slot of a step must not depend on the (run-time) layer index,
  // so a layer spans an even number of stages.
  static constexpr bool STAGED = STEPS % 32 == 0;
  static constexpr int SF = 16;
  static constexpr int D = 4;
  static_assert(!STAGED || STEPS % D == 0, "the read-ahead ring must tile a layer");
  static constexpr int RING_FRAGS = STAGED ? 2 * SF * 64 : 0;
  static constexpr int PF = STAGED ? D : (STEPS < 8 ? STEPS : 8);
  static_assert(STAGED || STEPS % PF == 0, "the prefetch ring must tile a layer");
  static_assert((size_t)(STAGED ? 2 * SF : PF) * 64 * sizeof(Frag) <= APK_PAD,
                "the stream reads past the padding");
  static constexpr size_t LDS_BYTES = (size_t)(RING_FRAGS + 4 * HANDOFF_FRAGS) * sizeof(Frag);

  // accumulator registers retired per K step by the pipelined epilogue
  static constexpr int QPK = KBS >= 16 ? 1 : 16 / KBS;
  static constexpr int KPQ = KBS >= 16 ? KBS / 16 : 1;

  struct Stream {
    // A is deliberately not __restrict__: the empty asm with a memory clobber
    // after every step pins each look-ahead load to its step (through a
    // restrict pointer the loads may sink to their uses and expose their
    // latency).
    const Frag* A;     // hidden-layer fragments, layer-major, APK_PAD readable bytes after the end
    Frag* ring;        // LDS ring (staged)
    int lane, wave;
    Frag pf[PF];

    // Stage `stage` of the stream into ring slot `slot`: each wave copies 4 of
    // the stage's 16 fragments (1 KB each, one 16-byte direct load per lane).
    __device__ inline void copy_stage(int64_t stage, int slot) {
#pragma unroll
      for (int i = 0; i < SF / 4; ++i) {
        const int f = (SF / 4) * wave + i;
        __builtin_amdgcn_global_load_lds((const void*)(A + (stage * SF + f) * 64 + lane),
                                         (lds_void*)(ring + (slot * SF + f) * 64), 16, 0, 0);
      }
    }
    // Make the copies issued so far visible to the whole workgroup.
    __device__ inline void barrier() {
      asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
      __syncthreads();
    }

    // Before layer 0: start the first two stages (or fill the register ring).
    __device__ inline void start() {
      if constexpr (STAGED) {
        copy_stage(0, 0);
        copy_stage(1, 1);
      } else {
#pragma unroll
        for (int d = 0; d < PF; ++d) pf[d] = A[d * 64 + lane];
      }
    }
    // After layer 0: stage 0 visible, first D fragments in registers.
    __device__ inline void begin() {
      if constexpr (STAGED) {
        barrier();
#pragma unroll
        for (int d = 0; d < D; ++d) pf[d] = ring[d * 64 + lane];
      }
    }
    // Fragment of step s of hidden layer j; issues the look-ahead.
    __device__ inline Frag next(int j, int s) {
      if constexpr (STAGED) {
        const int t = s + D;  // step whose fragment is read now
        if (t % SF == 0) {
          // every wave is past its last read of the slot about to be refilled
          barrier();
          copy_stage((int64_t)j * (STEPS / SF) + t / SF + 1, (t / SF + 1) % 2);
        }
        const Frag a = pf[s % D];
        pf[s % D] = ring[((t / SF) % 2 * SF + t % SF) * 64 + lane];
        return a;
      } else {
        const Frag a = pf[s % PF];
        pf[s % PF] = A[((size_t)j * STEPS + s + PF) * 64 + lane];
        return a;
      }
    }
  };

  // Element q of a 32x32 accumulator, as an operand index of the next layer.
  static __device__ inline void put(Frag (&b)[NC][KBS], int c, int rb, int q, float v) {
    S::set(b[c][2 * rb + q / 8], q % 8, v);
  }

  static __device__ inline void load16(const float* __restrict__ p, Acc& v) {
    const float4* p4 = reinterpret_cast<const float4*>(p);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float4 t = p4[i];
      v[4 * i] = t.x; v[4 * i + 1] = t.y; v[4 * i + 2] = t.z; v[4 * i + 3] = t.w;
    }
  }

  // Layer 0 into the B registers of layer 1 (or, with `out`, straight into the
  // output dot product when the level has one sine layer).
  static __device__ inline void first(const LevelDev& L, const float (&x)[NC][3], Frag (&b)[NC][KBS],
                                      int lane, int g, const float* __restrict__ wlc, float (&f)[NC],
                                      bool out) {
    float bxy[NC], bz1[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      bxy[c] = g == 0 ? x[c][0] : x[c][1];
      bz1[c] = g == 0 ? x[c][2] : 1.f;
    }
    Acc zero;
#pragma unroll
    for (int q = 0; q < 16; ++q) zero[q] = 0.f;
#pragma unroll
    for (int rb = 0; rb < RBS; ++rb) {
      const float2 a = L.w0m[rb * 64 + lane];
      Acc acc[NC];
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        acc[c] = __builtin_amdgcn_mfma_f32_32x32x2f32(a.x, bxy[c], zero, 0, 0, 0);
        acc[c] = __builtin_amdgcn_mfma_f32_32x32x2f32(a.y, bz1[c], acc[c], 0, 0, 0);
      }
      if (out) {
        Acc wl;
        load16(wlc + (rb * 2 + g) * 16, wl);
#pragma unroll
        for (int q = 0; q < 16; ++q)
#pragma unroll
          for (int c = 0; c < NC; ++c) f[c] = fmaf(wl[q], sin_rev(acc[c][q]), f[c]);
      } else {
#pragma unroll
        for (int q = 0; q < 16; ++q)
#pragma unroll
          for (int c = 0; c < NC; ++c) put(b, c, rb, q, sin_rev(acc[c][q]));
      }
    }
  }

  // Hidden GEMM j over all row blocks with the epilogue of block rb - 1 issued
  // in slices between the MFMAs of block rb.  epi(rb, q, acc) retires
  // accumulator register q of row block rb.
  template <class Epi>
  static __device__ inline void gemm(Stream& A, int j, const float* __restrict__ bc,
                                     const Frag (&b)[NC][KBS], int g, Epi epi) {
    Acc acc[2][NC];
    Acc bias;
    load16(bc + g * 16, bias);
#pragma unroll
    for (int rb = 0; rb < RBS; ++rb) {
      Acc bias_next;
#pragma unroll
      for (int kb = 0; kb < KBS; ++kb) {
        const Frag a = A.next(j, rb * KBS + kb);
#pragma unroll
        for (int c = 0; c < NC; ++c)
          acc[rb & 1][c] = S::mma(a, b[c][kb], kb == 0 ? bias : acc[rb & 1][c]);
        if (kb == 0 && rb + 1 < RBS) load16(bc + ((rb + 1) * 2 + g) * 16, bias_next);
        if (rb > 0 && kb % KPQ == 0) {
#pragma unroll
          for (int i = 0; i < QPK; ++i) epi(rb - 1, (kb / KPQ) * QPK + i, acc[(rb - 1) & 1]);
        }
        asm volatile("" ::: "memory");
      }
      if (rb + 1 < RBS) bias = bias_next;
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) epi(RBS - 1, q, acc[(RBS - 1) & 1]);
  }

  // f_level(x) for the NC points of this lane (both lane groups return the
  // same value for their column).  lds: the workgroup's LDS (ring first, then
  // one handoff region per wave).  Every wave of the workgroup must call it.
  static __device__ inline void eval(const LevelDev& L, const float (&x)[NC][3], float (&f)[NC],
                                     Frag* lds) {
    const int lane = threadIdx.x & 63;
    const int g = lane / 32;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const float bl = *L.blp;
#pragma unroll
    for (int c = 0; c < NC; ++c) f[c] = 0.f;
    Frag b[NC][KBS];
    const int n_gemm = L.n_hidden - 1;
    if (n_gemm == 0) {
      first(L, x, b, lane, g, L.wlc, f, true);
#pragma unroll
      for (int c = 0; c < NC; ++c) f[c] = group_sum(f[c]) + bl;
      return;
    }
    Frag* hand = lds + RING_FRAGS + wave * HANDOFF_FRAGS;
    Stream A{reinterpret_cast<const Frag*>(L.apk), lds, lane, wave, {}};
    A.start();
    first(L, x, b, lane, g, nullptr, f, false);
    A.begin();

    // hidden layers that feed another hidden layer: B registers -> LDS -> B registers
    for (int j = 0; j + 1 < n_gemm; ++j) {
      Frag st[NC];
      gemm(A, j, L.biasc + (size_t)j * W, b, g, [&](int rb, int q, const Acc (&acc)[NC]) {
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          S::set(st[c], q % 8, sin_rev(acc[c][q]));
          if (q % 8 == 7) hand[(c * KBS + 2 * rb + q / 8) * 64 + lane] = st[c];
        }
      });
#pragma unroll
      for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int kb = 0; kb < KBS; ++kb) b[c][kb] = hand[(c * KBS + kb) * 64 + lane];
    }

    // last hidden layer fused with the output layer (output weights of row
    // block rb + 1 are fetched while rb retires)
    Acc wl, wl_next;
    load16(L.wlc + g * 16, wl);
    gemm(A, n_gemm - 1, L.biasc + (size_t)(n_gemm - 1) * W, b, g,
         [&](int rb, int q, const Acc (&acc)[NC]) {
           if (q == 0 && rb + 1 < RBS) load16(L.wlc + ((rb + 1) * 2 + g) * 16, wl_next);
#pragma unroll
           for (int c = 0; c < NC; ++c) f[c] = fmaf(wl[q], sin_rev(acc[c][q]), f[c]);
           if (q == 15 && rb + 1 < RBS) wl = wl_next;
         });
    // the look-ahead copies past the end land in the ring; let them finish
    // before the workgroup can retire and its LDS be reassigned
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
#pragma unroll
    for (int c = 0; c < NC; ++c) f[c] = group_sum(f[c]) + bl;
  }

  static __device__ inline float group_sum(float v) { return v + __shfl_xor(v, 32); }
};

template <class S, int W, int NC>
__global__ __launch_bounds__(256) void k_level16(LevelDev L, EvalTarget t) {
  typedef Lvl16<S, W, NC> M;
  typedef typename S::Frag Frag;
  extern __shared__ uint4 lds_raw[];
  const int lane = threadIdx.x & 63;
  const int64_t n = (t.src == SRC_LIST) ? (int64_t)*t.list_count : t.n;
  // Workgroup-uniform exit: the waves of a workgroup share barriers, so a wave
  // whose tile lies past n still runs (on clamped points) and stores nothing.
  if ((int64_t)blockIdx.x * 4 * M::TILE >= n) return;
  const int64_t base = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * M::TILE;

  float x[NC][3];
  int64_t key[NC];
  bool valid[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const int64_t i = base + c * 32 + (lane % 32);
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
  M::eval(L, x, f, reinterpret_cast<Frag*>(lds_raw));

  // Lanes of group 0 own the store (both groups hold the same value).
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    bool flag = false;
    if (lane < 32 && valid[c]) {
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
static int launch16_typed(const LevelDev& L, const EvalTarget& t, hipStream_t stream) {
  typedef Lvl16<S, W, NC> M;
  const int64_t per_block = (int64_t)M::TILE * 4;
  const int64_t blocks = (t.n + per_block - 1) / per_block;
  if (blocks <= 0) return MP_OK;
  MP_REQUIRE(blocks < (1ll << 31), MP_EINVAL, "level launch: %lld points is too many", (long long)t.n);
  static_assert(M::LDS_BYTES <= 160 * 1024, "LDS budget of one workgroup");
  static bool attr_set = false;
  if (!attr_set) {
    MP_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_level16<S, W, NC>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)M::LDS_BYTES));
    attr_set = true;
  }
  hipLaunchKernelGGL((k_level16<S, W, NC>), dim3((unsigned)blocks), dim3(256), M::LDS_BYTES, stream, L, t);
  MP_LAUNCHED("k_level16");
  return MP_OK;
}

// Column blocks per wave: the B operand takes NC * W / 4 VGPRs (128 at the
// defaults), or one block when the stack asks for narrow tiles.
template <class S>
static int launch16_width(const LevelDev& L, int nc, const EvalTarget& t, hipStream_t stream) {
  if (nc == 1) {
    switch (L.width) {
      case 64: return launch16_typed<S, 64, 1>(L, t, stream);
      case 128: return launch16_typed<S, 128, 1>(L, t, stream);
      case 256: return launch16_typed<S, 256, 1>(L, t, stream);
      case 512: return launch16_typed<S, 512, 1>(L, t, stream);
    }
  } else {
    switch (L.width) {
      case 64: return launch16_typed<S, 64, 4>(L, t, stream);
      case 128: return launch16_typed<S, 128, 2>(L, t, stream);
      case 256: return launch16_typed<S, 256, 2>(L, t, stream);
      case 512: return launch16_typed<S, 512, 1>(L, t, stream);
    }
  }
  set_error("level launch: unsupported padded width %d", L.width);
  return MP_EUNSUPPORTED;
}

int launch_level16(const mp_stack* st, int level, const EvalTarget& t, hipStream_t stream) {
  const LevelDev& L = st->lv[level].dev;
  switch (st->shape) {
    case SH_32x32x16_BF16: return launch16_width<S32x32x16_bf16>(L, st->nc, t, stream);
    case SH_32x32x16_F16: return launch16_width<S32x32x16_f16>(L, st->nc, t, stream);
  }
  set_error("level launch: shape %d is not a 16-bit 32x32 shape", st->shape);
  return MP_EUNSUPPORTED;
}

}  // namespace mp
