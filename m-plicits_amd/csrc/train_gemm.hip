// Matrix-core GEMMs of the residual-level training step with the elementwise
// stages fused into their epilogues (the reference runs these products as
// autograd GEMMs: the forward, the create_graph input gradient and the double
// backward of experiment_scripts/train_sdf_on_neighborhood_fine.py:220-252).
//
//   C [M][N] = sum_k A(m, k) B(n, k)
//
// Both operands are read in 16-byte chunks (eight 16-bit or four FP32
// elements along k) from the blocked layout [K/KL][rows][KL]: element (r, k)
// at ((k / KL) * rows + r) * KL + k % KL.  For the layer products k runs over
// the units of the layer below (A is an activation block written
// "unit-blocked" by the previous epilogue, B the weights); for the weight
// gradients (TG_PART) k runs over the activation rows, and the operands are
// the "row-blocked" copies the epilogues write beside the unit-blocked ones.
//
// Rows 4p..4p+3 hold the value and the three coordinate tangents of point p.
// In the MFMA C layout a lane holds four consecutive rows of one column
// (32x32 blocks: rows 8i + 4g + j; 16x16 blocks: rows 4g + j), i.e. the value
// and the tangents of one point and unit, so every activation and its
// backward is lane-local.  The unit-blocked outputs need a unit run per row
// instead: the four lanes of a quad (units 4w..4w+3) exchange their values
// with two DPP quad permutations, after which lane t of the quad holds row t
// of its point for the four units (an 8-byte or 16-byte store, 128 contiguous
// bytes per 16-bit chunk over the two lane groups).  The row-blocked copy and
// the activation's derivative factors D (point-major [points][N][4] in the
// operand type: cos a and sin(a) ad_t, all the backward needs of the layer)
// are stored straight from the C layout.
//
// Workgroup: four waves in 2 x 2, each a 4 x 4 grid of MFMA blocks (16-bit:
// 32x32x16, 128 x 128 per wave; FP32: 16x16x4, 64 x 64 per wave).  Per K step
// of 16 the workgroup's operand tiles (16 KB) are copied once into an LDS
// ring by direct 16-byte global-to-LDS loads (several steps in flight), and
// every wave reads its fragments from there, one step ahead of its MFMAs:
// half the L2 traffic of per-wave register loads, and no registers held by
// the loads in flight.  FP32 blocks take a chunk of four consecutive k per
// lane group and run four MFMAs on it: instruction e reduces k = 16 s + 4 g + e
// over the lane groups g (the reduction order is free when both operands
// follow it).  The workgroups that read the same operand rows (the column
// tiles of one row tile; the tiles of one split-K chunk) are numbered to run
// on one XCD, so those rows are fetched from HBM into its L2 once.
// No bounds checks: the host pads M, N, K to whole tiles (widths to 256,
// points to 128).
#include <type_traits>

#include "common.h"
#include "mlp_core.h"
#include "train_gemm.h"

namespace mp {

template <class T> struct TgTraits;
template <> struct TgTraits<__bf16> {
  typedef bf16x8 Chunk;
  typedef S32x32x16_bf16 S;
};
template <> struct TgTraits<_Float16> {
  typedef f16x8 Chunk;
  typedef S32x32x16_f16 S;
};
template <> struct TgTraits<float> {
  typedef f32x4 Chunk;
  typedef S16x16x4_f32 S;
};

constexpr int TG_RING = 4;  // K steps per chunk of a split-K range (multiples keep the chunks whole)
constexpr int TG_STAGE_BYTES = 16384;
// LDS ring depth: 8 stages (128 KB, one workgroup per CU, as its registers
// allow anyway) for 16-bit operands; 4 (64 KB) in FP32, whose smaller
// register footprint lets two workgroups share a CU
template <class T>
constexpr int tg_stages() { return sizeof(T) == 4 ? 4 : 8; }
constexpr int TG_BR = 4, TG_BC = 4;

typedef __attribute__((address_space(3))) void lds_void;

// TgArgs plus the launch geometry
struct TgLaunch {
  TgArgs a;
  int64_t ksteps;           // K / 16
  int64_t mtiles;           // M / tile rows
  int ntn;                  // N / tile columns
  int64_t steps_per_split;  // TG_PART
};

template <class T>
__device__ inline void st4rows(T* p, float a, float b, float c, float d) {
  typedef T v4 __attribute__((ext_vector_type(4)));
  v4 o;
  o[0] = (T)a; o[1] = (T)b; o[2] = (T)c; o[3] = (T)d;
  *reinterpret_cast<v4*>(p) = o;
}

// DPP quad permutations: lane t of each quad of four lanes reads lane t ^ 2 / t ^ 1
__device__ inline uint32_t quad_xor2(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
}
__device__ inline uint32_t quad_xor1(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
}
__device__ inline float quad_xor2(float v) { return __builtin_bit_cast(float, quad_xor2(__builtin_bit_cast(uint32_t, v))); }
__device__ inline float quad_xor1(float v) { return __builtin_bit_cast(float, quad_xor1(__builtin_bit_cast(uint32_t, v))); }

template <class T>
__device__ inline uint32_t pack2(float lo, float hi) {
  return (uint32_t)__builtin_bit_cast(uint16_t, (T)lo) | ((uint32_t)__builtin_bit_cast(uint16_t, (T)hi) << 16);
}

// 4 x 4 transpose over a quad: lane t holds m[t][0..3] (its unit, rows 0..3),
// packed two 16-bit values per dword (d0 = rows 0, 1; d1 = rows 2, 3); returns
// row t of units 0..3 (x = units 0, 1; y = units 2, 3).  First the off-diagonal
// 2 x 2 blocks swap between lanes t and t ^ 2, then the elements within them
// between lanes t and t ^ 1.
// The half-word moves are byte permutes (v_perm_b32: selector bytes 0-3 pick
// the second operand's bytes, 4-7 the first's) whose selectors depend on the
// lane's place only, so the lanes of a quad take no branches.
__device__ inline uint2 quad_transpose16(uint32_t d0, uint32_t d1, int t) {
  const bool hi = (t & 2) != 0, odd = (t & 1) != 0;
  const uint32_t r = quad_xor2(hi ? d0 : d1);
  const uint32_t e0 = hi ? r : d0, e1 = hi ? d1 : r;
  // odd lanes send the low halves of e0, e1, even lanes the high halves
  const uint32_t q = quad_xor1(__builtin_amdgcn_perm(e1, e0, odd ? 0x05040100u : 0x07060302u));
  // odd: x = (q.lo, e0.hi), y = (q.hi, e1.hi); even: x = (e0.lo, q.lo), y = (e1.lo, q.hi)
  return make_uint2(__builtin_amdgcn_perm(e0, q, odd ? 0x07060100u : 0x01000504u),
                    __builtin_amdgcn_perm(e1, q, odd ? 0x07060302u : 0x03020504u));
}

// the same for FP32 values
__device__ inline float4 quad_transpose32(float m0, float m1, float m2, float m3, int t) {
  const bool hi = (t & 2) != 0, odd = (t & 1) != 0;
  const float r0 = quad_xor2(hi ? m0 : m2), r1 = quad_xor2(hi ? m1 : m3);
  const float a0 = hi ? r0 : m0, a1 = hi ? r1 : m1, a2 = hi ? m2 : r0, a3 = hi ? m3 : r1;
  const float q0 = quad_xor1(odd ? a0 : a1), q1 = quad_xor1(odd ? a2 : a3);
  return odd ? make_float4(q0, a1, q1, a3) : make_float4(a0, q0, a2, q1);
}

// Rows mv..mv+3 (one point) of unit n = the lane's column into the
// unit-blocked X [N/KL][M][KL]; t = n % 4 (the lane's place in its quad).
// Offsets are unsigned (rows and units are never negative).
template <class T>
__device__ inline void st_unit_blocked(T* X, int64_t M, int n, int64_t mv, int t, float v0, float v1, float v2,
                                       float v3) {
  const uint32_t nq = (uint32_t)(n - t);
  if constexpr (sizeof(T) == 4) {
    *reinterpret_cast<float4*>(X + ((uint64_t)(nq >> 2) * (uint64_t)M + (uint64_t)mv + t) * 4) =
        quad_transpose32(v0, v1, v2, v3, t);
  } else {
    *reinterpret_cast<uint2*>(X + ((uint64_t)(nq >> 3) * (uint64_t)M + (uint64_t)mv + t) * 8 + (nq & 7u)) =
        quad_transpose16(pack2<T>(v0, v1), pack2<T>(v2, v3), t);
  }
}

// The same from the four values already packed two per dword (16-bit T).
template <class T>
__device__ inline void st_unit_blocked16(T* X, int64_t M, int n, int64_t mv, int t, uint2 packed) {
  const uint32_t nq = (uint32_t)(n - t);
  *reinterpret_cast<uint2*>(X + ((uint64_t)(nq >> 3) * (uint64_t)M + (uint64_t)mv + t) * 8 + (nq & 7u)) =
      quad_transpose16(packed.x, packed.y, t);
}

// Rows mv..mv+3 of unit n into the row-blocked Xb [M/KL][N][KL].
template <class T>
__device__ inline void st_row_blocked(T* Xb, int N, int n, int64_t mv, float v0, float v1, float v2, float v3) {
  constexpr int KL = 16 / (int)sizeof(T);
  const uint64_t m = (uint64_t)mv;
  st4rows(Xb + ((m / KL) * (uint32_t)N + (uint32_t)n) * KL + m % KL, v0, v1, v2, v3);
}

// The same from the packed values (16-bit T: rows mv..mv+3 are 8 contiguous bytes).
template <class T>
__device__ inline void st_row_blocked16(T* Xb, int N, int n, int64_t mv, uint2 packed) {
  constexpr int KL = 16 / (int)sizeof(T);
  const uint64_t m = (uint64_t)mv;
  *reinterpret_cast<uint2*>(Xb + ((m / KL) * (uint32_t)N + (uint32_t)n) * KL + m % KL) = packed;
}

// Both layouts of one point's four rows, each value converted once.
template <class T>
__device__ inline void st_both(T* X, T* Xb, int64_t M, int N, int n, int64_t mv, int t, float v0, float v1, float v2,
                               float v3) {
  if constexpr (sizeof(T) == 4) {
    st_unit_blocked(X, M, n, mv, t, v0, v1, v2, v3);
    if (Xb) st_row_blocked(Xb, N, n, mv, v0, v1, v2, v3);
  } else {
    const uint2 packed = make_uint2(pack2<T>(v0, v1), pack2<T>(v2, v3));
    st_unit_blocked16(X, M, n, mv, t, packed);
    if (Xb) st_row_blocked16(Xb, N, n, mv, packed);
  }
}

// Sum of V values over the lanes of a lane group, exchanging with the lanes
// at xor distances M, M/2, ..., 1 (V <= 2M, powers of two): every exchange
// halves the values a lane keeps, so lane c ends with the sum of value
// (c / (2M / V)) % V, held by 2M / V neighbouring lanes.
template <int V, int M>
__device__ inline void lane_reduce(float* v, int col) {
  if constexpr (M >= 1) {
    if constexpr (V > 1) {
      const bool up = (col & M) != 0;
#pragma unroll
      for (int k = 0; k < V / 2; ++k) {
        const float send = up ? v[k] : v[k + V / 2];
        const float keep = up ? v[k + V / 2] : v[k];
        v[k] = keep + __shfl_xor(send, M);
      }
      lane_reduce<V / 2, M / 2>(v, col);
    } else {
      v[0] += __shfl_xor(v[0], M);
      lane_reduce<1, M / 2>(v, col);
    }
  }
}

template <class T, int EPI>
__global__ __launch_bounds__(256) void k_tgemm(TgLaunch L) {
  typedef TgTraits<T> TR;
  typedef typename TR::S S;
  typedef typename TR::Chunk Chunk;
  typedef typename S::Acc Acc;
  constexpr int KL = 16 / (int)sizeof(T);  // k per chunk
  constexpr int CPS = 16 / KL;             // chunks per K step along k (= lane groups)
  constexpr int R = S::ROWS;
  constexpr int WM = TG_BR * R, WN = TG_BC * R;
  constexpr int PTS = S::CREGS / 4;  // points per lane per block
  const TgArgs& a = L.a;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int g = lane / R, col = lane % R;
  // workgroup -> (group, index): the workgroups of a group share an XCD
  const int per_grp = EPI == TG_PART ? (int)(L.mtiles * L.ntn) : L.ntn;
  const int xcd = blockIdx.x % 8, slot = blockIdx.x / 8;
  const int64_t grp = xcd + 8 * (int64_t)(slot / per_grp);
  const int idx = slot % per_grp;
  int64_t mt, nt, chunk;
  if constexpr (EPI == TG_PART) {
    chunk = grp;
    if (chunk >= a.splits) return;  // workgroup-uniform
    mt = idx / L.ntn;
    nt = idx % L.ntn;
  } else {
    chunk = 0;
    mt = grp;
    if (mt >= L.mtiles) return;  // workgroup-uniform
    nt = idx;
  }
  const int64_t m0 = mt * (2 * WM) + (wave >> 1) * WM;
  const int n0 = (int)nt * (2 * WN) + (wave & 1) * WN;
  const int64_t s0 = chunk * L.steps_per_split;
  const int64_t s1 = s0 + L.steps_per_split < L.ksteps ? s0 + L.steps_per_split : L.ksteps;

  Acc acc[TG_BR][TG_BC];
#pragma unroll
  for (int i = 0; i < TG_BR; ++i)
#pragma unroll
    for (int j = 0; j < TG_BC; ++j)
#pragma unroll
      for (int q = 0; q < S::CREGS; ++q) acc[i][j][q] = 0.f;

  // LDS ring of SG K steps; a stage holds the workgroup's A tile (CPS planes
  // of TM chunks) and B tile (CPS planes of TN chunks), 16 KB, in the order of
  // the blocked layout, so both the copy and the fragment reads are
  // contiguous runs of 16-byte chunks.
  constexpr int SG = tg_stages<T>();
  __shared__ __attribute__((aligned(16))) char ring[SG * TG_STAGE_BYTES];
  if (s0 < s1) {
    constexpr int TM = 2 * WM, TN = 2 * WN;
    constexpr int PP = TM / 64;  // 1 KB pieces per plane (TM == TN)
    static_assert(TM == TN && CPS * TM * 16 == TG_STAGE_BYTES / 2, "stage = A tile + B tile");
    const int64_t lda = a.M, ldb = a.N;
    const int64_t mt0 = mt * TM, nt0 = nt * TN;
    const Chunk* A = reinterpret_cast<const Chunk*>(a.A);
    const Chunk* B = reinterpret_cast<const Chunk*>(a.B);
    // Stage `slot` <- K step ks: 16 pieces of 1 KB (one 16-byte direct load
    // per lane each), pieces 0..7 the A tile, 8..15 the B tile; wave w copies
    // pieces 4w..4w+3.
    auto copy = [&](int64_t ks, int slot) {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int piece = 4 * wave + k;
        const int q = piece & 7, plane = q / PP, r = (q % PP) * 64 + lane;
        const Chunk* src = piece < 8 ? A + (ks * CPS + plane) * lda + mt0 + r : B + (ks * CPS + plane) * ldb + nt0 + r;
        __builtin_amdgcn_global_load_lds((const void*)src, (lds_void*)(ring + slot * TG_STAGE_BYTES + piece * 1024), 16,
                                         0, 0);
      }
    };
    const int64_t nsteps = s1 - s0;
    auto clamp = [&](int64_t t) { return s0 + (t < nsteps ? t : nsteps - 1); };
    // this wave's fragment offsets within a stage (in chunks)
    const int fa = g * TM + (wave >> 1) * WM + col;
    const int fb = TM * CPS + g * TN + (wave & 1) * WN + col;
    Chunk ra[2][TG_BR], rb[2][TG_BC];
    auto frags = [&](int64_t t, int k) {  // k: a constant after inlining
      const Chunk* st = reinterpret_cast<const Chunk*>(ring + (t % SG) * TG_STAGE_BYTES);
#pragma unroll
      for (int i = 0; i < TG_BR; ++i) ra[k][i] = st[fa + i * R];
#pragma unroll
      for (int j = 0; j < TG_BC; ++j) rb[k][j] = st[fb + j * R];
    };
#pragma unroll
    for (int d = 0; d < SG - 1; ++d) copy(clamp(d), d);
    // step 0 in: this wave's copies of steps 1..SG-2 may still fly
    asm volatile("s_waitcnt vmcnt(%0)\n\ts_barrier" ::"n"(4 * (SG - 2)) : "memory");
    frags(0, 0);
    // one K step on fragment buffer K (compile-time, so the buffers stay in registers)
    auto step = [&](int64_t t, auto kc) {
      constexpr int K = decltype(kc)::value;
      // Step t + 1 in (this wave's copies of the SG - 3 steps after it may
      // still fly; after the barrier every wave's are), and every wave is
      // done with stage (t - 1) % SG, which the next copy overwrites.  Its
      // fragments are read while the MFMAs of step t run.
      asm volatile("s_waitcnt vmcnt(%0)\n\ts_waitcnt lgkmcnt(0)\n\ts_barrier" ::"n"(4 * (SG - 3)) : "memory");
      copy(clamp(t + SG - 1), (int)((t + SG - 1) % SG));
      frags(t + 1 < nsteps ? t + 1 : t, K ^ 1);
#pragma unroll
      for (int i = 0; i < TG_BR; ++i)
#pragma unroll
        for (int j = 0; j < TG_BC; ++j) {
          if constexpr (sizeof(T) == 4) {
#pragma unroll
            for (int e = 0; e < 4; ++e) acc[i][j] = S::mma(ra[K][i][e], rb[K][j][e], acc[i][j]);
          } else {
            acc[i][j] = S::mma(ra[K][i], rb[K][j], acc[i][j]);
          }
        }
    };
    // nsteps is even (a multiple of the split-K granule)
    for (int64_t t = 0; t < nsteps; t += 2) {
      step(t, std::integral_constant<int, 0>());
      step(t + 1, std::integral_constant<int, 1>());
    }
    // the tail's clamped copies land before the workgroup retires
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  }

  // ---- epilogues ----
  const int N = a.N;
  const int64_t M = a.M;
  if constexpr (EPI == TG_PART) {
    float* out = a.part + (chunk * M) * N;
#pragma unroll
    for (int i = 0; i < TG_BR; ++i)
#pragma unroll
      for (int j = 0; j < TG_BC; ++j)
#pragma unroll
        for (int q = 0; q < S::CREGS; ++q) out[(m0 + i * R + S::rowmap(g, q)) * N + n0 + j * R + col] = acc[i][j][q];
  } else {
    T* X = reinterpret_cast<T*>(a.X);
    T* Xb = reinterpret_cast<T*>(a.Xb);
    const int t4 = col & 3;
    // per-column sums over this wave's rows: db (TG_ACT_BWD); db, dw_out
    // (TG_OUT_BWD); dW_0[.][0..2], db_0 (TG_FIRST_BWD)
    constexpr int NC = EPI == TG_ACT_BWD ? 1 : EPI == TG_OUT_BWD ? 2 : EPI == TG_FIRST_BWD ? 4 : 0;
    constexpr int NCA = NC > 0 ? NC : 1;
    float csum[TG_BC][NCA];
#pragma unroll
    for (int j = 0; j < TG_BC; ++j)
#pragma unroll
      for (int c = 0; c < NCA; ++c) csum[j][c] = 0.f;
    // per-column constants of the lane's TG_BC columns, loaded once
    float cbias[TG_BC], cwl[TG_BC];
    float4 cw0[TG_BC];
#pragma unroll
    for (int j = 0; j < TG_BC; ++j) {
      const int n = n0 + j * R + col;
      cbias[j] = (EPI == TG_ACT || EPI == TG_OUT || EPI == TG_OUT_BWD) && a.bias ? a.bias[n] : 0.f;
      cwl[j] = (EPI == TG_OUT || EPI == TG_OUT_BWD) ? a.wl[n] : 0.f;
      cw0[j] = EPI == TG_FIRST_BWD ? a.w0[n] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
#pragma unroll
    for (int i = 0; i < TG_BR; ++i) {
      // the block row's loads first, so their latencies overlap: per-point
      // inputs (seeds for TG_OUT_BWD, coordinates for TG_FIRST_BWD) and the
      // pre-activations of TG_ACT_BWD
      typedef T v4 __attribute__((ext_vector_type(4)));
      float4 pv[PTS];
      v4 zb[TG_BC][PTS];
#pragma unroll
      for (int pt = 0; pt < PTS; ++pt) {
        const uint64_t p = (uint64_t)(m0 + i * R + S::rowmap(g, 4 * pt)) >> 2;
        pv[pt] = (EPI == TG_OUT_BWD || EPI == TG_FIRST_BWD) ? a.pt4[p] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int j = 0; j < TG_BC; ++j) {
          if constexpr (EPI == TG_ACT_BWD)
            zb[j][pt] = reinterpret_cast<const v4*>(a.D)[p * (uint32_t)N + (uint32_t)(n0 + j * R + col)];
        }
      }
      float fo[PTS * 4];  // TG_OUT: (f, grad f) of the lane's points over its columns
#pragma unroll
      for (int k = 0; k < PTS * 4; ++k) fo[k] = 0.f;
#pragma unroll
      for (int j = 0; j < TG_BC; ++j) {
        const int n = n0 + j * R + col;
        const float bias = cbias[j], wl = cwl[j];
        const float4 w0 = cw0[j];
#pragma unroll
        for (int pt = 0; pt < PTS; ++pt) {
          const int q0 = 4 * pt;
          const int64_t mv = m0 + i * R + S::rowmap(g, q0);  // the point's value row
          const float v0 = acc[i][j][q0], v1 = acc[i][j][q0 + 1], v2 = acc[i][j][q0 + 2], v3 = acc[i][j][q0 + 3];
          if constexpr (EPI == TG_ACT) {
            // a = W p + b, p' = sin a, pd'_t = cos(a) ad_t; D = (cos a, sin(a) ad_t)
            float sn, cs;
            sincos_reduced(v0 + bias, sn, cs);
            st4rows(reinterpret_cast<T*>(a.D) + (((uint64_t)mv >> 2) * (uint32_t)N + (uint32_t)n) * 4, cs, sn * v1,
                    sn * v2, sn * v3);
            const float x1 = cs * v1, x2 = cs * v2, x3 = cs * v3;
            st_both(X, Xb, M, N, n, mv, t4, sn, x1, x2, x3);
          } else if constexpr (EPI == TG_OUT) {
            // f = sum_u wl_u sin a_u (+ b_out later), grad f_t = sum_u wl_u cos(a_u) ad_t,u
            float sn, cs;
            sincos_reduced(v0 + bias, sn, cs);
            const float wc = wl * cs;
            fo[q0] = fmaf(wl, sn, fo[q0]);
            fo[q0 + 1] = fmaf(wc, v1, fo[q0 + 1]);
            fo[q0 + 2] = fmaf(wc, v2, fo[q0 + 2]);
            fo[q0 + 3] = fmaf(wc, v3, fo[q0 + 3]);
          } else if constexpr (EPI == TG_OUT_BWD) {
            // seeds s = dL/df, g_t = dL/d(grad f)_t; pbar = s wl, pdbar_t = g_t wl:
            // abar = wl (s cos a - sin a sum_t g_t ad_t), adbar_t = wl g_t cos a,
            // dw_out += s sin a + cos a sum_t g_t ad_t
            float sn, cs;
            sincos_reduced(v0 + bias, sn, cs);
            const float4 sg = pv[pt];
            const float gad = fmaf(sg.w, v3, fmaf(sg.z, v2, sg.y * v1));
            const float abar = wl * fmaf(sg.x, cs, -sn * gad);
            const float wc = wl * cs;
            const float ad1 = wc * sg.y, ad2 = wc * sg.z, ad3 = wc * sg.w;
            csum[j][0] += abar;
            csum[j][1] += fmaf(sg.x, sn, cs * gad);
            st_both(X, Xb, M, N, n, mv, t4, abar, ad1, ad2, ad3);
          } else if constexpr (EPI == TG_ACT_BWD) {
            // abar = cos(a) pbar - sin(a) sum_t ad_t pdbar_t, adbar_t = cos(a) pdbar_t
            const v4 d = zb[j][pt];
            const float cs = (float)d[0];
            const float abar = fmaf(cs, v0, -fmaf((float)d[3], v3, fmaf((float)d[2], v2, (float)d[1] * v1)));
            const float ad1 = cs * v1, ad2 = cs * v2, ad3 = cs * v3;
            csum[j][0] += abar;
            st_both(X, Xb, M, N, n, mv, t4, abar, ad1, ad2, ad3);
          } else {  // TG_FIRST_BWD
            // layer 0: a = W_0 x + b_0, ad_t = W_0[.][t] (coordinate tangents);
            // dW_0[u][i] += abar x_i + adbar_i, db_0[u] += abar
            const float4 xp = pv[pt];
            const float z = fmaf(w0.z, xp.z, fmaf(w0.y, xp.y, fmaf(w0.x, xp.x, w0.w)));
            float sn, cs;
            sincos_reduced(z, sn, cs);
            const float abar = fmaf(cs, v0, -sn * fmaf(w0.z, v3, fmaf(w0.y, v2, w0.x * v1)));
            csum[j][0] += fmaf(abar, xp.x, cs * v1);
            csum[j][1] += fmaf(abar, xp.y, cs * v2);
            csum[j][2] += fmaf(abar, xp.z, cs * v3);
            csum[j][3] += abar;
          }
        }
        // one block at a time: interleaving the blocks' epilogues costs registers
        asm volatile("" ::: "memory");
      }
      if constexpr (EPI == TG_OUT) {
        // over the R columns of the lane group: lane col ends with value
        // k = (col / (R / V)) % V of its points (value k % 4 of point k / 4)
        constexpr int V = PTS * 4;
        lane_reduce<V, R / 2>(fo, col);
        if (col % (R / V) == 0) {
          const int k = (col / (R / V)) % V;
          const int64_t p = (m0 + i * R + S::rowmap(g, 4 * (k / 4))) / 4;
          a.part[((int64_t)(n0 / WN) * (M / 4) + p) * 4 + k % 4] = fo[0];
        }
      }
    }
    if constexpr (NC > 0) {
      // over the lane groups of each column, then one row of partial sums per wave
#pragma unroll
      for (int j = 0; j < TG_BC; ++j)
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          float v = csum[j][c];
          if (S::GROUPS == 4) v += __shfl_xor(v, 16);
          v += __shfl_xor(v, 32);
          if (g == 0) a.part[((m0 / WM) * NC + c) * N + n0 + j * R + col] = v;
        }
    }
  }
}

template <class T, int EPI>
static int launch_tg(const TgArgs& a, hipStream_t s) {
  constexpr int R = TgTraits<T>::S::ROWS;
  constexpr int TM = 2 * TG_BR * R, TN = 2 * TG_BC * R;
  MP_REQUIRE(a.M > 0 && a.N > 0 && a.K > 0 && a.M % TM == 0 && a.N % TN == 0 && a.K % (16 * TG_RING) == 0,
             MP_EINVAL, "training GEMM: %lld x %d x %lld is not a whole number of %d x %d x %d tiles",
             (long long)a.M, a.N, (long long)a.K, TM, TN, 16 * TG_RING);
  MP_REQUIRE(a.A && a.B, MP_EINVAL, "training GEMM: NULL operand");
  TgLaunch L{};
  L.a = a;
  L.ksteps = a.K / 16;
  L.mtiles = a.M / TM;
  L.ntn = a.N / TN;
  int64_t groups, per;
  if (EPI == TG_PART) {
    MP_REQUIRE(a.splits >= 1 && a.part, MP_EINVAL, "training GEMM: %d K chunks", a.splits);
    groups = a.splits;
    per = L.mtiles * L.ntn;
    L.steps_per_split = ((L.ksteps + a.splits - 1) / a.splits + TG_RING - 1) / TG_RING * TG_RING;
  } else {
    groups = L.mtiles;
    per = L.ntn;
    L.a.splits = 1;
    L.steps_per_split = L.ksteps;
  }
  const int64_t blocks = 8 * per * ((groups + 7) / 8);
  MP_REQUIRE(blocks < (1ll << 31), MP_EINVAL, "training GEMM: too many tiles");
  hipLaunchKernelGGL((k_tgemm<T, EPI>), dim3((unsigned)blocks), dim3(256), 0, s, L);
  MP_LAUNCHED("k_tgemm");
  return MP_OK;
}

template <class T>
static int launch_epi(int epi, const TgArgs& a, hipStream_t s) {
  switch (epi) {
    case TG_ACT:
      MP_REQUIRE(a.D && a.X, MP_EINVAL, "training GEMM: TG_ACT needs D and X");
      return launch_tg<T, TG_ACT>(a, s);
    case TG_OUT:
      MP_REQUIRE(a.wl && a.part, MP_EINVAL, "training GEMM: TG_OUT needs wl and part");
      return launch_tg<T, TG_OUT>(a, s);
    case TG_OUT_BWD:
      MP_REQUIRE(a.wl && a.pt4 && a.X && a.part, MP_EINVAL, "training GEMM: TG_OUT_BWD needs wl, seeds, X, part");
      return launch_tg<T, TG_OUT_BWD>(a, s);
    case TG_ACT_BWD:
      MP_REQUIRE(a.D && a.X && a.part, MP_EINVAL, "training GEMM: TG_ACT_BWD needs D, X and part");
      return launch_tg<T, TG_ACT_BWD>(a, s);
    case TG_FIRST_BWD:
      MP_REQUIRE(a.pt4 && a.w0 && a.part, MP_EINVAL, "training GEMM: TG_FIRST_BWD needs points, W_0 and part");
      return launch_tg<T, TG_FIRST_BWD>(a, s);
    case TG_PART:
      return launch_tg<T, TG_PART>(a, s);
  }
  set_error("training GEMM: epilogue %d", epi);
  return MP_EINVAL;
}

int tgemm(int dtype, int epi, const TgArgs& a, hipStream_t s) {
  switch (dtype) {
    case 0: return launch_epi<float>(epi, a, s);
    case 1: return launch_epi<__bf16>(epi, a, s);
    case 2: return launch_epi<_Float16>(epi, a, s);
  }
  set_error("training GEMM: element type %d", dtype);
  return MP_EINVAL;
}

int tgemm_tile_rows(int dtype) { return 2 * tgemm_wave_rows(dtype); }
int tgemm_tile_cols(int dtype) { return 2 * tgemm_wave_cols(dtype); }

int64_t tgemm_part_floats(int dtype, int epi, int64_t M, int N, int splits) {
  const int64_t wr = M / tgemm_wave_rows(dtype);
  switch (epi) {
    case TG_OUT: return (int64_t)(N / tgemm_wave_cols(dtype)) * M;
    case TG_OUT_BWD: return wr * 2 * N;
    case TG_ACT_BWD: return wr * N;
    case TG_FIRST_BWD: return wr * 4 * N;
    case TG_PART: return (int64_t)splits * M * N;
  }
  return 0;
}

int tgemm_weight_grad_chunks(int dtype, int64_t rows, int Wout, int Win) {
  const int tiles = (Wout / tgemm_tile_rows(dtype)) * (Win / tgemm_tile_cols(dtype));
  if (tiles <= 0) return 0;
  const int64_t steps = rows / 16;
  int64_t chunks = (256 + tiles - 1) / tiles;  // one workgroup per CU (the LDS ring allows one)
  chunks = (chunks + 7) / 8 * 8;
  const int64_t max_chunks = steps / TG_RING;
  if (chunks > max_chunks) chunks = max_chunks;
  return chunks < 1 ? 1 : (int)chunks;
}

}  // namespace mp
