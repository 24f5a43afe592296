// The training step of one residual level as two library calls (mp_train_fwd,
// mp_train_bwd), replacing the autograd graph of loss_add_detail_fine
// (i3d/loss_functions.py:327-394): a forward, a create_graph input gradient
// (i3d/diff_operators.py:259-265) and a double backward, as the trainers run
// it (experiment_scripts/train_sdf_on_neighborhood_fine.py:220-252).
//
// Notation per sine layer j (frequency baked into W, b):
//   a_j = W_j p_{j-1} + b_j,     p_j = sin a_j,
//   ad_j,t = W_j pd_{j-1},t,     pd_j,t = cos(a_j) ad_j,t       (t = 1..3)
// with p_{-1} = x and pd_{-1},t = e_t (the coordinate directions, so
// ad_0,t = W_0[.][t]).  Output f = w_out . p_{H-1} + b_out, grad f_t =
// w_out . pd_{H-1},t.  Rows 4p..4p+3 of every activation block are the value
// and the three tangents of point p, so one GEMM over 4 n rows applies W_j to
// all of them.  Backward, given the loss seeds s = dL/df and g = dL/dgrad f:
//   pbar_{H-1} = s w_out, pdbar_{H-1},t = g_t w_out,
//   abar_j    = cos(a_j) pbar_j - sin(a_j) sum_t ad_j,t pdbar_j,t
//   adbar_j,t = cos(a_j) pdbar_j,t
//   dW_j = [abar; adbar]_j^T [p; pd]_{j-1},  db_j = sum abar_j,
//   [pbar; pdbar]_{j-1} = [abar; adbar]_j W_j.
//
// Launches (H sine layers; train_gemm.hip holds the GEMMs and their epilogues):
//   forward   k_tw_first (layer 0) -> TG_ACT (layers 1..H-2) -> TG_OUT
//             (layer H-1 and the output layer, reduced per point) -> k_tw_out_sum
//   backward  TG_OUT_BWD (layer H-1 recomputed with the output-layer backward)
//             -> for j = H-1..1: TG_PART (dW_j), then TG_ACT_BWD (j >= 2) or
//             TG_FIRST_BWD (j = 1, layer 0's backward and parameter gradients)
// Stored between them: the unit-blocked activations X_j (GEMM A operands) and
// their row-blocked copies Xb_j (weight-gradient operands) of layers
// 0..H-2, and for layers 1..H-2 the derivative factors D_j = (cos a,
// sin(a) ad_t) that the backward through the activation needs, in the
// operand type (half the traffic of FP32 pre-activations in 16 bit,
// and no sine or cosine in the backward).  Layer 0's and layer H-1's
// pre-activations are recomputed where they are needed (three FMAs per unit
// for layer 0; the layer H-1 GEMM in the backward).  Widths are padded to 256 and
// points to 128 (zero weights, zero seeds), so the GEMMs need no edge
// handling; the padded units and points contribute nothing.
#include "common.h"
#include "mlp_core.h"
#include "stack.h"
#include "train_gemm.h"

namespace mp {
namespace {

constexpr int TW_PAD = 256;
constexpr int TP_PAD = 128;
constexpr int MAX_LAYERS = 16;

inline int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

struct TrainPlan {
  int dtype = 0, esize = 4, H = 0;
  int w[MAX_LAYERS] = {}, wp[MAX_LAYERS] = {};  // true / padded widths of the sine layers
  int64_t n = 0, np = 0, rows = 0;
  int64_t woff[MAX_LAYERS + 1] = {}, boff[MAX_LAYERS + 1] = {}, n_params = 0;
  // workspace offsets (bytes)
  size_t x4 = 0, w0 = 0, wl = 0, sg = 0;
  size_t bh[MAX_LAYERS] = {}, wf[MAX_LAYERS] = {}, wt[MAX_LAYERS] = {};
  size_t D[MAX_LAYERS] = {}, X[MAX_LAYERS] = {}, Xb[MAX_LAYERS] = {};
  size_t A[2] = {}, Ab[2] = {}, part_out = 0, part_col = 0, part_dw = 0, tmp = 0, red = 0, scl = 0;
  size_t total = 0;
  int dw_chunks[MAX_LAYERS] = {};
};

constexpr int64_t TMP_FLOATS = (int64_t)64 * 512 * 512;  // slice sums of the largest reduction

int plan_training(const mp_stack* st, int level, int64_t n, TrainPlan& p) {
  MP_REQUIRE(st != nullptr, MP_EINVAL, "training: stack is NULL");
  MP_REQUIRE(level >= 0 && level < st->n_levels, MP_EINVAL, "training: level %d of a %d-level stack", level,
             st->n_levels);
  MP_REQUIRE(n >= 1, MP_EINVAL, "training: n must be >= 1, got %lld", (long long)n);
  MP_REQUIRE(st->precision != MP_BF16X3 && st->precision != MP_BF16X3F, MP_EUNSUPPORTED,
             "training: three-term BF16 (MP_BF16X3, MP_BF16X3F) is not served by the training entry points");
  const mp_stack::Level& L = st->lv[level];
  p.H = L.n_hidden;
  MP_REQUIRE(p.H >= 2 && p.H < MAX_LAYERS, MP_EUNSUPPORTED, "training: a level needs 2..%d sine layers, got %d",
             MAX_LAYERS - 1, p.H);
  p.dtype = st->precision == MP_BF16 ? 1 : st->precision == MP_FP16 ? 2 : 0;
  p.esize = p.dtype == 0 ? 4 : 2;
  p.n = n;
  p.np = round_up(n, TP_PAD);
  p.rows = 4 * p.np;
  int in = 3;
  for (int k = 0; k < p.H; ++k) {
    p.w[k] = L.hidden[k];
    p.wp[k] = (int)round_up(p.w[k], TW_PAD);
    MP_REQUIRE(p.wp[k] <= 512, MP_EUNSUPPORTED, "training: width %d exceeds 512", p.w[k]);
    p.woff[k + 1] = p.woff[k] + (int64_t)p.w[k] * in;
    p.boff[k + 1] = p.boff[k] + p.w[k];
    in = p.w[k];
  }
  p.woff[p.H + 1] = p.woff[p.H] + in;  // output layer [1][w_{H-1}]
  p.boff[p.H + 1] = p.boff[p.H] + 1;
  p.n_params = p.woff[p.H + 1] + p.boff[p.H + 1];
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off = (off + bytes + 255) & ~(size_t)255;
    return o;
  };
  const int64_t R = p.rows;
  const int H = p.H;
  p.x4 = take((size_t)p.np * 16);
  p.w0 = take((size_t)p.wp[0] * 16);
  for (int j = 1; j < H; ++j) {
    p.bh[j] = take((size_t)p.wp[j] * 4);
    p.wf[j] = take((size_t)p.wp[j] * p.wp[j - 1] * p.esize);
    p.wt[j] = take((size_t)p.wp[j] * p.wp[j - 1] * p.esize);
  }
  p.wl = take((size_t)p.wp[H - 1] * 4);
  p.sg = take((size_t)p.np * 16);
  int wmax = 0;
  for (int k = 0; k < H - 1; ++k) {
    p.X[k] = take((size_t)R * p.wp[k] * p.esize);
    p.Xb[k] = take((size_t)R * p.wp[k] * p.esize);
    if (k >= 1) p.D[k] = take((size_t)R * p.wp[k] * p.esize);
  }
  for (int k = 1; k < H; ++k) wmax = p.wp[k] > wmax ? p.wp[k] : wmax;
  for (int c = 0; c < 2; ++c) {
    p.A[c] = take((size_t)R * wmax * p.esize);
    p.Ab[c] = take((size_t)R * wmax * p.esize);
  }
  p.part_out = take((size_t)tgemm_part_floats(p.dtype, TG_OUT, R, p.wp[H - 1], 0) * 4);
  int64_t pc = tgemm_part_floats(p.dtype, TG_OUT_BWD, R, p.wp[H - 1], 0);
  for (int j = 2; j < H; ++j) {
    const int64_t v = tgemm_part_floats(p.dtype, TG_ACT_BWD, R, p.wp[j - 1], 0);
    pc = v > pc ? v : pc;
  }
  const int64_t v0 = tgemm_part_floats(p.dtype, TG_FIRST_BWD, R, p.wp[0], 0);
  p.part_col = take((size_t)(v0 > pc ? v0 : pc) * 4);
  size_t dw = 0;
  for (int j = 1; j < H; ++j) {
    p.dw_chunks[j] = tgemm_weight_grad_chunks(p.dtype, R, p.wp[j], p.wp[j - 1]);
    const size_t b = (size_t)tgemm_part_floats(p.dtype, TG_PART, p.wp[j], p.wp[j - 1], p.dw_chunks[j]) * 4;
    dw = b > dw ? b : dw;
  }
  p.part_dw = take(dw);
  p.tmp = take((size_t)TMP_FLOATS * 4);
  p.red = take(256 * 4);
  p.scl = take(2 * 4);
  p.total = off;
  return MP_OK;
}

int ensure_train_ws(mp_stack* st, size_t bytes) {
  if (st->train_ws_bytes >= bytes) return MP_OK;
  if (st->train_ws) {
    (void)hipFree(st->train_ws);
    st->train_ws = nullptr;
    st->train_ws_bytes = 0;
  }
  st->kept_level = -1;
  const size_t b = (bytes + (1u << 20) - 1) & ~(size_t)((1u << 20) - 1);
  hipError_t e = hipMalloc(&st->train_ws, b);
  if (e != hipSuccess) {
    set_error("training workspace allocation of %zu bytes failed: %s", b, hipGetErrorString(e));
    return MP_ENOMEM;
  }
  st->train_ws_bytes = b;
  return MP_OK;
}

inline unsigned blocks_of(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// dst[i] = i < n ? src[i] : 0 for i < np
__global__ void k_tw_pad(const float* __restrict__ src, int64_t n, float* __restrict__ dst, int64_t np) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < np) dst[i] = i < n ? src[i] : 0.f;
}

// x4[p] = (x, y, z, 0) of point p (zero past n)
__global__ void k_tw_points(const float* __restrict__ x, int64_t n, float4* __restrict__ x4, int64_t np) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < np) x4[p] = p < n ? make_float4(x[3 * p], x[3 * p + 1], x[3 * p + 2], 0.f) : make_float4(0.f, 0.f, 0.f, 0.f);
}

// w0[u] = (W_0[u][0..2], b_0[u]) (zero past w)
__global__ void k_tw_first_w(const float* __restrict__ W0, const float* __restrict__ b0, int w, float4* __restrict__ w0,
                             int wp) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u < wp) w0[u] = u < w ? make_float4(W0[3 * u], W0[3 * u + 1], W0[3 * u + 2], b0[u]) : make_float4(0.f, 0.f, 0.f, 0.f);
}

// sg[p] = (dL/df, dL/dgrad f) of point p (NULL seeds and points past n: zero)
__global__ void k_tw_seeds(const float* __restrict__ s, const float* __restrict__ g, int64_t n, float4* __restrict__ sg,
                           int64_t np) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
  if (p < n) {
    if (s) v.x = s[p];
    if (g) {
      v.y = g[3 * p];
      v.z = g[3 * p + 1];
      v.w = g[3 * p + 2];
    }
  }
  sg[p] = v;
}

// FP16 seed scaling.  The backward stores every adjoint in the operand type;
// the seeds of a mean loss over n points are O(1/n), and in FP16 (smallest
// normal 6.1e-5) such adjoints lose their precision or flush to zero.  The
// seeds are therefore scaled by a power of two S that brings max |seed| into
// [2^10, 2^11), and the parameter gradients by 1/S at the end; both scalings
// are exact.  The pre-activation adjoints are the seeds times the output
// weights and cosines, so they sit well below the seeds (max |adjoint| /
// max |seed| ~ 0.006 on a 512x3 level with output weights x 0.1, ~0.06 with
// unscaled ones) and grow slowly through the layers: 2^10 keeps the small
// ones out of the subnormal range (at 2^4 the worst parameter tensor's error
// was 1.6x larger) with room up to FP16's 65504.  Block maxima of |sg| into
// red[0..255]:
__global__ __launch_bounds__(256) void k_tw_absmax(const float* __restrict__ src, int64_t n, float* __restrict__ out) {
  __shared__ float red[256];
  float m = 0.f;
  const int64_t step = (int64_t)gridDim.x * 256;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += step) m = fmaxf(m, fabsf(src[i]));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + h]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// scl = (S, 1 / S) from the 256 block maxima (S = 1 for all-zero seeds)
__global__ __launch_bounds__(256) void k_tw_seed_scale(const float* __restrict__ blockmax, float* __restrict__ scl) {
  __shared__ float red[256];
  red[threadIdx.x] = blockmax[threadIdx.x];
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + h]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const float m = red[0];
    int e = 0;
    if (m > 0.f) (void)frexpf(m, &e);  // m in [2^(e-1), 2^e)
    int k = m > 0.f ? 11 - e : 0;  // m S in [2^10, 2^11)
    k = k < -100 ? -100 : (k > 100 ? 100 : k);
    scl[0] = ldexpf(1.f, k);
    scl[1] = ldexpf(1.f, -k);
  }
}

// v[i] *= scl[which] (grid-stride)
__global__ void k_tw_scale(float* __restrict__ v, int64_t n, const float* __restrict__ scl, int which) {
  const float f = scl[which];
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step) v[i] *= f;
}

// hidden layer j as the B operands of the training GEMMs, zero padded, in
// the element type and the blocked layout [K/KL][N][KL]: Wf (forward: N = out,
// K = in) and Wt (backward: N = in, K = out)
template <class T>
__global__ void k_tw_hidden(const float* __restrict__ W, int wo, int wi, int wpo, int wpi, T* __restrict__ Wf,
                            T* __restrict__ Wt) {
  constexpr int KL = 16 / (int)sizeof(T);
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)wpo * wpi) return;
  const int i = (int)(e / wpi), k = (int)(e % wpi);
  const float v = (i < wo && k < wi) ? W[(int64_t)i * wi + k] : 0.f;
  Wf[((int64_t)(k / KL) * wpo + i) * KL + k % KL] = (T)v;
  Wt[((int64_t)(i / KL) * wpi + k) * KL + i % KL] = (T)v;
}

// Layer 0: a = W_0 x + b_0 and the rows (sin a, cos(a) W_0[u][0..2]) into X
// (unit-blocked [W/KL][4 np][KL]) and Xb (row-blocked [4 np / KL][W][KL]).
// One wave per 16 points x 64 units: lane u computes its unit's rows for the
// 16 points into LDS, then every store instruction writes 1 KB of contiguous
// memory: X by rows (lane l = row l of the 64 rows, one 16-byte chunk of KL
// units), Xb by units (lane l = unit l, one row group).  The grid is exact
// (np / 16 x W / 64 waves, a multiple of four).
template <class T>
__global__ __launch_bounds__(256) void k_tw_first(const float4* __restrict__ x4, const float4* __restrict__ w0,
                                                  int64_t np, int W, T* __restrict__ X, T* __restrict__ Xb) {
  constexpr int KL = 16 / (int)sizeof(T);
  constexpr int PW = 16;  // points per wave
  typedef T Vec __attribute__((ext_vector_type(KL)));
  typedef T V4 __attribute__((ext_vector_type(4)));
  __shared__ T lds[4][PW][65][4];  // [wave][point][unit (+1 against bank conflicts)][row]
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t wid = (int64_t)blockIdx.x * 4 + wave;
  const int ublocks = W / 64;
  const int64_t p0 = (wid / ublocks) * PW;
  const int u0 = (int)(wid % ublocks) * 64;
  const int64_t M = 4 * np;
  const float4 w = w0[u0 + lane];
#pragma unroll 4
  for (int pt = 0; pt < PW; ++pt) {
    const float4 xp = x4[p0 + pt];
    float sn, cs;
    sincos_reduced(fmaf(w.z, xp.z, fmaf(w.y, xp.y, fmaf(w.x, xp.x, w.w))), sn, cs);
    V4 o;
    o[0] = (T)sn;
    o[1] = (T)(cs * w.x);
    o[2] = (T)(cs * w.y);
    o[3] = (T)(cs * w.z);
    *reinterpret_cast<V4*>(&lds[wave][pt][lane][0]) = o;
  }
  __syncthreads();
  // X: chunk c holds units u0 + c KL .. + KL - 1; lane = row 4 (lane / 4) + lane % 4
#pragma unroll
  for (int c = 0; c < 64 / KL; ++c) {
    Vec o;
#pragma unroll
    for (int k = 0; k < KL; ++k) o[k] = lds[wave][lane / 4][c * KL + k][lane % 4];
    *reinterpret_cast<Vec*>(X + ((int64_t)(u0 / KL + c) * M + 4 * p0 + lane) * KL) = o;
  }
  // Xb: row groups of KL rows (two points in 16 bit, one in FP32); lane = unit
#pragma unroll
  for (int rg = 0; rg < 4 * PW / KL; ++rg) {
    Vec o;
#pragma unroll
    for (int k = 0; k < KL; ++k) o[k] = lds[wave][(rg * KL + k) / 4][lane][k % 4];
    *reinterpret_cast<Vec*>(Xb + ((4 * p0 / KL + rg) * W + u0 + lane) * KL) = o;
  }
}

// f[p] = b_out + sum of the column-range partials, grad[p] likewise (p < n)
__global__ void k_tw_out_sum(const float4* __restrict__ part, int slots, int64_t np, const float* __restrict__ b_out,
                             int64_t n, float* __restrict__ f, float* __restrict__ grad) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  float4 acc = part[p];
  for (int k = 1; k < slots; ++k) {
    const float4 v = part[(int64_t)k * np + p];
    acc.x += v.x;
    acc.y += v.y;
    acc.z += v.z;
    acc.w += v.w;
  }
  if (f) f[p] = acc.x + b_out[0];
  if (grad) {
    grad[3 * p] = acc.y;
    grad[3 * p + 1] = acc.z;
    grad[3 * p + 2] = acc.w;
  }
}

// out[b] = sum of src[i * stride] over the stripe i = b, b + gridDim.x, ...
// of i < n (workgroup b; 256 threads)
__global__ __launch_bounds__(256) void k_tw_sum(const float* __restrict__ src, int64_t n, int stride,
                                                float* __restrict__ out) {
  __shared__ float red[256];
  float acc = 0.f;
  const int64_t step = (int64_t)gridDim.x * 256;
#pragma unroll 4
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += step) acc += src[i * stride];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// tmp[slice][c] = sum over parts p = slice, slice + S, ... of part[p][c]
__global__ void k_sum_slices(const float* __restrict__ part, int64_t P, int64_t C, int S, float* __restrict__ tmp) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int slice = blockIdx.y;
  if (c >= C) return;
  float acc = 0.f;
#pragma unroll 8
  for (int64_t p = slice; p < P; p += S) acc += part[p * C + c];
  tmp[(int64_t)slice * C + c] = acc;
}

// out[r * out_rs + c] = sum_s tmp[s][off + r * rs + c * cs], r < rows, c < cols
__global__ void k_gather_sum(const float* __restrict__ tmp, int S, int64_t C, int rows, int cols, int64_t rs,
                             int64_t cs, int64_t off, float* __restrict__ out, int64_t out_rs) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)rows * cols) return;
  const int r = (int)(e / cols), c = (int)(e % cols);
  const int64_t src = off + r * rs + c * cs;
  float acc = 0.f;
#pragma unroll 8
  for (int sl = 0; sl < S; ++sl) acc += tmp[(int64_t)sl * C + src];
  out[(int64_t)r * out_rs + c] = acc;
}

// Sum P parts of C floats each, then write a rows x cols window of the sum.
int reduce_parts(const float* part, int64_t P, int64_t C, float* tmp, int rows, int cols, int64_t rs, int64_t cs,
                 int64_t off, float* out, int64_t out_rs, hipStream_t s) {
  int S = (int)(P / 64);
  S = S < 1 ? 1 : S > 64 ? 64 : S;
  MP_REQUIRE((int64_t)S * C <= TMP_FLOATS, MP_EINVAL, "training: reduction of %lld x %lld", (long long)P,
             (long long)C);
  hipLaunchKernelGGL(k_sum_slices, dim3(blocks_of(C, 256), S), dim3(256), 0, s, part, P, C, S, tmp);
  MP_LAUNCHED("k_sum_slices");
  hipLaunchKernelGGL(k_gather_sum, dim3(blocks_of((int64_t)rows * cols, 256)), dim3(256), 0, s, tmp, S, C, rows,
                     cols, rs, cs, off, out, out_rs);
  MP_LAUNCHED("k_gather_sum");
  return MP_OK;
}

inline char* at(const mp_stack* st, size_t off) { return static_cast<char*>(st->train_ws) + off; }
inline float* atf(const mp_stack* st, size_t off) { return reinterpret_cast<float*>(at(st, off)); }
inline float4* at4(const mp_stack* st, size_t off) { return reinterpret_cast<float4*>(at(st, off)); }

// padded operands from the level's FP32 weights (device) and the points
int stage_inputs(const mp_stack* st, const mp_stack::Level& L, const TrainPlan& p, const float* x, hipStream_t s) {
  const int H = p.H;
  hipLaunchKernelGGL(k_tw_points, dim3(blocks_of(p.np, 256)), dim3(256), 0, s, x, p.n, at4(st, p.x4), p.np);
  hipLaunchKernelGGL(k_tw_first_w, dim3(blocks_of(p.wp[0], 256)), dim3(256), 0, s, L.weights, L.biases, p.w[0],
                     at4(st, p.w0), p.wp[0]);
  for (int j = 1; j < H; ++j) {
    hipLaunchKernelGGL(k_tw_pad, dim3(blocks_of(p.wp[j], 256)), dim3(256), 0, s, L.biases + p.boff[j],
                       (int64_t)p.w[j], atf(st, p.bh[j]), (int64_t)p.wp[j]);
    const unsigned gb = blocks_of((int64_t)p.wp[j] * p.wp[j - 1], 256);
    const float* Wj = L.weights + p.woff[j];
    if (p.dtype == 1)
      hipLaunchKernelGGL(k_tw_hidden<__bf16>, dim3(gb), dim3(256), 0, s, Wj, p.w[j], p.w[j - 1], p.wp[j], p.wp[j - 1],
                         reinterpret_cast<__bf16*>(at(st, p.wf[j])), reinterpret_cast<__bf16*>(at(st, p.wt[j])));
    else if (p.dtype == 2)
      hipLaunchKernelGGL(k_tw_hidden<_Float16>, dim3(gb), dim3(256), 0, s, Wj, p.w[j], p.w[j - 1], p.wp[j],
                         p.wp[j - 1], reinterpret_cast<_Float16*>(at(st, p.wf[j])),
                         reinterpret_cast<_Float16*>(at(st, p.wt[j])));
    else
      hipLaunchKernelGGL(k_tw_hidden<float>, dim3(gb), dim3(256), 0, s, Wj, p.w[j], p.w[j - 1], p.wp[j], p.wp[j - 1],
                         atf(st, p.wf[j]), atf(st, p.wt[j]));
  }
  hipLaunchKernelGGL(k_tw_pad, dim3(blocks_of(p.wp[H - 1], 256)), dim3(256), 0, s, L.weights + p.woff[H],
                     (int64_t)p.w[H - 1], atf(st, p.wl), (int64_t)p.wp[H - 1]);
  MP_LAUNCHED("training inputs");
  return MP_OK;
}

TgArgs gemm_args(const void* A, const void* B, int64_t M, int N, int K) {
  TgArgs a{};
  a.A = A;
  a.B = B;
  a.M = M;
  a.N = N;
  a.K = K;
  return a;
}

// the forward with tangents: X_k, Xb_k (k <= H-2), Z_k (1 <= k <= H-2); with
// f or grad, also the output (layer H-1 and the output layer)
int run_forward(const mp_stack* st, const mp_stack::Level& L, const TrainPlan& p, float* f, float* grad,
                hipStream_t s) {
  const int H = p.H;
  const int64_t R = p.rows;
  const unsigned g0 = (unsigned)(p.np / 16 * (p.wp[0] / 64) / 4);  // exact: np % 128 == 0, wp % 256 == 0
  if (p.dtype == 1)
    hipLaunchKernelGGL(k_tw_first<__bf16>, dim3(g0), dim3(256), 0, s, at4(st, p.x4), at4(st, p.w0), p.np, p.wp[0],
                       reinterpret_cast<__bf16*>(at(st, p.X[0])), reinterpret_cast<__bf16*>(at(st, p.Xb[0])));
  else if (p.dtype == 2)
    hipLaunchKernelGGL(k_tw_first<_Float16>, dim3(g0), dim3(256), 0, s, at4(st, p.x4), at4(st, p.w0), p.np, p.wp[0],
                       reinterpret_cast<_Float16*>(at(st, p.X[0])), reinterpret_cast<_Float16*>(at(st, p.Xb[0])));
  else
    hipLaunchKernelGGL(k_tw_first<float>, dim3(g0), dim3(256), 0, s, at4(st, p.x4), at4(st, p.w0), p.np, p.wp[0],
                       atf(st, p.X[0]), atf(st, p.Xb[0]));
  MP_LAUNCHED("k_tw_first");
  int rc = MP_OK;
  for (int j = 1; j < H - 1; ++j) {
    TgArgs a = gemm_args(at(st, p.X[j - 1]), at(st, p.wf[j]), R, p.wp[j], p.wp[j - 1]);
    a.bias = atf(st, p.bh[j]);
    a.D = at(st, p.D[j]);
    a.X = at(st, p.X[j]);
    a.Xb = at(st, p.Xb[j]);
    rc = tgemm(p.dtype, TG_ACT, a, s);
    if (rc != MP_OK) return rc;
  }
  if (!f && !grad) return MP_OK;
  TgArgs a = gemm_args(at(st, p.X[H - 2]), at(st, p.wf[H - 1]), R, p.wp[H - 1], p.wp[H - 2]);
  a.bias = atf(st, p.bh[H - 1]);
  a.wl = atf(st, p.wl);
  a.part = atf(st, p.part_out);
  rc = tgemm(p.dtype, TG_OUT, a, s);
  if (rc != MP_OK) return rc;
  const int slots = p.wp[H - 1] / tgemm_wave_cols(p.dtype);
  hipLaunchKernelGGL(k_tw_out_sum, dim3(blocks_of(p.n, 256)), dim3(256), 0, s,
                     reinterpret_cast<const float4*>(at(st, p.part_out)), slots, p.np, L.biases + p.boff[H], p.n, f,
                     grad);
  MP_LAUNCHED("k_tw_out_sum");
  return MP_OK;
}

}  // namespace
}  // namespace mp

extern "C" {

int64_t mp_train_param_count(const mp_stack* st, int level) {
  mp::TrainPlan p;
  if (mp::plan_training(st, level, 1, p) != MP_OK) return -1;
  return p.n_params;
}

int mp_train_fwd(mp_stack* st, int level, const float* x, int64_t n, float* f, float* grad, int keep,
                 void* stream) {
  mp::TrainPlan p;
  int rc = mp::plan_training(st, level, n, p);
  if (rc != MP_OK) return rc;
  MP_REQUIRE(x != nullptr, MP_EINVAL, "mp_train_fwd: x is NULL");
  rc = mp::ensure_train_ws(st, p.total);
  if (rc != MP_OK) return rc;
  hipStream_t s = mp::as_stream(stream);
  st->kept_level = -1;
  const mp_stack::Level& L = st->lv[level];
  rc = mp::stage_inputs(st, L, p, x, s);
  if (rc != MP_OK) return rc;
  rc = mp::run_forward(st, L, p, f, grad, s);
  if (rc != MP_OK) return rc;
  if (keep) {
    st->kept_level = level;
    st->kept_x = x;
    st->kept_n = n;
    st->kept_version = st->version;
  }
  return MP_OK;
}

int mp_train_bwd(mp_stack* st, int level, const float* x, int64_t n, const float* dL_df, const float* dL_dgrad,
                 float* dparams, void* stream) {
  using namespace mp;
  TrainPlan p;
  int rc = plan_training(st, level, n, p);
  if (rc != MP_OK) return rc;
  MP_REQUIRE(x != nullptr && dparams != nullptr, MP_EINVAL, "mp_train_bwd: x and dparams must be non-NULL");
  MP_REQUIRE(dL_df != nullptr || dL_dgrad != nullptr, MP_EINVAL, "mp_train_bwd: no seeds (dL_df and dL_dgrad NULL)");
  hipStream_t s = as_stream(stream);
  const mp_stack::Level& L = st->lv[level];
  const bool kept = st->train_ws && st->kept_level == level && st->kept_x == x && st->kept_n == n &&
                    st->kept_version == st->version && st->train_ws_bytes >= p.total;
  if (!kept) {
    rc = ensure_train_ws(st, p.total);
    if (rc != MP_OK) return rc;
    rc = stage_inputs(st, L, p, x, s);
    if (rc != MP_OK) return rc;
    rc = run_forward(st, L, p, nullptr, nullptr, s);
    if (rc != MP_OK) return rc;
  }
  const int H = p.H;
  const int64_t R = p.rows;
  const int64_t WR = R / tgemm_wave_rows(p.dtype);  // partial-sum rows of the column epilogues
  float* tmp = atf(st, p.tmp);
  float* part = atf(st, p.part_col);
  float* dW = dparams;                  // weights of every layer, then
  float* dB = dparams + p.woff[H + 1];  // the biases
  hipLaunchKernelGGL(k_tw_seeds, dim3(blocks_of(p.np, 256)), dim3(256), 0, s, dL_df, dL_dgrad, n, at4(st, p.sg), p.np);
  MP_LAUNCHED("k_tw_seeds");
  const bool scaled = p.dtype == 2;  // FP16 adjoints: seed scaling (see k_tw_absmax)
  if (scaled) {
    hipLaunchKernelGGL(k_tw_absmax, dim3(256), dim3(256), 0, s, atf(st, p.sg), 4 * p.np, atf(st, p.red));
    hipLaunchKernelGGL(k_tw_seed_scale, dim3(1), dim3(256), 0, s, atf(st, p.red), atf(st, p.scl));
    hipLaunchKernelGGL(k_tw_scale, dim3(256), dim3(256), 0, s, atf(st, p.sg), 4 * p.np, atf(st, p.scl), 0);
    MP_LAUNCHED("k_tw_seed_scale");
  }
  // b_out: the sum of the value seeds
  hipLaunchKernelGGL(k_tw_sum, dim3(256), dim3(256), 0, s, atf(st, p.sg), p.np, 4, atf(st, p.red));
  hipLaunchKernelGGL(k_tw_sum, dim3(1), dim3(256), 0, s, atf(st, p.red), (int64_t)256, 1, dB + p.boff[H]);
  MP_LAUNCHED("k_tw_sum");
  // layer H-1 recomputed, through the output layer: Abar_{H-1}, db_{H-1}, dw_out
  {
    TgArgs a = gemm_args(at(st, p.X[H - 2]), at(st, p.wf[H - 1]), R, p.wp[H - 1], p.wp[H - 2]);
    a.bias = atf(st, p.bh[H - 1]);
    a.wl = atf(st, p.wl);
    a.pt4 = at4(st, p.sg);
    a.X = at(st, p.A[0]);
    a.Xb = at(st, p.Ab[0]);
    a.part = part;
    rc = tgemm(p.dtype, TG_OUT_BWD, a, s);
    if (rc != MP_OK) return rc;
    const int64_t C = 2 * (int64_t)p.wp[H - 1];
    rc = reduce_parts(part, WR, C, tmp, 1, p.w[H - 1], 0, 1, 0, dB + p.boff[H - 1], 0, s);
    if (rc != MP_OK) return rc;
    rc = reduce_parts(part, WR, C, tmp, 1, p.w[H - 1], 0, 1, p.wp[H - 1], dW + p.woff[H], 0, s);
    if (rc != MP_OK) return rc;
  }
  int cur = 0;
  for (int j = H - 1; j >= 1; --j) {
    // dW_j = Abar_j^T [p; pd]_{j-1} over all rows (split-K partials)
    {
      TgArgs a = gemm_args(at(st, p.Ab[cur]), at(st, p.Xb[j - 1]), p.wp[j], p.wp[j - 1], R);
      a.part = atf(st, p.part_dw);
      a.splits = p.dw_chunks[j];
      rc = tgemm(p.dtype, TG_PART, a, s);
      if (rc != MP_OK) return rc;
      rc = reduce_parts(atf(st, p.part_dw), p.dw_chunks[j], (int64_t)p.wp[j] * p.wp[j - 1], tmp, p.w[j], p.w[j - 1],
                        p.wp[j - 1], 1, 0, dW + p.woff[j], p.w[j - 1], s);
      if (rc != MP_OK) return rc;
    }
    // [pbar; pdbar]_{j-1} = [abar; adbar]_j W_j, then through layer j-1's activation
    TgArgs a = gemm_args(at(st, p.A[cur]), at(st, p.wt[j]), R, p.wp[j - 1], p.wp[j]);
    a.part = part;
    if (j >= 2) {
      a.D = at(st, p.D[j - 1]);
      a.X = at(st, p.A[1 - cur]);
      a.Xb = at(st, p.Ab[1 - cur]);
      rc = tgemm(p.dtype, TG_ACT_BWD, a, s);
      if (rc != MP_OK) return rc;
      rc = reduce_parts(part, WR, p.wp[j - 1], tmp, 1, p.w[j - 1], 0, 1, 0, dB + p.boff[j - 1], 0, s);
      if (rc != MP_OK) return rc;
      cur = 1 - cur;
    } else {
      // layer 0 with its parameter gradients: part [WR][4][wp0]
      a.pt4 = at4(st, p.x4);
      a.w0 = at4(st, p.w0);
      rc = tgemm(p.dtype, TG_FIRST_BWD, a, s);
      if (rc != MP_OK) return rc;
      const int64_t C = 4 * (int64_t)p.wp[0];
      rc = reduce_parts(part, WR, C, tmp, p.w[0], 3, 1, p.wp[0], 0, dW, 3, s);
      if (rc != MP_OK) return rc;
      rc = reduce_parts(part, WR, C, tmp, 1, p.w[0], 0, 1, 3 * (int64_t)p.wp[0], dB, 0, s);
      if (rc != MP_OK) return rc;
    }
  }
  if (scaled) {
    hipLaunchKernelGGL(k_tw_scale, dim3(256), dim3(256), 0, s, dparams, p.n_params, atf(st, p.scl), 1);
    MP_LAUNCHED("k_tw_scale");
  }
  return MP_OK;
}

}  // extern "C"
