// Elementwise stages of the residual-level training step (forward with input
// tangents and its hand-written backward, including the second-order terms of
// the input-gradient losses).  Replaces the autograd graph of
// loss_add_detail_fine (i3d/loss_functions.py:327-394: a forward, a
// create_graph gradient and a double backward) in the trainers
// (experiment_scripts/train_sdf_on_neighborhood_fine.py:220-252).
//
// Notation per sine layer j (frequency baked into W, b):
//   a_j = W_j p_{j-1} + b_j,     p_j = sin a_j,
//   ad_j,t = W_j pd_{j-1},t,     pd_j,t = cos(a_j) ad_j,t       (tangent t = 1..T)
// with p_{-1} = x and pd_{-1},t = v_t (the tangent directions).  Every
// activation block is laid out [n][1 + T][W]: row 0 the value, rows 1..T the
// tangents, so one GEMM over n (1 + T) rows applies W_j to all of them.
// Backward, given pbar_j (value row) and pdbar_j,t (tangent rows):
//   abar_j    = cos(a_j) pbar_j - sin(a_j) sum_t ad_j,t pdbar_j,t
//   adbar_j,t = cos(a_j) pdbar_j,t
// then dW_j = [abar; adbar]^T [p_{j-1}; pd_{j-1}], db_j = sum abar_j and
// [pbar; pdbar]_{j-1} = [abar; adbar] W_j (GEMMs, issued by the host).
#include "common.h"
#include "mlp_core.h"

namespace mp {

__device__ inline void st(float* p, int64_t i, float v) { p[i] = v; }
__device__ inline void st(__bf16* p, int64_t i, float v) { p[i] = (__bf16)v; }
__device__ inline void st(_Float16* p, int64_t i, float v) { p[i] = (_Float16)v; }

// Layer 0: Z = [a0; ad0] (float32), X = [sin a0; cos a0 * ad0].
template <class T>
__global__ void k_train_first(const float* __restrict__ x, const float* __restrict__ v, const float* __restrict__ W0,
                              const float* __restrict__ b0, int64_t n, int W, int nt, float* __restrict__ Z,
                              T* __restrict__ X) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * W) return;
  const int64_t p = i / W;
  const int u = (int)(i % W);
  const float w0 = W0[3 * u], w1 = W0[3 * u + 1], w2 = W0[3 * u + 2];
  const float a = fmaf(w2, x[3 * p + 2], fmaf(w1, x[3 * p + 1], fmaf(w0, x[3 * p], b0[u])));
  float s, c;
  sincos_reduced(a, s, c);
  const int64_t base = p * (1 + nt) * W + u;
  Z[base] = a;
  st(X, base, s);
  for (int t = 0; t < nt; ++t) {
    const float* vt = v + (p * nt + t) * 3;
    const float ad = fmaf(w2, vt[2], fmaf(w1, vt[1], w0 * vt[0]));
    Z[base + (int64_t)(1 + t) * W] = ad;
    st(X, base + (int64_t)(1 + t) * W, c * ad);
  }
}

// Hidden layer: Z holds W_j X_{j-1} (no bias); adds the bias to the value
// row in place and writes X_j.
template <class T>
__global__ void k_train_act(float* __restrict__ Z, const float* __restrict__ b, int64_t n, int W, int nt,
                            T* __restrict__ X) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * W) return;
  const int64_t p = i / W;
  const int u = (int)(i % W);
  const int64_t base = p * (1 + nt) * W + u;
  const float a = Z[base] + b[u];
  Z[base] = a;
  float s, c;
  sincos_reduced(a, s, c);
  st(X, base, s);
  for (int t = 0; t < nt; ++t) {
    const int64_t k = base + (int64_t)(1 + t) * W;
    st(X, k, c * Z[k]);
  }
}

// Backward through sin / the tangent products: Pbar -> Abar.
template <class T>
__global__ void k_train_act_bwd(const float* __restrict__ Z, const float* __restrict__ Pbar, int64_t n, int W,
                                int nt, T* __restrict__ Abar, float* __restrict__ abar_val) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * W) return;
  const int64_t p = i / W;
  const int u = (int)(i % W);
  const int64_t base = p * (1 + nt) * W + u;
  float s, c;
  sincos_reduced(Z[base], s, c);
  float acc = c * Pbar[base];
  for (int t = 0; t < nt; ++t) {
    const int64_t k = base + (int64_t)(1 + t) * W;
    const float pdb = Pbar[k];
    acc = fmaf(-s * Z[k], pdb, acc);
    st(Abar, k, c * pdb);
  }
  st(Abar, base, acc);
  if (abar_val) abar_val[p * W + u] = acc;
}

// Output layer seeds: pbar = s * w_L (value row), pdbar_t = g_t * w_L.
__global__ void k_train_seed(const float* __restrict__ s, const float* __restrict__ g, const float* __restrict__ wl,
                             int64_t n, int W, int nt, float* __restrict__ Pbar) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * W) return;
  const int64_t p = i / W;
  const int u = (int)(i % W);
  const int64_t base = p * (1 + nt) * W + u;
  const float w = wl[u];
  Pbar[base] = s ? s[p] * w : 0.f;
  for (int t = 0; t < nt; ++t) Pbar[base + (int64_t)(1 + t) * W] = g ? g[p * nt + t] * w : 0.f;
}

// Four consecutive units per thread (W % 4 == 0): 16-byte loads of the FP32
// blocks, 8- or 16-byte stores of the GEMM operands.  NT = tangent rows.
template <class T> struct Vec4;
template <> struct Vec4<float> { typedef float4 type; };
template <> struct Vec4<__bf16> { typedef __bf16 type __attribute__((ext_vector_type(4))); };
template <> struct Vec4<_Float16> { typedef _Float16 type __attribute__((ext_vector_type(4))); };

template <class T>
__device__ inline void st4(T* p, int64_t i, float4 v) {
  typename Vec4<T>::type o;
  o[0] = (T)v.x; o[1] = (T)v.y; o[2] = (T)v.z; o[3] = (T)v.w;
  *reinterpret_cast<typename Vec4<T>::type*>(p + i) = o;
}
template <>
__device__ inline void st4<float>(float* p, int64_t i, float4 v) { *reinterpret_cast<float4*>(p + i) = v; }
__device__ inline float4 ld4(const float* p, int64_t i) { return *reinterpret_cast<const float4*>(p + i); }
__device__ inline float4 ld4(const __bf16* p, int64_t i) {
  const Vec4<__bf16>::type v = *reinterpret_cast<const Vec4<__bf16>::type*>(p + i);
  return make_float4((float)v[0], (float)v[1], (float)v[2], (float)v[3]);
}
__device__ inline float4 ld4(const _Float16* p, int64_t i) {
  const Vec4<_Float16>::type v = *reinterpret_cast<const Vec4<_Float16>::type*>(p + i);
  return make_float4((float)v[0], (float)v[1], (float)v[2], (float)v[3]);
}

// K-blocked copy of a [n][4][W] operand block (three tangents), the layout of
// the weight-gradient GEMM: element (row r = 4 p + t, unit u) at
// ((r / 8) * W + u) * 8 + r % 8, so eight consecutive rows of one unit are 16
// contiguous bytes.  vals[t][e]: rows t = 0..3 of point p, units u..u+3.
template <class T>
__device__ inline void st_blocked(T* Xb, int64_t p, int W, int u, const float (&vals)[4][4]) {
  typedef T v4 __attribute__((ext_vector_type(4)));
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    v4 o;
    o[0] = (T)vals[0][e]; o[1] = (T)vals[1][e]; o[2] = (T)vals[2][e]; o[3] = (T)vals[3][e];
    *reinterpret_cast<v4*>(Xb + ((p >> 1) * W + u + e) * 8 + 4 * (p & 1)) = o;
  }
}

// Two consecutive points (8 rows, point pair `pair`) at once: one 16-byte
// store per unit.  lo / hi: rows of the even / odd point.
template <class T>
__device__ inline void st_blocked8(T* Xb, int64_t pair, int W, int u, const float (&lo)[4][4],
                                   const float (&hi)[4][4]) {
  typedef T v8 __attribute__((ext_vector_type(8)));
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    v8 o;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      o[t] = (T)lo[t][e];
      o[4 + t] = (T)hi[t][e];
    }
    *reinterpret_cast<v8*>(Xb + (pair * W + u + e) * 8) = o;
  }
}

// Blocked copy of P points' rows (P = 1: the point's half of each 16-byte
// chunk; P = 2: whole chunks).
template <class T, int P>
__device__ inline void st_blocked_p(T* Xb, int64_t p_first, int W, int u, const float (&rows)[P][4][4]) {
  if constexpr (P == 2) st_blocked8(Xb, p_first >> 1, W, u, rows[0], rows[1]);
  else st_blocked(Xb, p_first, W, u, rows[0]);
}

template <class T, int NT, int P>
__global__ __launch_bounds__(256) void k_train_first4(const float* __restrict__ x, const float* __restrict__ v,
                                                      const float* __restrict__ W0, const float* __restrict__ b0,
                                                      int64_t n, int W, float* __restrict__ Z, T* __restrict__ X,
                                                      T* __restrict__ Xb) {
  // P consecutive points x 4 units per thread
  const int W4 = W / 4;
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (n / P) * W4) return;
  const int64_t pq = q / W4;
  const int u = (int)(q - pq * W4) * 4;
  float rows[P][4][4];
#pragma unroll
  for (int pp = 0; pp < P; ++pp) {
    const int64_t p = pq * P + pp;
    const int64_t base = p * (1 + NT) * W + u;
    const float x0 = x[3 * p], x1 = x[3 * p + 1], x2 = x[3 * p + 2];
    float a[4], s[4], c[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const float* w = W0 + 3 * (u + e);
      a[e] = fmaf(w[2], x2, fmaf(w[1], x1, fmaf(w[0], x0, b0[u + e])));
      sincos_reduced(a[e], s[e], c[e]);
    }
    *reinterpret_cast<float4*>(Z + base) = make_float4(a[0], a[1], a[2], a[3]);
    st4(X, base, make_float4(s[0], s[1], s[2], s[3]));
#pragma unroll
    for (int e = 0; e < 4; ++e) rows[pp][0][e] = s[e];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const float* vt = v + (p * NT + t) * 3;
      float ad[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const float* w = W0 + 3 * (u + e);
        ad[e] = fmaf(w[2], vt[2], fmaf(w[1], vt[1], w[0] * vt[0]));
      }
      const int64_t k = base + (int64_t)(1 + t) * W;
      *reinterpret_cast<float4*>(Z + k) = make_float4(ad[0], ad[1], ad[2], ad[3]);
      st4(X, k, make_float4(c[0] * ad[0], c[1] * ad[1], c[2] * ad[2], c[3] * ad[3]));
      if constexpr (NT == 3) {
#pragma unroll
        for (int e = 0; e < 4; ++e) rows[pp][1 + t][e] = c[e] * ad[e];
      }
    }
  }
  if constexpr (NT == 3) {
    if (Xb) st_blocked_p<T, P>(Xb, pq * P, W, u, rows);
  }
}

template <class T, int NT, int P>
__global__ __launch_bounds__(256) void k_train_act4(float* __restrict__ Z, const float* __restrict__ b, int64_t n,
                                                    int W, T* __restrict__ X, T* __restrict__ Xb) {
  // P consecutive points x 4 units per thread
  const int W4 = W / 4;
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (n / P) * W4) return;
  const int64_t pq = q / W4;
  const int u = (int)(q - pq * W4) * 4;
  const float4 bb = *reinterpret_cast<const float4*>(b + u);
  float rows[P][4][4];
#pragma unroll
  for (int pp = 0; pp < P; ++pp) {
    const int64_t base = (pq * P + pp) * (1 + NT) * W + u;
    const float4 z = ld4(Z, base);
    const float a[4] = {z.x + bb.x, z.y + bb.y, z.z + bb.z, z.w + bb.w};
    float s[4], c[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) sincos_reduced(a[e], s[e], c[e]);
    *reinterpret_cast<float4*>(Z + base) = make_float4(a[0], a[1], a[2], a[3]);
    st4(X, base, make_float4(s[0], s[1], s[2], s[3]));
#pragma unroll
    for (int e = 0; e < 4; ++e) rows[pp][0][e] = s[e];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int64_t k = base + (int64_t)(1 + t) * W;
      const float4 ad = ld4(Z, k);
      const float o[4] = {c[0] * ad.x, c[1] * ad.y, c[2] * ad.z, c[3] * ad.w};
      st4(X, k, make_float4(o[0], o[1], o[2], o[3]));
      if constexpr (NT == 3) {
#pragma unroll
        for (int e = 0; e < 4; ++e) rows[pp][1 + t][e] = o[e];
      }
    }
  }
  if constexpr (NT == 3) {
    if (Xb) st_blocked_p<T, P>(Xb, pq * P, W, u, rows);
  }
}

template <class T, int NT>
__global__ __launch_bounds__(256) void k_train_act_bwd4(const float* __restrict__ Z, const float* __restrict__ Pbar,
                                                        int64_t n, int W, T* __restrict__ Abar,
                                                        float* __restrict__ abar_val, T* __restrict__ Ab) {
  const int W4 = W / 4;
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n * W4) return;
  const int64_t p = q / W4;
  const int u = (int)(q - p * W4) * 4;
  const int64_t base = p * (1 + NT) * W + u;
  const float4 z = ld4(Z, base), pb = ld4(Pbar, base);
  const float za[4] = {z.x, z.y, z.z, z.w}, pa[4] = {pb.x, pb.y, pb.z, pb.w};
  float s[4], c[4], acc[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    sincos_reduced(za[e], s[e], c[e]);
    acc[e] = c[e] * pa[e];
  }
  float rows[4][4];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int64_t k = base + (int64_t)(1 + t) * W;
    const float4 zt = ld4(Z, k), pt = ld4(Pbar, k);
    const float zv[4] = {zt.x, zt.y, zt.z, zt.w}, pv[4] = {pt.x, pt.y, pt.z, pt.w};
#pragma unroll
    for (int e = 0; e < 4; ++e) acc[e] = fmaf(-s[e] * zv[e], pv[e], acc[e]);
    const float o[4] = {c[0] * pv[0], c[1] * pv[1], c[2] * pv[2], c[3] * pv[3]};
    st4(Abar, k, make_float4(o[0], o[1], o[2], o[3]));
    if constexpr (NT == 3) {
#pragma unroll
      for (int e = 0; e < 4; ++e) rows[1 + t][e] = o[e];
    }
  }
  st4(Abar, base, make_float4(acc[0], acc[1], acc[2], acc[3]));
  if (abar_val) *reinterpret_cast<float4*>(abar_val + p * W + u) = make_float4(acc[0], acc[1], acc[2], acc[3]);
  if constexpr (NT == 3) {
    if (Ab) {
#pragma unroll
      for (int e = 0; e < 4; ++e) rows[0][e] = acc[e];
      st_blocked(Ab, p, W, u, rows);
    }
  }
}

static inline unsigned grid_for(int64_t n) { return (unsigned)((n + 255) / 256); }

// Launch K<T, NT> for the run-time tangent count.
#define MP_TRAIN_NT_DISPATCH(KERNEL, T, nt, grid, s, ...)                                     \
  do {                                                                                        \
    switch (nt) {                                                                             \
      case 0: hipLaunchKernelGGL((KERNEL<T, 0>), dim3(grid), dim3(256), 0, s, __VA_ARGS__); break; \
      case 1: hipLaunchKernelGGL((KERNEL<T, 1>), dim3(grid), dim3(256), 0, s, __VA_ARGS__); break; \
      case 2: hipLaunchKernelGGL((KERNEL<T, 2>), dim3(grid), dim3(256), 0, s, __VA_ARGS__); break; \
      default: hipLaunchKernelGGL((KERNEL<T, 3>), dim3(grid), dim3(256), 0, s, __VA_ARGS__); break; \
    }                                                                                         \
  } while (0)

// The same for kernels with P points per thread.
#define MP_TRAIN_NTP_DISPATCH(KERNEL, T, P, nt, grid, s, ...)                                       \
  do {                                                                                              \
    switch (nt) {                                                                                   \
      case 0: hipLaunchKernelGGL((KERNEL<T, 0, P>), dim3(grid), dim3(256), 0, s, __VA_ARGS__); break;   \
      case 1: hipLaunchKernelGGL((KERNEL<T, 1, P>), dim3(grid), dim3(256), 0, s, __VA_ARGS__); break;   \
      case 2: hipLaunchKernelGGL((KERNEL<T, 2, P>), dim3(grid), dim3(256), 0, s, __VA_ARGS__); break;   \
      default: hipLaunchKernelGGL((KERNEL<T, 3, P>), dim3(grid), dim3(256), 0, s, __VA_ARGS__); break;  \
    }                                                                                               \
  } while (0)

}  // namespace mp

using namespace mp;

extern "C" {

int mp_train_first(const float* x, const float* v, const float* W0, const float* b0, int64_t n, int W,
                   int n_tangents, float* Z, void* X, int x_dtype, void* Xb, void* stream) {
  MP_REQUIRE(x && W0 && b0 && Z && X && n >= 0 && W > 0 && n_tangents >= 0 && n_tangents <= 3, MP_EINVAL,
             "mp_train_first: bad arguments");
  MP_REQUIRE(n_tangents == 0 || v, MP_EINVAL, "mp_train_first: tangent directions are NULL");
  MP_REQUIRE(!Xb || (W % 4 == 0 && n_tangents == 3 && n % 2 == 0 && x_dtype != 0), MP_EINVAL,
             "mp_train_first: the blocked copy needs 16-bit X, 3 tangents, even n and W % 4 == 0");
  if (n == 0) return MP_OK;
  hipStream_t s = as_stream(stream);
  if (W % 4 == 0) {
    // with the blocked copy: two points per thread (whole 16-byte chunks)
    const unsigned g = grid_for((Xb ? n / 2 : n) * (W / 4));
    if (x_dtype == 1 && Xb)
      MP_TRAIN_NTP_DISPATCH(k_train_first4, __bf16, 2, n_tangents, g, s, x, v, W0, b0, n, W, Z,
                            reinterpret_cast<__bf16*>(X), reinterpret_cast<__bf16*>(Xb));
    else if (x_dtype == 1)
      MP_TRAIN_NTP_DISPATCH(k_train_first4, __bf16, 1, n_tangents, g, s, x, v, W0, b0, n, W, Z,
                            reinterpret_cast<__bf16*>(X), static_cast<__bf16*>(nullptr));
    else if (x_dtype == 2 && Xb)
      MP_TRAIN_NTP_DISPATCH(k_train_first4, _Float16, 2, n_tangents, g, s, x, v, W0, b0, n, W, Z,
                            reinterpret_cast<_Float16*>(X), reinterpret_cast<_Float16*>(Xb));
    else if (x_dtype == 2)
      MP_TRAIN_NTP_DISPATCH(k_train_first4, _Float16, 1, n_tangents, g, s, x, v, W0, b0, n, W, Z,
                            reinterpret_cast<_Float16*>(X), static_cast<_Float16*>(nullptr));
    else
      MP_TRAIN_NTP_DISPATCH(k_train_first4, float, 1, n_tangents, g, s, x, v, W0, b0, n, W, Z,
                            reinterpret_cast<float*>(X), static_cast<float*>(nullptr));
    MP_LAUNCHED("k_train_first4");
    return MP_OK;
  }
  if (x_dtype == 1)
    hipLaunchKernelGGL(k_train_first<__bf16>, dim3(grid_for(n * W)), dim3(256), 0, s, x, v, W0, b0, n, W,
                       n_tangents, Z, reinterpret_cast<__bf16*>(X));
  else if (x_dtype == 2)
    hipLaunchKernelGGL(k_train_first<_Float16>, dim3(grid_for(n * W)), dim3(256), 0, s, x, v, W0, b0, n, W,
                       n_tangents, Z, reinterpret_cast<_Float16*>(X));
  else
    hipLaunchKernelGGL(k_train_first<float>, dim3(grid_for(n * W)), dim3(256), 0, s, x, v, W0, b0, n, W,
                       n_tangents, Z, reinterpret_cast<float*>(X));
  MP_LAUNCHED("k_train_first");
  return MP_OK;
}

int mp_train_act(float* Z, const float* b, int64_t n, int W, int n_tangents, void* X, int x_dtype, void* Xb,
                 void* stream) {
  MP_REQUIRE(Z && b && X && n >= 0 && W > 0 && n_tangents >= 0 && n_tangents <= 3, MP_EINVAL,
             "mp_train_act: bad arguments");
  MP_REQUIRE(!Xb || (W % 4 == 0 && n_tangents == 3 && n % 2 == 0 && x_dtype != 0), MP_EINVAL,
             "mp_train_act: the blocked copy needs 16-bit X, 3 tangents, even n and W % 4 == 0");
  if (n == 0) return MP_OK;
  hipStream_t s = as_stream(stream);
  if (W % 4 == 0) {
    // with the blocked copy: two points per thread (whole 16-byte chunks)
    const unsigned g = grid_for((Xb ? n / 2 : n) * (W / 4));
    if (x_dtype == 1 && Xb)
      MP_TRAIN_NTP_DISPATCH(k_train_act4, __bf16, 2, n_tangents, g, s, Z, b, n, W, reinterpret_cast<__bf16*>(X),
                            reinterpret_cast<__bf16*>(Xb));
    else if (x_dtype == 1)
      MP_TRAIN_NTP_DISPATCH(k_train_act4, __bf16, 1, n_tangents, g, s, Z, b, n, W, reinterpret_cast<__bf16*>(X),
                            static_cast<__bf16*>(nullptr));
    else if (x_dtype == 2 && Xb)
      MP_TRAIN_NTP_DISPATCH(k_train_act4, _Float16, 2, n_tangents, g, s, Z, b, n, W,
                            reinterpret_cast<_Float16*>(X), reinterpret_cast<_Float16*>(Xb));
    else if (x_dtype == 2)
      MP_TRAIN_NTP_DISPATCH(k_train_act4, _Float16, 1, n_tangents, g, s, Z, b, n, W,
                            reinterpret_cast<_Float16*>(X), static_cast<_Float16*>(nullptr));
    else
      MP_TRAIN_NTP_DISPATCH(k_train_act4, float, 1, n_tangents, g, s, Z, b, n, W, reinterpret_cast<float*>(X),
                            static_cast<float*>(nullptr));
    MP_LAUNCHED("k_train_act4");
    return MP_OK;
  }
  if (x_dtype == 1)
    hipLaunchKernelGGL(k_train_act<__bf16>, dim3(grid_for(n * W)), dim3(256), 0, s, Z, b, n, W, n_tangents,
                       reinterpret_cast<__bf16*>(X));
  else if (x_dtype == 2)
    hipLaunchKernelGGL(k_train_act<_Float16>, dim3(grid_for(n * W)), dim3(256), 0, s, Z, b, n, W, n_tangents,
                       reinterpret_cast<_Float16*>(X));
  else
    hipLaunchKernelGGL(k_train_act<float>, dim3(grid_for(n * W)), dim3(256), 0, s, Z, b, n, W, n_tangents,
                       reinterpret_cast<float*>(X));
  MP_LAUNCHED("k_train_act");
  return MP_OK;
}

int mp_train_act_bwd(const float* Z, const float* Pbar, int64_t n, int W, int n_tangents, void* Abar,
                     int a_dtype, float* abar_val, void* Ab, void* stream) {
  MP_REQUIRE(Z && Pbar && Abar && n >= 0 && W > 0 && n_tangents >= 0 && n_tangents <= 3, MP_EINVAL,
             "mp_train_act_bwd: bad arguments");
  MP_REQUIRE(!Ab || (W % 4 == 0 && n_tangents == 3 && n % 2 == 0 && a_dtype != 0), MP_EINVAL,
             "mp_train_act_bwd: the blocked copy needs 16-bit A, 3 tangents, even n and W % 4 == 0");
  if (n == 0) return MP_OK;
  hipStream_t s = as_stream(stream);
  if (W % 4 == 0) {
    const unsigned g = grid_for(n * (W / 4));
    if (a_dtype == 1)
      MP_TRAIN_NT_DISPATCH(k_train_act_bwd4, __bf16, n_tangents, g, s, Z, Pbar, n, W, reinterpret_cast<__bf16*>(Abar),
                           abar_val, reinterpret_cast<__bf16*>(Ab));
    else if (a_dtype == 2)
      MP_TRAIN_NT_DISPATCH(k_train_act_bwd4, _Float16, n_tangents, g, s, Z, Pbar, n, W,
                           reinterpret_cast<_Float16*>(Abar), abar_val, reinterpret_cast<_Float16*>(Ab));
    else
      MP_TRAIN_NT_DISPATCH(k_train_act_bwd4, float, n_tangents, g, s, Z, Pbar, n, W, reinterpret_cast<float*>(Abar),
                           abar_val, static_cast<float*>(nullptr));
    MP_LAUNCHED("k_train_act_bwd4");
    return MP_OK;
  }
  if (a_dtype == 1)
    hipLaunchKernelGGL(k_train_act_bwd<__bf16>, dim3(grid_for(n * W)), dim3(256), 0, s, Z, Pbar, n, W, n_tangents,
                       reinterpret_cast<__bf16*>(Abar), abar_val);
  else if (a_dtype == 2)
    hipLaunchKernelGGL(k_train_act_bwd<_Float16>, dim3(grid_for(n * W)), dim3(256), 0, s, Z, Pbar, n, W,
                       n_tangents, reinterpret_cast<_Float16*>(Abar), abar_val);
  else
    hipLaunchKernelGGL(k_train_act_bwd<float>, dim3(grid_for(n * W)), dim3(256), 0, s, Z, Pbar, n, W, n_tangents,
                       reinterpret_cast<float*>(Abar), abar_val);
  MP_LAUNCHED("k_train_act_bwd");
  return MP_OK;
}

int mp_train_seed(const float* s_val, const float* g_tan, const float* wl, int64_t n, int W, int n_tangents,
                  float* Pbar, void* stream) {
  MP_REQUIRE(wl && Pbar && n >= 0 && W > 0 && n_tangents >= 0 && n_tangents <= 3, MP_EINVAL,
             "mp_train_seed: bad arguments");
  if (n == 0) return MP_OK;
  hipLaunchKernelGGL(k_train_seed, dim3(grid_for(n * W)), dim3(256), 0, as_stream(stream), s_val, g_tan, wl, n, W,
                     n_tangents, Pbar);
  MP_LAUNCHED("k_train_seed");
  return MP_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Output layer (W -> 1) of the training forward / backward, fused:
//   forward:  f = w_L . sin(a) + b_L,  fd_t = w_L . (cos(a) ad_t)
//   backward: Pbar = [s w_L; g_t w_L] and the partial sums of
//             dw_L = sum_points s sin(a) + sum_t g_t cos(a) ad_t
// One wave per point for the forward (lane-strided dot products, then a
// wave reduction); the backward accumulates dw_L per block.
// ---------------------------------------------------------------------------
namespace mp {

// b_sine (may be NULL): bias of the last sine layer, added to the value row
// of Z in place first (Z then holds that layer's full pre-activations).
__global__ __launch_bounds__(256) void k_train_out_fwd(float* __restrict__ Z, const float* __restrict__ b_sine,
                                                       const float* __restrict__ wl, const float* __restrict__ bl,
                                                       int64_t n, int W, int nt, float* __restrict__ f,
                                                       float* __restrict__ fd) {
  const int64_t p = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (p >= n) return;
  float acc[4] = {0.f, 0.f, 0.f, 0.f};
  float* z = Z + p * (1 + nt) * W;
  for (int u = lane; u < W; u += 64) {
    float a = z[u];
    if (b_sine) {
      a += b_sine[u];
      z[u] = a;
    }
    float s, c;
    sincos_reduced(a, s, c);
    const float w = wl[u];
    acc[0] = fmaf(w, s, acc[0]);
    for (int t = 0; t < nt; ++t) acc[1 + t] = fmaf(w * c, z[(1 + t) * W + u], acc[1 + t]);
  }
  for (int k = 0; k <= nt; ++k)
    for (int o = 32; o > 0; o >>= 1) acc[k] += __shfl_xor(acc[k], o);
  if (lane == 0) {
    f[p] = acc[0] + bl[0];
    for (int t = 0; t < nt; ++t) fd[p * nt + t] = acc[1 + t];
  }
}

// The same with 16 lanes per point (W % 64 == 0): each lane reads 4 units
// per row with one 16-byte load, a row's 16 lanes 256 contiguous bytes, and
// the sums reduce over 16 lanes; 16 points per block.
template <int NT>
__global__ __launch_bounds__(256) void k_train_out_fwd16(float* __restrict__ Z, const float* __restrict__ b_sine,
                                                         const float* __restrict__ wl, const float* __restrict__ bl,
                                                         int64_t n, int W, float* __restrict__ f,
                                                         float* __restrict__ fd) {
  const int64_t p = (int64_t)blockIdx.x * 16 + (threadIdx.x >> 4);
  const int j = threadIdx.x & 15;
  if (p >= n) return;  // the whole 16-lane group
  float acc[1 + NT];
#pragma unroll
  for (int k = 0; k <= NT; ++k) acc[k] = 0.f;
  float* z = Z + p * (1 + NT) * W;
  for (int u = 4 * j; u < W; u += 64) {
    float4 a = ld4(z, u);
    if (b_sine) {
      const float4 bb = ld4(b_sine, u);
      a = make_float4(a.x + bb.x, a.y + bb.y, a.z + bb.z, a.w + bb.w);
      *reinterpret_cast<float4*>(z + u) = a;
    }
    const float4 w4 = ld4(wl, u);
    const float av[4] = {a.x, a.y, a.z, a.w}, wv[4] = {w4.x, w4.y, w4.z, w4.w};
    float sn[4], cs[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      sincos_reduced(av[e], sn[e], cs[e]);
      acc[0] = fmaf(wv[e], sn[e], acc[0]);
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const float4 zt = ld4(z, (int64_t)(1 + t) * W + u);
      const float zv[4] = {zt.x, zt.y, zt.z, zt.w};
#pragma unroll
      for (int e = 0; e < 4; ++e) acc[1 + t] = fmaf(wv[e] * cs[e], zv[e], acc[1 + t]);
    }
  }
#pragma unroll
  for (int k = 0; k <= NT; ++k)
    for (int o = 8; o > 0; o >>= 1) acc[k] += __shfl_xor(acc[k], o);
  if (j == 0) {
    f[p] = acc[0] + bl[0];
#pragma unroll
    for (int t = 0; t < NT; ++t) fd[p * NT + t] = acc[1 + t];
  }
}

// grid.x = ceil(n / rows_per_block), block = 256 threads over W (W <= 512:
// two units per thread); dwl_part [gridDim.x][W].
__global__ __launch_bounds__(256) void k_train_out_bwd(const float* __restrict__ Z, const float* __restrict__ s,
                                                       const float* __restrict__ g, const float* __restrict__ wl,
                                                       int64_t n, int W, int nt, int rows_per_block,
                                                       float* __restrict__ Pbar, float* __restrict__ dwl_part) {
  const int64_t p0 = (int64_t)blockIdx.x * rows_per_block;
  for (int u = threadIdx.x; u < W; u += blockDim.x) {
    const float w = wl[u];
    float acc = 0.f;
    for (int r = 0; r < rows_per_block; ++r) {
      const int64_t p = p0 + r;
      if (p >= n) break;
      const float* z = Z + p * (1 + nt) * W;
      float* pb = Pbar + p * (1 + nt) * W;
      float sn, cs;
      sincos_reduced(z[u], sn, cs);
      const float sv = s ? s[p] : 0.f;
      pb[u] = sv * w;
      acc = fmaf(sv, sn, acc);
      for (int t = 0; t < nt; ++t) {
        const float gv = g ? g[p * nt + t] : 0.f;
        pb[(1 + t) * W + u] = gv * w;
        acc = fmaf(gv * cs, z[(1 + t) * W + u], acc);
      }
    }
    dwl_part[(int64_t)blockIdx.x * W + u] = acc;
  }
}

// Rows (points) per block of the fused reductions below.
constexpr int RED_ROWS = 64;

// Output layer backward, four units per thread (W % 4 == 0, blockDim = W / 4).
template <int NT>
__global__ __launch_bounds__(128) void k_train_out_bwd4(const float* __restrict__ Z, const float* __restrict__ s,
                                                        const float* __restrict__ g, const float* __restrict__ wl,
                                                        int64_t n, int W, float* __restrict__ Pbar,
                                                        float* __restrict__ dwl_part) {
  const int u = threadIdx.x * 4;
  if (u >= W) return;
  const float4 w4 = ld4(wl, u);
  float acc[4] = {0.f, 0.f, 0.f, 0.f};
  const int64_t p0 = (int64_t)blockIdx.x * RED_ROWS;
  for (int r = 0; r < RED_ROWS; ++r) {
    const int64_t p = p0 + r;
    if (p >= n) break;
    const int64_t base = p * (1 + NT) * W + u;
    const float4 z = ld4(Z, base);
    const float za[4] = {z.x, z.y, z.z, z.w};
    float sn[4], cs[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) sincos_reduced(za[e], sn[e], cs[e]);
    const float sv = s ? s[p] : 0.f;
    *reinterpret_cast<float4*>(Pbar + base) = make_float4(sv * w4.x, sv * w4.y, sv * w4.z, sv * w4.w);
#pragma unroll
    for (int e = 0; e < 4; ++e) acc[e] = fmaf(sv, sn[e], acc[e]);
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int64_t k = base + (int64_t)(1 + t) * W;
      const float gv = g ? g[p * NT + t] : 0.f;
      *reinterpret_cast<float4*>(Pbar + k) = make_float4(gv * w4.x, gv * w4.y, gv * w4.z, gv * w4.w);
      const float4 zt = ld4(Z, k);
      const float zv[4] = {zt.x, zt.y, zt.z, zt.w};
#pragma unroll
      for (int e = 0; e < 4; ++e) acc[e] = fmaf(gv * cs[e], zv[e], acc[e]);
    }
  }
  *reinterpret_cast<float4*>(dwl_part + (int64_t)blockIdx.x * W + u) = make_float4(acc[0], acc[1], acc[2], acc[3]);
}

// Output layer and last sine layer backward in one pass (W % 4 == 0,
// blockDim = W / 4).  The seeds of the last sine layer are rank one,
// pbar = s w_L and pdbar_t = g_t w_L, so they are formed in registers:
//   abar = cos a pbar - sin a sum_t ad_t pdbar_t,  adbar_t = cos a pdbar_t
// are written as the GEMM operand (and its K-blocked copy), and
// part [blocks][2][W] receives per-block partial sums of
// dw_L = sum s sin a + sum_t g_t cos a ad_t (row 0) and db = sum abar (row 1).
template <class T, int NT, int P>
__global__ __launch_bounds__(128) void k_train_out_act_bwd4(const float* __restrict__ Z, const float* __restrict__ s,
                                                            const float* __restrict__ g, const float* __restrict__ wl,
                                                            int64_t n, int W, T* __restrict__ Abar, T* __restrict__ Ab,
                                                            float* __restrict__ part) {
  const int u = threadIdx.x * 4;
  if (u >= W) return;
  const float4 w4 = ld4(wl, u);
  const float wv[4] = {w4.x, w4.y, w4.z, w4.w};
  float dwl[4] = {0.f, 0.f, 0.f, 0.f}, db[4] = {0.f, 0.f, 0.f, 0.f};
  const int64_t p0 = (int64_t)blockIdx.x * RED_ROWS;
  for (int r = 0; r < RED_ROWS; r += P) {
    if (p0 + r >= n) break;  // n is even when P = 2
    float rows[P][4][4];
#pragma unroll
    for (int pp = 0; pp < P; ++pp) {
      const int64_t p = p0 + r + pp;
      const int64_t base = p * (1 + NT) * W + u;
      const float4 z = ld4(Z, base);
      const float za[4] = {z.x, z.y, z.z, z.w};
      float sn[4], cs[4], abar[4];
      const float sv = s ? s[p] : 0.f;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        sincos_reduced(za[e], sn[e], cs[e]);
        abar[e] = cs[e] * (sv * wv[e]);
        dwl[e] = fmaf(sv, sn[e], dwl[e]);
      }
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const int64_t k = base + (int64_t)(1 + t) * W;
        const float gv = g ? g[p * NT + t] : 0.f;
        const float4 zt = ld4(Z, k);
        const float zv[4] = {zt.x, zt.y, zt.z, zt.w};
        float o[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const float pt = gv * wv[e];
          abar[e] = fmaf(-sn[e] * zv[e], pt, abar[e]);
          o[e] = cs[e] * pt;
          dwl[e] = fmaf(gv * cs[e], zv[e], dwl[e]);
        }
        st4(Abar, k, make_float4(o[0], o[1], o[2], o[3]));
        if constexpr (NT == 3) {
#pragma unroll
          for (int e = 0; e < 4; ++e) rows[pp][1 + t][e] = o[e];
        }
      }
      st4(Abar, base, make_float4(abar[0], abar[1], abar[2], abar[3]));
#pragma unroll
      for (int e = 0; e < 4; ++e) db[e] += abar[e];
#pragma unroll
      for (int e = 0; e < 4; ++e) rows[pp][0][e] = abar[e];
    }
    if constexpr (NT == 3) {
      if (Ab) st_blocked_p<T, P>(Ab, p0 + r, W, u, rows);
    }
  }
  float* o = part + (int64_t)blockIdx.x * 2 * W + u;
  *reinterpret_cast<float4*>(o) = make_float4(dwl[0], dwl[1], dwl[2], dwl[3]);
  *reinterpret_cast<float4*>(o + W) = make_float4(db[0], db[1], db[2], db[3]);
}

// Hidden-layer backward through sin / the tangent products (as
// k_train_act_bwd4) with the bias gradient reduced per block: blockDim =
// W / 4, db_part [blocks][W].  P: element type of Pbar.
template <class T, class PB, int NT, int P>
__global__ __launch_bounds__(128) void k_train_act_bwd_part4(const float* __restrict__ Z,
                                                             const PB* __restrict__ Pbar, int64_t n, int W,
                                                             T* __restrict__ Abar, T* __restrict__ Ab,
                                                             float* __restrict__ db_part) {
  const int u = threadIdx.x * 4;
  if (u >= W) return;
  float db[4] = {0.f, 0.f, 0.f, 0.f};
  const int64_t p0 = (int64_t)blockIdx.x * RED_ROWS;
  for (int r = 0; r < RED_ROWS; r += P) {
    if (p0 + r >= n) break;  // n is even when P = 2
    float rows[P][4][4];
#pragma unroll
    for (int pp = 0; pp < P; ++pp) {
      const int64_t p = p0 + r + pp;
      const int64_t base = p * (1 + NT) * W + u;
      const float4 z = ld4(Z, base), pb = ld4(Pbar, base);
      const float za[4] = {z.x, z.y, z.z, z.w}, pa[4] = {pb.x, pb.y, pb.z, pb.w};
      float sn[4], cs[4], abar[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        sincos_reduced(za[e], sn[e], cs[e]);
        abar[e] = cs[e] * pa[e];
      }
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const int64_t k = base + (int64_t)(1 + t) * W;
        const float4 zt = ld4(Z, k), pt = ld4(Pbar, k);
        const float zv[4] = {zt.x, zt.y, zt.z, zt.w}, pv[4] = {pt.x, pt.y, pt.z, pt.w};
        float o[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          abar[e] = fmaf(-sn[e] * zv[e], pv[e], abar[e]);
          o[e] = cs[e] * pv[e];
        }
        st4(Abar, k, make_float4(o[0], o[1], o[2], o[3]));
        if constexpr (NT == 3) {
#pragma unroll
          for (int e = 0; e < 4; ++e) rows[pp][1 + t][e] = o[e];
        }
      }
      st4(Abar, base, make_float4(abar[0], abar[1], abar[2], abar[3]));
#pragma unroll
      for (int e = 0; e < 4; ++e) db[e] += abar[e];
#pragma unroll
      for (int e = 0; e < 4; ++e) rows[pp][0][e] = abar[e];
    }
    if constexpr (NT == 3) {
      if (Ab) st_blocked_p<T, P>(Ab, p0 + r, W, u, rows);
    }
  }
  *reinterpret_cast<float4*>(db_part + (int64_t)blockIdx.x * W + u) = make_float4(db[0], db[1], db[2], db[3]);
}

// Layer-0 backward fused with its parameter gradients (W % 4 == 0,
// blockDim = W / 4): abar = cos a pbar - sin a sum_t ad_t pdbar_t and
// adbar_t = cos a pdbar_t are reduced in place against the layer input
// [x; v_t], giving per-block partial sums part [blocks][W][4] of
// (dW0[u][0..2], db0[u]).
template <class P, int NT>
__global__ __launch_bounds__(128) void k_train_first_bwd4(const float* __restrict__ Z, const P* __restrict__ Pbar,
                                                          const float* __restrict__ x, const float* __restrict__ v,
                                                          int64_t n, int W, float* __restrict__ part) {
  const int u = threadIdx.x * 4;
  if (u >= W) return;
  float acc[4][4];
#pragma unroll
  for (int e = 0; e < 4; ++e)
#pragma unroll
    for (int k = 0; k < 4; ++k) acc[e][k] = 0.f;
  const int64_t p0 = (int64_t)blockIdx.x * RED_ROWS;
  for (int r = 0; r < RED_ROWS; ++r) {
    const int64_t p = p0 + r;
    if (p >= n) break;
    const int64_t base = p * (1 + NT) * W + u;
    const float4 z = ld4(Z, base), pb = ld4(Pbar, base);
    const float za[4] = {z.x, z.y, z.z, z.w}, pa[4] = {pb.x, pb.y, pb.z, pb.w};
    float sn[4], cs[4], abar[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      sincos_reduced(za[e], sn[e], cs[e]);
      abar[e] = cs[e] * pa[e];
    }
    const float x0 = x[3 * p], x1 = x[3 * p + 1], x2 = x[3 * p + 2];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int64_t k = base + (int64_t)(1 + t) * W;
      const float4 zt = ld4(Z, k), pt = ld4(Pbar, k);
      const float zv[4] = {zt.x, zt.y, zt.z, zt.w}, pv[4] = {pt.x, pt.y, pt.z, pt.w};
      const float* vt = v + (p * NT + t) * 3;
      const float v0 = vt[0], v1 = vt[1], v2 = vt[2];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        abar[e] = fmaf(-sn[e] * zv[e], pv[e], abar[e]);
        const float ad = cs[e] * pv[e];
        acc[e][0] = fmaf(ad, v0, acc[e][0]);
        acc[e][1] = fmaf(ad, v1, acc[e][1]);
        acc[e][2] = fmaf(ad, v2, acc[e][2]);
      }
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      acc[e][0] = fmaf(abar[e], x0, acc[e][0]);
      acc[e][1] = fmaf(abar[e], x1, acc[e][1]);
      acc[e][2] = fmaf(abar[e], x2, acc[e][2]);
      acc[e][3] += abar[e];
    }
  }
  float* o = part + ((int64_t)blockIdx.x * W + u) * 4;
#pragma unroll
  for (int e = 0; e < 4; ++e)
    *reinterpret_cast<float4*>(o + 4 * e) = make_float4(acc[e][0], acc[e][1], acc[e][2], acc[e][3]);
}

}  // namespace mp

extern "C" {

int mp_train_out_fwd(float* Z, const float* b_sine, const float* w_last, const float* b_last, int64_t n, int W,
                     int n_tangents, float* f, float* fd, void* stream) {
  MP_REQUIRE(Z && w_last && b_last && f && n >= 0 && W > 0 && n_tangents >= 0 && n_tangents <= 3, MP_EINVAL,
             "mp_train_out_fwd: bad arguments");
  MP_REQUIRE(n_tangents == 0 || fd, MP_EINVAL, "mp_train_out_fwd: fd is NULL");
  if (n == 0) return MP_OK;
  hipStream_t st = as_stream(stream);
  if (W % 64 == 0) {
    const dim3 g((unsigned)((n + 15) / 16));
    switch (n_tangents) {
      case 0: hipLaunchKernelGGL(k_train_out_fwd16<0>, g, dim3(256), 0, st, Z, b_sine, w_last, b_last, n, W, f, fd); break;
      case 1: hipLaunchKernelGGL(k_train_out_fwd16<1>, g, dim3(256), 0, st, Z, b_sine, w_last, b_last, n, W, f, fd); break;
      case 2: hipLaunchKernelGGL(k_train_out_fwd16<2>, g, dim3(256), 0, st, Z, b_sine, w_last, b_last, n, W, f, fd); break;
      default: hipLaunchKernelGGL(k_train_out_fwd16<3>, g, dim3(256), 0, st, Z, b_sine, w_last, b_last, n, W, f, fd); break;
    }
    MP_LAUNCHED("k_train_out_fwd16");
    return MP_OK;
  }
  hipLaunchKernelGGL(k_train_out_fwd, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, st, Z, b_sine,
                     w_last, b_last, n, W, n_tangents, f, fd);
  MP_LAUNCHED("k_train_out_fwd");
  return MP_OK;
}

int64_t mp_train_out_bwd_blocks(int64_t n) { return (n + RED_ROWS - 1) / RED_ROWS; }

int mp_train_out_bwd(const float* Z, const float* s, const float* g, const float* w_last, int64_t n, int W,
                     int n_tangents, float* Pbar, float* dwl_part, void* stream) {
  MP_REQUIRE(Z && w_last && Pbar && dwl_part && n >= 0 && W > 0 && n_tangents >= 0 && n_tangents <= 3, MP_EINVAL,
             "mp_train_out_bwd: bad arguments");
  if (n == 0) return MP_OK;
  const int64_t blocks = mp_train_out_bwd_blocks(n);
  hipStream_t st = as_stream(stream);
  if (W % 4 == 0 && W <= 512) {
    switch (n_tangents) {
      case 0: hipLaunchKernelGGL(k_train_out_bwd4<0>, dim3((unsigned)blocks), dim3(W / 4), 0, st, Z, s, g, w_last, n, W, Pbar, dwl_part); break;
      case 1: hipLaunchKernelGGL(k_train_out_bwd4<1>, dim3((unsigned)blocks), dim3(W / 4), 0, st, Z, s, g, w_last, n, W, Pbar, dwl_part); break;
      case 2: hipLaunchKernelGGL(k_train_out_bwd4<2>, dim3((unsigned)blocks), dim3(W / 4), 0, st, Z, s, g, w_last, n, W, Pbar, dwl_part); break;
      default: hipLaunchKernelGGL(k_train_out_bwd4<3>, dim3((unsigned)blocks), dim3(W / 4), 0, st, Z, s, g, w_last, n, W, Pbar, dwl_part); break;
    }
    MP_LAUNCHED("k_train_out_bwd4");
    return MP_OK;
  }
  hipLaunchKernelGGL(k_train_out_bwd, dim3((unsigned)blocks), dim3(256), 0, st, Z, s, g, w_last, n,
                     W, n_tangents, RED_ROWS, Pbar, dwl_part);
  MP_LAUNCHED("k_train_out_bwd");
  return MP_OK;
}

// Launch K<T, NT, P> over `blocks` blocks of W / 4 threads.
#define MP_TRAIN_REDP_DISPATCH(KERNEL, T, P, nt, blocks, W, s, ...)                                                  \
  do {                                                                                                              \
    switch (nt) {                                                                                                   \
      case 0: hipLaunchKernelGGL((KERNEL<T, 0, P>), dim3((unsigned)(blocks)), dim3((W) / 4), 0, s, __VA_ARGS__); break; \
      case 1: hipLaunchKernelGGL((KERNEL<T, 1, P>), dim3((unsigned)(blocks)), dim3((W) / 4), 0, s, __VA_ARGS__); break; \
      case 2: hipLaunchKernelGGL((KERNEL<T, 2, P>), dim3((unsigned)(blocks)), dim3((W) / 4), 0, s, __VA_ARGS__); break; \
      default: hipLaunchKernelGGL((KERNEL<T, 3, P>), dim3((unsigned)(blocks)), dim3((W) / 4), 0, s, __VA_ARGS__); break; \
    }                                                                                                               \
  } while (0)

// Launch K<T, NT> over `blocks` blocks of W / 4 threads.
#define MP_TRAIN_RED_DISPATCH(KERNEL, T, nt, blocks, W, s, ...)                                              \
  do {                                                                                                      \
    switch (nt) {                                                                                           \
      case 0: hipLaunchKernelGGL((KERNEL<T, 0>), dim3((unsigned)(blocks)), dim3((W) / 4), 0, s, __VA_ARGS__); break; \
      case 1: hipLaunchKernelGGL((KERNEL<T, 1>), dim3((unsigned)(blocks)), dim3((W) / 4), 0, s, __VA_ARGS__); break; \
      case 2: hipLaunchKernelGGL((KERNEL<T, 2>), dim3((unsigned)(blocks)), dim3((W) / 4), 0, s, __VA_ARGS__); break; \
      default: hipLaunchKernelGGL((KERNEL<T, 3>), dim3((unsigned)(blocks)), dim3((W) / 4), 0, s, __VA_ARGS__); break; \
    }                                                                                                       \
  } while (0)

int mp_train_out_act_bwd(const float* Z, const float* s, const float* g, const float* w_last, int64_t n, int W,
                         int n_tangents, void* Abar, int a_dtype, void* Ab, float* part, void* stream) {
  MP_REQUIRE(Z && w_last && Abar && part && n >= 0 && W > 0 && n_tangents >= 0 && n_tangents <= 3, MP_EINVAL,
             "mp_train_out_act_bwd: bad arguments");
  MP_REQUIRE(a_dtype >= 0 && a_dtype <= 2, MP_EINVAL, "mp_train_out_act_bwd: a_dtype %d", a_dtype);
  MP_REQUIRE(W % 4 == 0 && W <= 512, MP_EUNSUPPORTED, "mp_train_out_act_bwd: width %d (needs a multiple of 4, <= 512)",
             W);
  MP_REQUIRE(!Ab || (n_tangents == 3 && n % 2 == 0 && a_dtype != 0), MP_EINVAL,
             "mp_train_out_act_bwd: the blocked copy needs 16-bit A, 3 tangents and even n");
  if (n == 0) return MP_OK;
  const int64_t blocks = mp_train_out_bwd_blocks(n);
  hipStream_t st = as_stream(stream);
  // with the blocked copy: two points per step (whole 16-byte chunks)
  if (a_dtype == 1 && Ab)
    MP_TRAIN_REDP_DISPATCH(k_train_out_act_bwd4, __bf16, 2, n_tangents, blocks, W, st, Z, s, g, w_last, n, W,
                           reinterpret_cast<__bf16*>(Abar), reinterpret_cast<__bf16*>(Ab), part);
  else if (a_dtype == 1)
    MP_TRAIN_REDP_DISPATCH(k_train_out_act_bwd4, __bf16, 1, n_tangents, blocks, W, st, Z, s, g, w_last, n, W,
                           reinterpret_cast<__bf16*>(Abar), static_cast<__bf16*>(nullptr), part);
  else if (a_dtype == 2 && Ab)
    MP_TRAIN_REDP_DISPATCH(k_train_out_act_bwd4, _Float16, 2, n_tangents, blocks, W, st, Z, s, g, w_last, n, W,
                           reinterpret_cast<_Float16*>(Abar), reinterpret_cast<_Float16*>(Ab), part);
  else if (a_dtype == 2)
    MP_TRAIN_REDP_DISPATCH(k_train_out_act_bwd4, _Float16, 1, n_tangents, blocks, W, st, Z, s, g, w_last, n, W,
                           reinterpret_cast<_Float16*>(Abar), static_cast<_Float16*>(nullptr), part);
  else
    MP_TRAIN_REDP_DISPATCH(k_train_out_act_bwd4, float, 1, n_tangents, blocks, W, st, Z, s, g, w_last, n, W,
                           reinterpret_cast<float*>(Abar), static_cast<float*>(nullptr), part);
  MP_LAUNCHED("k_train_out_act_bwd4");
  return MP_OK;
}

}  // extern "C"

namespace mp {

template <class T, class PB, int P>
static void launch_act_bwd_part_p(int nt, int64_t blocks, int W, hipStream_t st, const float* Z, const void* Pbar,
                                  int64_t n, void* Abar, void* Ab, float* db_part) {
  const PB* pb = reinterpret_cast<const PB*>(Pbar);
  T* a = reinterpret_cast<T*>(Abar);
  T* ab = reinterpret_cast<T*>(Ab);
  const dim3 g((unsigned)blocks), b(W / 4);
  switch (nt) {
    case 0: hipLaunchKernelGGL((k_train_act_bwd_part4<T, PB, 0, P>), g, b, 0, st, Z, pb, n, W, a, ab, db_part); break;
    case 1: hipLaunchKernelGGL((k_train_act_bwd_part4<T, PB, 1, P>), g, b, 0, st, Z, pb, n, W, a, ab, db_part); break;
    case 2: hipLaunchKernelGGL((k_train_act_bwd_part4<T, PB, 2, P>), g, b, 0, st, Z, pb, n, W, a, ab, db_part); break;
    default: hipLaunchKernelGGL((k_train_act_bwd_part4<T, PB, 3, P>), g, b, 0, st, Z, pb, n, W, a, ab, db_part); break;
  }
}

// with the blocked copy: two points per step (whole 16-byte chunks)
template <class T, class PB>
static void launch_act_bwd_part(int nt, int64_t blocks, int W, hipStream_t st, const float* Z, const void* Pbar,
                                int64_t n, void* Abar, void* Ab, float* db_part) {
  if (Ab && sizeof(T) == 2) launch_act_bwd_part_p<T, PB, 2>(nt, blocks, W, st, Z, Pbar, n, Abar, Ab, db_part);
  else launch_act_bwd_part_p<T, PB, 1>(nt, blocks, W, st, Z, Pbar, n, Abar, nullptr, db_part);
}

}  // namespace mp

extern "C" {

int mp_train_act_bwd_part(const float* Z, const void* Pbar, int p_dtype, int64_t n, int W, int n_tangents,
                          void* Abar, int a_dtype, void* Ab, float* db_part, void* stream) {
  MP_REQUIRE(Z && Pbar && Abar && db_part && n >= 0 && W > 0 && n_tangents >= 0 && n_tangents <= 3, MP_EINVAL,
             "mp_train_act_bwd_part: bad arguments");
  MP_REQUIRE(a_dtype >= 0 && a_dtype <= 2 && p_dtype >= 0 && p_dtype <= 2, MP_EINVAL,
             "mp_train_act_bwd_part: a_dtype %d, p_dtype %d", a_dtype, p_dtype);
  MP_REQUIRE(W % 4 == 0 && W <= 512, MP_EUNSUPPORTED,
             "mp_train_act_bwd_part: width %d (needs a multiple of 4, <= 512)", W);
  MP_REQUIRE(!Ab || (n_tangents == 3 && n % 2 == 0 && a_dtype != 0), MP_EINVAL,
             "mp_train_act_bwd_part: the blocked copy needs 16-bit A, 3 tangents and even n");
  if (n == 0) return MP_OK;
  const int64_t blocks = mp_train_out_bwd_blocks(n);
  hipStream_t st = as_stream(stream);
#define MP_ACT_BWD_P(T)                                                                                     \
  do {                                                                                                    \
    if (p_dtype == 1) launch_act_bwd_part<T, __bf16>(n_tangents, blocks, W, st, Z, Pbar, n, Abar, Ab, db_part);    \
    else if (p_dtype == 2) launch_act_bwd_part<T, _Float16>(n_tangents, blocks, W, st, Z, Pbar, n, Abar, Ab, db_part); \
    else launch_act_bwd_part<T, float>(n_tangents, blocks, W, st, Z, Pbar, n, Abar, Ab, db_part);          \
  } while (0)
  if (a_dtype == 1) MP_ACT_BWD_P(__bf16);
  else if (a_dtype == 2) MP_ACT_BWD_P(_Float16);
  else MP_ACT_BWD_P(float);
#undef MP_ACT_BWD_P
  MP_LAUNCHED("k_train_act_bwd_part4");
  return MP_OK;
}

int64_t mp_train_first_bwd_blocks(int64_t n) { return (n + RED_ROWS - 1) / RED_ROWS; }

}  // extern "C"

namespace mp {

template <class P>
static void launch_first_bwd(int nt, int64_t blocks, int W, hipStream_t st, const float* Z, const void* Pbar,
                             const float* x, const float* v, int64_t n, float* part) {
  const P* pb = reinterpret_cast<const P*>(Pbar);
  const dim3 g((unsigned)blocks), b(W / 4);
  switch (nt) {
    case 0: hipLaunchKernelGGL((k_train_first_bwd4<P, 0>), g, b, 0, st, Z, pb, x, v, n, W, part); break;
    case 1: hipLaunchKernelGGL((k_train_first_bwd4<P, 1>), g, b, 0, st, Z, pb, x, v, n, W, part); break;
    case 2: hipLaunchKernelGGL((k_train_first_bwd4<P, 2>), g, b, 0, st, Z, pb, x, v, n, W, part); break;
    default: hipLaunchKernelGGL((k_train_first_bwd4<P, 3>), g, b, 0, st, Z, pb, x, v, n, W, part); break;
  }
}

}  // namespace mp

extern "C" {

int mp_train_first_bwd(const float* Z, const void* Pbar, int p_dtype, const float* x, const float* v, int64_t n,
                       int W, int n_tangents, float* part, void* stream) {
  MP_REQUIRE(Z && Pbar && x && part && n >= 0 && W > 0 && n_tangents >= 0 && n_tangents <= 3, MP_EINVAL,
             "mp_train_first_bwd: bad arguments");
  MP_REQUIRE(p_dtype >= 0 && p_dtype <= 2, MP_EINVAL, "mp_train_first_bwd: p_dtype %d", p_dtype);
  MP_REQUIRE(n_tangents == 0 || v, MP_EINVAL, "mp_train_first_bwd: tangent directions are NULL");
  MP_REQUIRE(W % 4 == 0 && W <= 512, MP_EUNSUPPORTED, "mp_train_first_bwd: width %d (needs a multiple of 4, <= 512)", W);
  if (n == 0) return MP_OK;
  const int64_t blocks = mp_train_first_bwd_blocks(n);
  hipStream_t st = as_stream(stream);
  if (p_dtype == 1) launch_first_bwd<__bf16>(n_tangents, blocks, W, st, Z, Pbar, x, v, n, part);
  else if (p_dtype == 2) launch_first_bwd<_Float16>(n_tangents, blocks, W, st, Z, Pbar, x, v, n, part);
  else launch_first_bwd<float>(n_tangents, blocks, W, st, Z, Pbar, x, v, n, part);
  MP_LAUNCHED("k_train_first_bwd4");
  return MP_OK;
}

}  // extern "C"
