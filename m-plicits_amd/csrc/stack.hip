// Level-stack creation / packing, point evaluation and grid coordinates.
//
// mp_stack_create replaces i3d.util.from_pth(...).to(device) for the GPU path
// (i3d/util.py:79-142) and the renderer's NeuralImplicit::init
// (renderer/cuda/gemms.cu:111-134): the float32 weights stay where the caller
// keeps them (a torch state_dict on the device) and are packed into MFMA
// fragment order by the kernels below.
#include <stdlib.h>
#include <string.h>

#include <mutex>
#include <string>

#include "stack.h"
#include "eval_dispatch.h"

namespace mp {

static thread_local std::string g_error;

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_error = buf;
}
const char* get_error() { return g_error.c_str(); }

int default_nc(int shape, int width) {
  (void)shape;
  return width <= 256 ? 2 : 1;
}

// 1 / (2 pi): hidden layers are packed in revolutions for v_sin_f32.
static constexpr float INV_2PI = 0.159154943091895336f;
__host__ __device__ inline bool first_in_revs(int shape) { return shape_info(shape).elem_bytes == 2; }

int ensure_scratch(mp_stack* st, size_t bytes) {
  if (st->scratch_bytes >= bytes) return MP_OK;
  if (st->scratch) {
    // the previous contents are not needed; work on the stream is ordered
    // before the free by the device-wide synchronisation hipFree performs
    (void)hipFree(st->scratch);
    st->scratch = nullptr;
    st->scratch_bytes = 0;
  }
  const size_t b = (bytes + (1u << 20) - 1) & ~(size_t)((1u << 20) - 1);
  hipError_t e = hipMalloc(&st->scratch, b);
  if (e != hipSuccess) {
    set_error("scratch allocation of %zu bytes failed: %s", b, hipGetErrorString(e));
    return MP_ENOMEM;
  }
  st->scratch_bytes = b;
  return MP_OK;
}

// ---------------------------------------------------------------------------
// Packing kernels.  Layer k of the level has shape [out_k][in_k]; offsets into
// the borrowed weight / bias buffers follow the renderer .bin order.
// ---------------------------------------------------------------------------
struct PackArgs {
  const float* W;   // weights of the level
  const float* B;   // biases of the level
  int shape;
  int width;        // padded width
  int n_hidden;
  int dims[17];      // layer widths: dims[0] = 3, dims[k] = hidden[k-1], dims[H+1] = 1
  int64_t woff[17];                  // offset of W_k in W
  int64_t boff[17];                  // offset of b_k in B
  float4* w0v;
  float2* w0m;
  void* apk;
  void* apkt;
  float* biasc;
  float* wlc;
  float* wlv;
  float* blp;
};

__global__ void k_pack_small(PackArgs a) {
  const ShapeInfo s = shape_info(a.shape);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int H = a.n_hidden;
  if (i < a.width) {
    // layer 0 in virtual order
    const int u = shape_perm(s, i);
    float4 w = make_float4(0.f, 0.f, 0.f, 0.f);
    if (u < a.dims[1]) {
      const float* row = a.W + a.woff[0] + (int64_t)u * 3;
      w = make_float4(row[0], row[1], row[2], a.B[a.boff[0] + u]);
      if (first_in_revs(a.shape)) w = make_float4(w.x * INV_2PI, w.y * INV_2PI, w.z * INV_2PI, w.w * INV_2PI);
    }
    a.w0v[i] = w;
    // output layer, virtual order
    a.wlv[i] = (u < a.dims[H]) ? a.W[a.woff[H] + u] : 0.f;
    // output layer, C order: i = (rb * GROUPS + g) * CREGS + q
    const int q = i % s.cregs, g = (i / s.cregs) % s.groups, rb = i / (s.cregs * s.groups);
    const int r = rb * s.rows + shape_rowmap(s, g, q);
    a.wlc[i] = (r < a.dims[H]) ? a.W[a.woff[H] + r] : 0.f;
    // hidden biases, C order
    for (int j = 1; j < H; ++j)
      a.biasc[(int64_t)(j - 1) * a.width + i] = (r < a.dims[j + 1]) ? a.B[a.boff[j] + r] * INV_2PI : 0.f;
  }
  // layer 0 as the A operands of two 32x32x2 FP32 MFMAs per 32-row block:
  // lane l holds row r = 32 rb + l % 32 and reduction indices (g, 2 + g),
  // g = l / 32, over the operand (x, y, z, 1): (w_x | w_y, w_z | b).
  if (i < a.width * 2) {
    const int rb = i / 64, lane = i % 64, r = rb * 32 + lane % 32, g = lane / 32;
    float2 m = make_float2(0.f, 0.f);
    if (r < a.dims[1]) {
      const float* row = a.W + a.woff[0] + (int64_t)r * 3;
      m = make_float2(row[g], g == 0 ? row[2] : a.B[a.boff[0] + r]);
      if (first_in_revs(a.shape)) m = make_float2(m.x * INV_2PI, m.y * INV_2PI);
    }
    a.w0m[i] = m;
  }
  if (i == 0) *a.blp = a.B[a.boff[H]];
}

// Hidden layer j (1..H-1): element ((rb * KBS + kb) * 64 + lane) * EPL + e.
// transpose = 0: A = W_j / (2 pi) (forward, revolutions), rows = layer-j units.
// transpose = 1: A = W_j^T (backward, radians), rows = layer-(j-1) units.
// Three-term BF16, forward: fragment 2 (rb * KBS + kb) holds bf16(w), the
// next one bf16(w - bf16(w)); the transposed copy holds bf16(w) only.
__global__ void k_pack_hidden(PackArgs a, int j, int transpose) {
  const ShapeInfo s = shape_info(a.shape);
  const int64_t per_layer = (int64_t)a.width * a.width;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= per_layer) return;
  const int e = (int)(i % s.epl);
  const int lane = (int)((i / s.epl) % 64);
  const int64_t blk = i / (s.epl * 64);
  const int kbs = a.width / s.kb;
  const int kb = (int)(blk % kbs), rb = (int)(blk / kbs);
  const int r = rb * s.rows + lane % s.rows;
  const int v = kb * s.kb + (lane / s.rows) * s.epl + e;
  const int u = shape_perm(s, v);
  const int out_j = a.dims[j + 1], in_j = a.dims[j];
  float w;
  if (!transpose) w = (r < out_j && u < in_j) ? a.W[a.woff[j] + (int64_t)r * in_j + u] * INV_2PI : 0.f;
  else w = (u < out_j && r < in_j) ? a.W[a.woff[j] + (int64_t)u * in_j + r] : 0.f;
  const int64_t o = (int64_t)(j - 1) * per_layer + i;
  void* dst = transpose ? a.apkt : a.apk;
  if (a.shape == SH_32x32x16_BF16X3F && !transpose && j == a.n_hidden - 1) {
    // the last hidden product of the mixed shape: one FP16 fragment per step,
    // at the layer's offset in the pair layout
    const int64_t oh = ((int64_t)(j - 1) * per_layer * 2) + (blk * 64 + lane) * s.epl + e;
    reinterpret_cast<_Float16*>(dst)[oh] = (_Float16)w;
    return;
  }
  if ((a.shape == SH_32x32x16_BF16X3 || a.shape == SH_32x32x16_BF16X3F) && !transpose) {
    const __bf16 hi = (__bf16)w;
    const int64_t oh = ((int64_t)(j - 1) * per_layer * 2) + ((blk * 2) * 64 + lane) * s.epl + e;
    reinterpret_cast<__bf16*>(dst)[oh] = hi;
    reinterpret_cast<__bf16*>(dst)[oh + 64 * s.epl] = (__bf16)(w - (float)hi);
    return;
  }
  switch (a.shape) {
    case SH_32x32x16_BF16X3: case SH_32x32x16_BF16X3F:
    case SH_32x32x16_BF16: case SH_16x16x32_BF16: reinterpret_cast<__bf16*>(dst)[o] = (__bf16)w; break;
    case SH_32x32x16_F16: case SH_16x16x32_F16: reinterpret_cast<_Float16*>(dst)[o] = (_Float16)w; break;
    default: reinterpret_cast<float*>(dst)[o] = w; break;
  }
}

static int pack_level(const mp_stack* st, const mp_stack::Level& L, hipStream_t stream) {
  PackArgs a{};
  a.W = L.weights;
  a.B = L.biases;
  a.shape = L.shape;
  a.width = L.width;
  a.n_hidden = L.n_hidden;
  a.dims[0] = 3;
  for (int k = 0; k < L.n_hidden; ++k) a.dims[k + 1] = L.hidden[k];
  a.dims[L.n_hidden + 1] = 1;
  int64_t wo = 0, bo = 0;
  for (int k = 0; k <= L.n_hidden; ++k) {
    a.woff[k] = wo;
    a.boff[k] = bo;
    wo += (int64_t)a.dims[k + 1] * a.dims[k];
    bo += a.dims[k + 1];
  }
  a.w0v = const_cast<float4*>(L.dev.w0v);
  a.w0m = const_cast<float2*>(L.dev.w0m);
  a.apk = const_cast<void*>(L.dev.apk);
  a.apkt = const_cast<void*>(L.dev.apkt);
  a.biasc = const_cast<float*>(L.dev.biasc);
  a.wlc = const_cast<float*>(L.dev.wlc);
  a.wlv = const_cast<float*>(L.dev.wlv);
  a.blp = const_cast<float*>(L.dev.blp);
  hipLaunchKernelGGL(k_pack_small, dim3((2 * L.width + 255) / 256), dim3(256), 0, stream, a);
  MP_LAUNCHED("k_pack_small");
  const int64_t per_layer = (int64_t)L.width * L.width;
  for (int j = 1; j < L.n_hidden; ++j) {
    for (int tr = 0; tr < 2; ++tr) {
      hipLaunchKernelGGL(k_pack_hidden, dim3((unsigned)((per_layer + 255) / 256)), dim3(256), 0,
                         stream, a, j, tr);
      MP_LAUNCHED("k_pack_hidden");
    }
  }
  return MP_OK;
}

}  // namespace mp

using namespace mp;

extern "C" {

int mp_abi_version(void) { return MP_ABI_VERSION; }

const char* mp_last_error(void) { return mp::get_error(); }

int mp_stack_levels(const mp_stack* stack) { return stack ? stack->n_levels : 0; }

void mp_stack_destroy(mp_stack* st) {
  if (!st) return;
  for (int l = 0; l < st->n_levels; ++l)
    if (st->lv[l].buf) (void)hipFree(st->lv[l].buf);
  if (st->scratch) (void)hipFree(st->scratch);
  if (st->train_ws) (void)hipFree(st->train_ws);
  delete st;
}

int mp_stack_create(const mp_level_desc* levels, int n_levels, int precision, void* stream,
                    mp_stack** out) {
  MP_REQUIRE(out != nullptr, MP_EINVAL, "mp_stack_create: out is NULL");
  *out = nullptr;
  MP_REQUIRE(levels != nullptr && n_levels >= 1 && n_levels <= MP_MAX_LEVELS, MP_EINVAL,
             "mp_stack_create: n_levels must be in [1, %d], got %d", MP_MAX_LEVELS, n_levels);
  MP_REQUIRE(precision == MP_FP32 || precision == MP_BF16 || precision == MP_FP16 || precision == MP_BF16X3 ||
                 precision == MP_BF16X3F,
             MP_EINVAL, "mp_stack_create: unknown precision %d", precision);
  mp_stack* st = new mp_stack();
  st->n_levels = n_levels;
  st->precision = precision;
  st->shape = precision == MP_BF16 ? SH_32x32x16_BF16
              : precision == MP_FP16 ? SH_32x32x16_F16
              : precision == MP_BF16X3 ? SH_32x32x16_BF16X3
              : precision == MP_BF16X3F ? SH_32x32x16_BF16X3F
                                     : SH_16x16x4_F32;
  int maxw = 0;
  for (int l = 0; l < n_levels; ++l) {
    const mp_level_desc& d = levels[l];
    mp_stack::Level& L = st->lv[l];
    if (d.n_in != 3 || d.n_out != 1 || d.n_hidden < 1 || d.n_hidden > 15 || !d.hidden ||
        !d.weights || !d.biases) {
      mp::set_error("mp_stack_create: level %d: need n_in=3, n_out=1, 1<=n_hidden<=15 and "
                    "non-NULL hidden/weights/biases (got n_in=%d n_out=%d n_hidden=%d)",
                    l, d.n_in, d.n_out, d.n_hidden);
      mp_stack_destroy(st);
      return MP_EINVAL;
    }
    int w = 0;
    for (int k = 0; k < d.n_hidden; ++k) {
      if (d.hidden[k] < 1) {
        mp::set_error("mp_stack_create: level %d: hidden width %d is < 1", l, d.hidden[k]);
        mp_stack_destroy(st);
        return MP_EINVAL;
      }
      w = d.hidden[k] > w ? d.hidden[k] : w;
      L.hidden.push_back(d.hidden[k]);
    }
    L.width = padded_width(w);
    if (L.width < 0) {
      mp::set_error("mp_stack_create: level %d: width %d exceeds the supported maximum 512", l, w);
      mp_stack_destroy(st);
      return MP_EUNSUPPORTED;
    }
    maxw = L.width > maxw ? L.width : maxw;
    L.n_hidden = d.n_hidden;
    // three-term BF16 up to width 256; wider levels FP16; the mixed shape
    // (FP16 last product) for levels with two hidden products, the others
    // three-term throughout
    const bool split = st->shape == SH_32x32x16_BF16X3 || st->shape == SH_32x32x16_BF16X3F;
    L.shape = (split && L.width > 256) ? SH_32x32x16_F16
              : (st->shape == SH_32x32x16_BF16X3F && d.n_hidden != 3) ? SH_32x32x16_BF16X3
                                                                       : st->shape;
    const ShapeInfo si = shape_info(L.shape);
    const size_t parts = (L.shape == SH_32x32x16_BF16X3 || L.shape == SH_32x32x16_BF16X3F) ? 2 : 1;
    L.weights = d.weights;
    L.biases = d.biases;
    const size_t W = (size_t)L.width;
    const size_t n_gemm = (size_t)(L.n_hidden - 1);
    const size_t sz_w0v = W * sizeof(float4);
    const size_t sz_apk = n_gemm * W * W * si.elem_bytes * parts;
    const size_t sz_bias = n_gemm * W * sizeof(float);
    const size_t sz_wl = W * sizeof(float);
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    // the level kernels prefetch A fragments up to APK_PAD bytes past the last layer
    const size_t total = 2 * al(sz_w0v) + al(sz_apk + APK_PAD) + al(sz_apk) + al(sz_bias) + 2 * al(sz_wl) + 256;
    hipError_t e = hipMalloc(&L.buf, total);
    if (e != hipSuccess) {
      mp::set_error("mp_stack_create: hipMalloc(%zu) failed: %s", total, hipGetErrorString(e));
      mp_stack_destroy(st);
      return MP_ENOMEM;
    }
    L.buf_bytes = total;
    char* p = reinterpret_cast<char*>(L.buf);
    L.dev.width = L.width;
    L.dev.n_hidden = L.n_hidden;
    L.dev.w0v = reinterpret_cast<const float4*>(p); p += al(sz_w0v);
    L.dev.w0m = reinterpret_cast<const float2*>(p); p += al(sz_w0v);
    L.dev.apk = p; p += al(sz_apk + APK_PAD);
    L.dev.apkt = p; p += al(sz_apk);
    L.dev.biasc = reinterpret_cast<const float*>(p); p += al(sz_bias);
    L.dev.wlc = reinterpret_cast<const float*>(p); p += al(sz_wl);
    L.dev.wlv = reinterpret_cast<const float*>(p); p += al(sz_wl);
    L.dev.blp = reinterpret_cast<const float*>(p);
  }
  // two column blocks per wave for the 16-bit shapes; launch_level drops to
  // one for levels wider than 256 (their activations would not fit the VGPRs)
  st->nc = default_nc(st->shape, 0);
  (void)maxw;
  int rc = mp_stack_update(st, stream);
  if (rc != MP_OK) {
    mp_stack_destroy(st);
    return rc;
  }
  *out = st;
  return MP_OK;
}

int mp_stack_set_tiles(mp_stack* st, int column_blocks) {
  MP_REQUIRE(st != nullptr, MP_EINVAL, "mp_stack_set_tiles: stack is NULL");
  MP_REQUIRE(column_blocks == 1 || column_blocks == 2, MP_EINVAL,
             "mp_stack_set_tiles: column_blocks must be 1 or 2, got %d", column_blocks);
  st->nc = column_blocks;
  return MP_OK;
}

int mp_stack_set_kernel_tiles(mp_stack* st, int family, int tiles) {
  MP_REQUIRE(st != nullptr, MP_EINVAL, "mp_stack_set_kernel_tiles: stack is NULL");
  switch (family) {
    case MP_TILES_EXTRACT:
      return mp_stack_set_tiles(st, tiles);
    case MP_TILES_BATCH:
      MP_REQUIRE(tiles == 1 || tiles == 2, MP_EINVAL,
                 "mp_stack_set_kernel_tiles: batched tiles must be 1 (default) or 2 (half), got %d", tiles);
      st->batch_tiles = tiles;
      return MP_OK;
    case MP_TILES_TRACE:
      MP_REQUIRE(tiles >= 0 && tiles <= 2, MP_EINVAL,
                 "mp_stack_set_kernel_tiles: tracing tiles must be 0, 1 or 2, got %d", tiles);
      st->trace_tiles = tiles;
      return MP_OK;
  }
  set_error("mp_stack_set_kernel_tiles: unknown kernel family %d", family);
  return MP_EINVAL;
}

int mp_stack_update(mp_stack* st, void* stream) {
  MP_REQUIRE(st != nullptr, MP_EINVAL, "mp_stack_update: stack is NULL");
  ++st->version;  // drops the activations kept for training
  for (int l = 0; l < st->n_levels; ++l) {
    int rc = pack_level(st, st->lv[l], as_stream(stream));
    if (rc != MP_OK) return rc;
  }
  return MP_OK;
}

int mp_eval(const mp_stack* st, const float* xyz, int64_t n, int lvl0, int lvl1, float* sdf,
            float* grad, void* stream) {
  MP_REQUIRE(st != nullptr, MP_EINVAL, "mp_eval: stack is NULL");
  MP_REQUIRE(n >= 0, MP_EINVAL, "mp_eval: n must be >= 0, got %lld", (long long)n);
  MP_REQUIRE(0 <= lvl0 && lvl0 < lvl1 && lvl1 <= st->n_levels, MP_EINVAL,
             "mp_eval: need 0 <= lvl0 < lvl1 <= %d, got [%d, %d)", st->n_levels, lvl0, lvl1);
  if (n == 0) return MP_OK;
  MP_REQUIRE(xyz != nullptr && sdf != nullptr, MP_EINVAL, "mp_eval: xyz and sdf must be non-NULL");
  // with gradients one pass per level gives both
  if (grad) return launch_grad(st, lvl0, lvl1, xyz, n, sdf, grad, as_stream(stream));
  for (int l = lvl0; l < lvl1; ++l) {
    EvalTarget t{};
    t.src = SRC_POINTS;
    t.xyz = xyz;
    t.n = n;
    t.out = sdf;
    t.accumulate = l > lvl0;
    int rc = launch_level(st, l, t, as_stream(stream));
    if (rc != MP_OK) return rc;
  }
  return MP_OK;
}

}  // extern "C"
