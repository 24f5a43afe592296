// The renderer's fixed-iteration frame: every pixel takes the same number of
// sphere-tracing steps per phase, then the pixels whose shading distance is
// below the threshold are shaded with the analytic normal of the selected
// levels.  Replaces MIPplicit::SdfInference (renderer/cuda/state.cu:816-1055),
// which runs the CUTLASS GEMM chain of every level on all pixels each step
// (IterateSphereTracing_kernel, renderer/cuda/kernels.cu:315-398), computes
// the normals with per-coordinate gradient networks (sum_residual_normals,
// kernels.cu:65-120) and shades with Shade_kernel (kernels.cu:122-275).
//
// Here one wave owns a tile of NC * ROWS pixels for the whole schedule, its
// positions in registers and every level evaluated by the fused LevelCore.
// Phase 0 (level 0, step f0 - delta), phase 1 (levels 0 and 1) and phase 2
// (levels 0, 1 for residual models, and 2) follow the reference's step rule;
// phases 1 and 2 step by the residual sum (standalone models: by the finer
// level) where f0 < band_scale * delta and by f0 elsewhere.  A finer level is
// evaluated on a step only when some pixel of the tile is inside that band, or
// when the step is the one whose output the shading distance reads: the
// reference evaluates it on every step but reads it nowhere else, so the
// frame is the same.  The shading distance is the last evaluated output
// (taken before the last step), as in the reference.
//
// No host synchronisation and no allocation after the first call with a
// given size: a frame can be captured in a hipGraph.
#include "trace_core.h"

namespace mp {

struct RenderArgs {
  const float* cams;  // [n_cams][32]: inverse view, inverse projection (column-major)
  int width, height;
  int64_t total_rays;
  int iters[3];
  float delta, lim;   // band of level 0 and band_scale * delta
  float bound_radius;
  int residual, skip_lod0, lod2_alias, flip_y;
  int last1_phase;    // phase of the last level-1 evaluation (0: none)
  int use1, use2;     // residual models: terms of the shading distance
  int dist_src;       // standalone models: level whose output is the shading distance
  LevelDev levels[3];
  float* pos;         // [rays][3] final positions
  float* dist;        // [rays] shading distance
  unsigned long long* ctr;  // [0..2] level evaluations, [7] shaded pixels
};

__global__ void k_render_reset(unsigned long long* ctr) {
  if (threadIdx.x < 8) ctr[threadIdx.x] = 0ull;
}

template <class S, int NC>
__global__ __launch_bounds__(256) void k_render(RenderArgs a) {
  extern __shared__ uint4 lds_raw[];
  typename S::Frag* hand =
      reinterpret_cast<typename S::Frag*>(reinterpret_cast<char*>(lds_raw) + (threadIdx.x >> 6) * TRACE_HAND_BYTES);
  const int lane = threadIdx.x & 63;
  const int col = lane % S::ROWS;
  const int64_t tile = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * (NC * S::ROWS);
  if (tile >= a.total_rays) return;  // the whole wave: no barriers below
  int64_t rid[NC];
  bool act[NC];
  float p[NC][3], d[NC][3];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    rid[c] = tile + c * S::ROWS + col;
    act[c] = rid[c] < a.total_rays;
    float t0;
    bool inside;
    if (act[c]) {
      // a ray that never meets the bounding sphere starts at the eye
      ray_gen(a.cams, a.width, a.height, 0, a.flip_y, a.bound_radius, rid[c], p[c], d[c], t0, inside);
    } else {
      p[c][0] = p[c][1] = p[c][2] = 0.f;
      d[c][0] = d[c][1] = d[c][2] = 0.f;
    }
  }
  const bool counter = lane < S::ROWS;
  unsigned long long ev[3] = {0, 0, 0};
  int n_act = 0;
#pragma unroll
  for (int c = 0; c < NC; ++c) n_act += act[c] ? 1 : 0;
  float o0[NC], o1[NC], o2[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) o0[c] = o1[c] = o2[c] = 0.f;
  // The phases as one loop nest with a single evaluation call site (one
  // inlined copy of each width's code).
  for (int ph = 0; ph < 3; ++ph) {
    const int n_it = (ph == 0 && a.skip_lod0) ? 0 : a.iters[ph];
    for (int it = 0; it < n_it; ++it) {
      const bool last = it == n_it - 1;
      bool near = false;
      for (int l = 0; l <= ph; ++l) {
        bool run = l == 0;
        if (l == 1) run = ph == 1 ? (near || (a.last1_phase == 1 && last)) : (a.residual && (near || last));
        if (l == 2) run = !a.lod2_alias && (near || last);
        if (!run) continue;
        float r[NC];
        eval_level_any<S, NC>(a.levels[l], p, r, hand);
        if (counter) ev[l] += n_act;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          if (l == 0) o0[c] = r[c];
          else if (l == 1) o1[c] = r[c];
          else o2[c] = r[c];
        }
        if (l == 0) {
          bool any = false;
#pragma unroll
          for (int c = 0; c < NC; ++c) any |= act[c] && o0[c] < a.lim;
          near = __ballot(any) != 0;
        }
      }
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        float s;
        if (ph == 0) {
          s = o0[c] - a.delta;
        } else {
          const float fine = ph == 1 ? (a.residual ? o0[c] + o1[c] : o1[c])
                                     : (a.residual ? o0[c] + o1[c] + (a.lod2_alias ? o0[c] : o2[c])
                                                   : (a.lod2_alias ? o0[c] : o2[c]));
          s = o0[c] < a.lim ? fine : o0[c];
        }
        p[c][0] += s * d[c][0];
        p[c][1] += s * d[c][1];
        p[c][2] += s * d[c][2];
      }
    }
  }
  if (counter) {
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      if (!act[c]) continue;
      const float r2 = a.lod2_alias ? o0[c] : o2[c];
      float s;
      if (a.residual) {
        s = o0[c];
        if (a.use1) s += o1[c];
        if (a.use2) s += r2;
      } else {
        s = a.dist_src == 2 ? r2 : (a.dist_src == 1 ? o1[c] : o0[c]);
      }
      const int64_t id = rid[c];
      a.pos[3 * id + 0] = p[c][0];
      a.pos[3 * id + 1] = p[c][1];
      a.pos[3 * id + 2] = p[c][2];
      a.dist[id] = s;
    }
    for (int l = 0; l < 3; ++l)
      if (ev[l]) atomicAdd(&a.ctr[l], ev[l]);
  }
}

// White background for every pixel; the pixels that pass Shade_kernel's test
// (|dist| < threshold and the position inside [-1, 1]^3) compacted into
// ids / cpos.
__global__ void k_render_collect(const float* __restrict__ pos, const float* __restrict__ dist, float threshold,
                                 int64_t n, uint32_t* __restrict__ rgba, float* __restrict__ normals,
                                 uint32_t* __restrict__ ids, float* __restrict__ cpos,
                                 unsigned long long* __restrict__ count) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  bool shade = false;
  float x = 0.f, y = 0.f, z = 0.f;
  if (i < n) {
    rgba[i] = 0xFFFFFFFFu;
    if (normals) {
      normals[3 * i] = 0.f; normals[3 * i + 1] = 0.f; normals[3 * i + 2] = 0.f;
    }
    x = pos[3 * i]; y = pos[3 * i + 1]; z = pos[3 * i + 2];
    shade = fabsf(dist[i]) < threshold && fabsf(x) < 1.f && fabsf(y) < 1.f && fabsf(z) < 1.f;
  }
  const uint64_t m = __ballot(shade);
  if (!m) return;
  unsigned long long base = 0;
  if (lane == 0) base = atomicAdd(count, (unsigned long long)__popcll(m));
  base = __shfl(base, 0);
  if (shade) {
    const uint64_t below = (lane == 0) ? 0ull : (m & ((~0ull) >> (64 - lane)));
    const unsigned long long k = base + __popcll(below);
    ids[k] = (uint32_t)i;
    cpos[3 * k + 0] = x;
    cpos[3 * k + 1] = y;
    cpos[3 * k + 2] = z;
  }
}

// RgbaFloatToInt (kernels.cu:28-39): clamp, scale by 255, truncate.
__device__ inline uint32_t rgba8_trunc(float r, float g, float b) {
  auto q = [](float v) { return (uint32_t)(fminf(fmaxf(v, 0.f), 1.f) * 255.f); };
  return q(r) | (q(g) << 8) | (q(b) << 16) | (255u << 24);
}

// Shade_kernel's colours at the compacted pixels: NORMALS (|n_x|, |n_z|,
// |n_y|) or PHONG with the light and the viewer at the eye (ambient 0.2,
// diffuse 0.54, specular 0.5 with exponent 35).
__global__ void k_render_shade(const float* __restrict__ cams, int width, int height, int flip_y, float R,
                               int shading, const uint32_t* __restrict__ ids, const float* __restrict__ grad,
                               const unsigned long long* __restrict__ count, uint32_t* __restrict__ rgba,
                               float* __restrict__ normals) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (int64_t)*count) return;
  const int64_t id = ids[k];
  float nx = grad[3 * k], ny = grad[3 * k + 1], nz = grad[3 * k + 2];
  const float len = sqrtf(nx * nx + ny * ny + nz * nz);
  const float inv = len > 0.f ? 1.0f / len : 0.f;
  nx *= inv; ny *= inv; nz *= inv;
  uint32_t c;
  if (shading == MP_SHADE_NORMALS) {
    c = rgba8_trunc(fabsf(nx), fabsf(nz), fabsf(ny));
  } else {
    float o[3], d[3], t0;
    bool inside;
    ray_gen(cams, width, height, 0, flip_y, R, id, o, d, t0, inside);
    const float lx = -d[0], ly = -d[1], lz = -d[2];
    const float dot_d = lx * nx + ly * ny + lz * nz;
    const float rx = 2.f * dot_d * nx - lx, ry = 2.f * dot_d * ny - ly, rz = 2.f * dot_d * nz - lz;
    const float ls = lx * rx + ly * ry + lz * rz;
    const float dot_s = ls > 0.f ? powf(ls, 35.f) : 0.f;
    const float v = 0.2f + fmaxf(dot_d, 0.f) * 0.54f + 0.5f * dot_s;
    c = rgba8_trunc(v, v, v);
  }
  rgba[id] = c;
  if (normals) {
    normals[3 * id] = nx; normals[3 * id + 1] = ny; normals[3 * id + 2] = nz;
  }
}

template <class S, int NC>
static int launch_render(const RenderArgs& a, hipStream_t s) {
  static std::atomic<uint64_t> attr_devs{0};
  const size_t lds = 4 * TRACE_HAND_BYTES;
  MP_HIP(set_lds_once(attr_devs, reinterpret_cast<const void*>(&k_render<S, NC>), (int)lds));
  const int64_t per_block = 4 * NC * S::ROWS;
  hipLaunchKernelGGL((k_render<S, NC>), dim3((unsigned)((a.total_rays + per_block - 1) / per_block)), dim3(256), lds,
                     s, a);
  MP_LAUNCHED("k_render");
  return MP_OK;
}

}  // namespace mp

using namespace mp;

extern "C" {

int mp_render(mp_stack* st, const float* cams, int n_cams, const mp_render_params* prm, uint32_t* rgba, float* pos,
              float* dist, float* normals, int64_t* stats, void* stream) {
  MP_REQUIRE(st != nullptr && cams != nullptr && prm != nullptr && rgba != nullptr, MP_EINVAL,
             "mp_render: NULL argument");
  MP_REQUIRE(n_cams >= 1 && prm->width >= 1 && prm->height >= 1, MP_EINVAL, "mp_render: empty image");
  MP_REQUIRE(st->n_levels <= 3, MP_EUNSUPPORTED, "mp_render: the renderer draws at most three levels, got %d",
             st->n_levels);
  for (int l = 0; l < 3; ++l)
    MP_REQUIRE(prm->iters[l] >= 0, MP_EINVAL, "mp_render: negative iteration count for level %d", l);
  MP_REQUIRE(prm->iters[1] == 0 || st->n_levels >= 2, MP_EINVAL, "mp_render: level-1 iterations need 2 levels");
  MP_REQUIRE(prm->iters[2] == 0 || st->n_levels >= 3, MP_EINVAL, "mp_render: level-2 iterations need 3 levels");
  MP_REQUIRE(prm->normal_lod >= 0 && prm->normal_lod < st->n_levels, MP_EINVAL,
             "mp_render: normal_lod %d outside 0..%d", prm->normal_lod, st->n_levels - 1);
  MP_REQUIRE(prm->shading == MP_SHADE_NORMALS || prm->shading == MP_SHADE_PHONG, MP_EINVAL,
             "mp_render: unknown shading %d", prm->shading);
  MP_REQUIRE(prm->bound_radius > 0.f && prm->threshold >= 0.f, MP_EINVAL, "mp_render: bad parameters");
  MP_REQUIRE(st->precision != MP_BF16X3F, MP_EUNSUPPORTED,
             "mp_render: MP_BF16X3F (FP16 last product) serves extraction and values only; use MP_BF16X3");
  hipStream_t s = as_stream(stream);
  const int64_t total = (int64_t)n_cams * prm->width * prm->height;
  MP_REQUIRE(total < (1ll << 32), MP_EUNSUPPORTED, "mp_render: too many rays");

  // scratch: counters | pos [total][3] | dist [total] | ids [total] | cpos [total][3] | grad [total][3]
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t off_pos = 256;
  const size_t off_dist = off_pos + al((size_t)total * 12);
  const size_t off_ids = off_dist + al((size_t)total * 4);
  const size_t off_cpos = off_ids + al((size_t)total * 4);
  const size_t off_grad = off_cpos + al((size_t)total * 12);
  int rc = ensure_scratch(st, off_grad + (size_t)total * 12);
  if (rc != MP_OK) return rc;
  char* base = reinterpret_cast<char*>(st->scratch);
  unsigned long long* ctr = reinterpret_cast<unsigned long long*>(base);
  float* pos_buf = pos ? pos : reinterpret_cast<float*>(base + off_pos);
  float* dist_buf = dist ? dist : reinterpret_cast<float*>(base + off_dist);
  uint32_t* ids = reinterpret_cast<uint32_t*>(base + off_ids);
  float* cpos = reinterpret_cast<float*>(base + off_cpos);
  float* grad = reinterpret_cast<float*>(base + off_grad);

  RenderArgs a{};
  a.cams = cams;
  a.width = prm->width;
  a.height = prm->height;
  a.total_rays = total;
  for (int l = 0; l < 3; ++l) a.iters[l] = prm->iters[l];
  a.delta = prm->delta;
  a.lim = prm->band_scale * prm->delta;
  a.bound_radius = prm->bound_radius;
  a.residual = prm->residual != 0;
  a.skip_lod0 = prm->skip_lod0 != 0;
  a.lod2_alias = prm->lod2_reads_lod0 != 0;
  a.flip_y = prm->flip_y != 0;
  // state.cu:936-1054: which outputs make up the shading distance
  a.use1 = prm->iters[1] > 0 || prm->iters[2] > 0;
  a.use2 = prm->iters[2] > 0;
  a.dist_src = prm->iters[2] > 0 ? 2 : (prm->iters[1] > 0 ? 1 : 0);
  a.last1_phase = (a.residual && prm->iters[2] > 0) ? 2 : (prm->iters[1] > 0 ? 1 : 0);
  int maxw = 0;
  for (int l = 0; l < st->n_levels; ++l) {
    a.levels[l] = st->lv[l].dev;
    maxw = st->lv[l].width > maxw ? st->lv[l].width : maxw;
  }
  a.pos = pos_buf;
  a.dist = dist_buf;
  a.ctr = ctr;

  hipLaunchKernelGGL(k_render_reset, dim3(1), dim3(64), 0, s, ctr);
  MP_LAUNCHED("k_render_reset");
  const int nc = (st->nc == 2 && maxw <= 256 && st->shape != SH_32x32x16_BF16X3) ? 2 : 1;
  switch (st->shape) {
    case SH_32x32x16_BF16:
      rc = nc == 2 ? launch_render<S32x32x16_bf16, 2>(a, s) : launch_render<S32x32x16_bf16, 1>(a, s);
      break;
    case SH_32x32x16_F16:
      rc = nc == 2 ? launch_render<S32x32x16_f16, 2>(a, s) : launch_render<S32x32x16_f16, 1>(a, s);
      break;
    case SH_16x16x4_F32:
      rc = launch_render<S16x16x4_f32, 1>(a, s);
      break;
    case SH_32x32x16_BF16X3:
      rc = launch_render<S32x32x16_bf16x3, 1>(a, s);
      break;
    default:
      set_error("mp_render: unsupported fragment shape %d", st->shape);
      return MP_EUNSUPPORTED;
  }
  if (rc != MP_OK) return rc;
  const unsigned grid = (unsigned)((total + 255) / 256);
  hipLaunchKernelGGL(k_render_collect, dim3(grid), dim3(256), 0, s, pos_buf, dist_buf, prm->threshold, total, rgba,
                     normals, ids, cpos, &ctr[7]);
  MP_LAUNCHED("k_render_collect");
  // normals at the shaded pixels: the sum of the levels up to normal_lod for
  // residual models (sum_residual_normals), that level alone otherwise
  const int l0 = prm->residual ? 0 : prm->normal_lod;
  for (int l = l0; l <= prm->normal_lod; ++l) {
    rc = launch_level_grad(st, l, cpos, total, nullptr, grad, l > l0, s, &ctr[7]);
    if (rc != MP_OK) return rc;
  }
  hipLaunchKernelGGL(k_render_shade, dim3(grid), dim3(256), 0, s, cams, prm->width, prm->height, a.flip_y,
                     prm->bound_radius, prm->shading, ids, grad, &ctr[7], rgba, normals);
  MP_LAUNCHED("k_render_shade");
  if (stats) {
    unsigned long long h[8] = {0};
    MP_HIP(hipMemcpyAsync(h, ctr, sizeof(h), hipMemcpyDeviceToHost, s));
    MP_HIP(hipStreamSynchronize(s));
    for (int l = 0; l < 3; ++l) stats[l] = (int64_t)h[l];
    stats[3] = (int64_t)h[7];
  }
  return MP_OK;
}

}  // extern "C"
