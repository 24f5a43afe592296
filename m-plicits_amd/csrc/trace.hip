// Multiscale sphere tracing with a persistent, self-compacting ray queue.
//
// Replaces the renderer's graph-captured tracing loop (MIPplicit::SdfInference,
// renderer/cuda/state.cu:848-881: a fixed number of steps on every pixel, the
// SIREN GEMM chain of each level on all pixels every step) and
// IterateSphereTracing_kernel (renderer/cuda/kernels.cu:315-398).  Each wave
// holds a tile of rays (one per MFMA column); after every step the rays that
// hit or left the scene are written out and their slots refilled from the
// wave's pool of ray ids (reserved from a global ray counter 64 at a time), so
// the matrix cores only ever see live rays.  The coarse level
// is evaluated for every live ray; the residual levels only when some ray of
// the tile is inside the coarse band.  A stack with residual levels is traced
// in two launches: every ray on the coarse level until it enters the coarse
// band (parked there), then the parked rays on all levels, so that far-field
// rays (coarse only) and near-surface rays (every level) do not share tiles.
//
// Step rule: IterateSphereTracing_kernel's (kernels.cu:315-398) with its
// fixed-length phases replaced by a per-ray switch, plus a hit test.
//   F_0 = f_0(p), banded = F_0 < s * delta_0 (s = band_scale, 1.4 in the
//   renderer, kernels.cu:335,374).
//   lod-0 phase (kernels.cu:331-336,369-373), until the ray is first banded:
//   step by F_0 - delta_0.
//   residual phase (kernels.cu:337-358,374-394), from then on: banded, step
//   by F = F_0 + f_1(p) + ... + f_{L-1}(p) (every residual level; the
//   renderer's delta_lod_1 is unused); otherwise step by F_0.
//   A banded ray hits when |F| < hit_eps; a ray misses when it leaves the
//   bounding sphere or runs out of steps.  One level: step by F_0, hit when
//   |F_0| < hit_eps.
// Rays start where they enter the sphere of radius bound_radius
// (InitializeRays_kernel, kernels.cu:400-487).
#include "trace_core.h"

#include <vector>

struct mp_trace_batch {
  int shape = 0, n_jobs = 0, blocks = 0;
  void* mem = nullptr;             // tables, counters, per-job scratch (one allocation)
  void* jobs = nullptr;            // TraceArgs [n_jobs] (device)
  void* outs = nullptr;            // TraceOut [n_jobs] (device)
  unsigned long long* ctr = nullptr;  // [n_jobs][8]: next ray, evaluations per level, hits
  std::vector<mp_stack*> stacks;
  std::vector<int64_t> rays;
  std::vector<float*> cpos, grad;
};

namespace mp {

__global__ void k_fill_ull(unsigned long long* p, int n, unsigned long long v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Trace the rays of one job until its queue is empty and this wave's slots
// have all finished.  hand: the wave's LDS handoff region.
template <class S, int NC>
__device__ __forceinline__ void trace_rays(const TraceArgs& a, typename S::Frag* hand) {
  const int lane = threadIdx.x & 63;
  const int col = lane % S::ROWS;
  // a launch after the first traces the rays the launch before parked
  const int64_t total = a.park_in ? (int64_t)*a.n_park_in : a.total_rays;
  const int n_eval = a.eval_levels > 0 ? a.eval_levels : a.n_levels;
  int64_t rid[NC];
  float p[NC][3], d[NC][3], t[NC], t0[NC];
  int steps[NC];
  // a parked ray's first evaluation repeats, at the same point, the coarse
  // level the launch before evaluated: that one is not counted again
  bool fresh[NC];
  // the ray has been in the coarse band (its lod-0 phase is over)
  bool entered[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    rid[c] = -1;
    fresh[c] = false;
    entered[c] = false;
  }
  const float R2 = a.bound_radius * a.bound_radius * 1.0001f;
  unsigned long long evals[MP_MAX_LEVELS] = {0, 0, 0, 0};
  // ray ids this wave has reserved from the global queue and not handed out
  // yet: [pool, pool_end) (wave-uniform)
  constexpr unsigned long long POOL = 64;
  unsigned long long pool = 0, pool_end = 0;

  while (true) {
    // refill empty slots (group-0 lanes vote) from the wave's pool, reserving
    // POOL more ids from the global queue when it runs dry, in rounds: a ray
    // that misses the bounding sphere is written out at once and its slot
    // votes again in the next round (every round hands out at least one id, so
    // the rounds end)
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      bool want = rid[c] < 0;
      while (true) {
        const uint64_t m = __ballot(lane < S::ROWS && want);
        if (m == 0) break;
        if (pool == pool_end) {
          unsigned long long base = 0;
          if (lane == 0) base = atomicAdd(a.next_ray, POOL);
          pool = __shfl(base, 0);
          pool_end = pool + POOL;
        }
        const unsigned long long take =
            (unsigned long long)__popcll(m) < pool_end - pool ? (unsigned long long)__popcll(m) : pool_end - pool;
        const uint64_t below = (col == 0) ? 0ull : (m & ((~0ull) >> (64 - col)));
        const unsigned long long rank = (unsigned long long)__popcll(below);
        if (lane < S::ROWS && want && rank < take) {
          const int64_t nid = (int64_t)(pool + rank);
          want = false;
          if (nid >= total) {
            rid[c] = -1;  // the queue is empty
          } else if (a.park_in) {
            const ParkedRay& r = a.park_in[nid];
            const int64_t id = r.rid;
            bool inside;
            ray_init(a, id, p[c], d[c], t0[c], inside);
            p[c][0] = r.p[0]; p[c][1] = r.p[1]; p[c][2] = r.p[2];
            t[c] = r.t;
            steps[c] = r.steps;
            rid[c] = id;
            fresh[c] = true;
            entered[c] = true;  // parked where it entered the coarse band
          } else {
            bool inside;
            ray_init(a, nid, p[c], d[c], t0[c], inside);
            if (inside) {
              rid[c] = nid;
              t[c] = 0.f;
              steps[c] = 0;
              entered[c] = false;
            } else {
              // the ray never meets the bounding sphere: a miss
              a.depth[nid] = -1.0f;
              a.steps[nid] = 0;
              want = true;
            }
          }
        }
        pool += take;
      }
    }
    // group-1 lanes copy the slot state of their group-0 partner so that both
    // groups evaluate the same points (the divergent refill above may differ)
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int src = col;  // lane in group 0
      rid[c] = (int64_t)__shfl((long long)rid[c], src);
      p[c][0] = __shfl(p[c][0], src); p[c][1] = __shfl(p[c][1], src); p[c][2] = __shfl(p[c][2], src);
      d[c][0] = __shfl(d[c][0], src); d[c][1] = __shfl(d[c][1], src); d[c][2] = __shfl(d[c][2], src);
      t[c] = __shfl(t[c], src); t0[c] = __shfl(t0[c], src);
      steps[c] = __shfl(steps[c], src);
      entered[c] = __shfl((int)entered[c], src) != 0;
    }
    bool live = false;
#pragma unroll
    for (int c = 0; c < NC; ++c) live |= rid[c] >= 0;
    if (__ballot(live) == 0) break;

    float x[NC][3], F[NC], Fsum[NC];
    bool need[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const bool act = rid[c] >= 0;
      x[c][0] = act ? p[c][0] : 0.f;
      x[c][1] = act ? p[c][1] : 0.f;
      x[c][2] = act ? p[c][2] : 0.f;
    }
    // one call site for every level keeps a single copy of each width's code
    for (int l = 0; l < n_eval; ++l) {
      // column blocks with a ray that needs level l
      unsigned cols = ~0u;
      if (l > 0) {
        cols = 0;
#pragma unroll
        for (int c = 0; c < NC; ++c) cols |= (__ballot(need[c]) != 0 ? 1u : 0u) << c;
        if (cols == 0) break;
      }
      float r[NC];
#pragma unroll
      for (int c = 0; c < NC; ++c) r[c] = 0.f;
      eval_level_any<S, NC>(a.levels[l], x, r, hand, cols);
      if (l == 0) {
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          F[c] = r[c];
          Fsum[c] = r[c];
          need[c] = rid[c] >= 0 && a.n_levels > 1 && r[c] < a.band_scale * a.deltas[0];
          if (lane < S::ROWS && rid[c] >= 0 && !fresh[c]) evals[0]++;
        }
      } else {
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          if (need[c]) {
            Fsum[c] += r[c];
            if (lane < S::ROWS) evals[l]++;
          }
          need[c] = need[c] && l + 1 < a.n_levels;
        }
      }
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) fresh[c] = false;
    // rays that need the next launch's level are parked before their step
    // (need[c]: level n_eval is due at this point)
    if (a.park_out) {
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const bool park = need[c];
        const uint64_t m = __ballot(lane < S::ROWS && park);
        if (m == 0) continue;
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(a.n_park_out, (unsigned long long)__popcll(m));
        base = __shfl(base, 0);
        if (lane < S::ROWS && park) {
          const uint64_t below = (col == 0) ? 0ull : (m & ((~0ull) >> (64 - col)));
          ParkedRay& r = a.park_out[base + __popcll(below)];
          r.rid = (uint32_t)rid[c];
          r.steps = steps[c];
          r.t = t[c];
          r.p[0] = p[c][0]; r.p[1] = p[c][1]; r.p[2] = p[c][2];
        }
        if (park) rid[c] = -1;
      }
    }
    // step, hit and exit tests
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      if (rid[c] < 0) continue;
      const bool banded = a.n_levels > 1 && F[c] < a.band_scale * a.deltas[0];
      float s;
      bool hit;
      if (a.n_levels == 1) {
        s = F[c];
        hit = fabsf(F[c]) < a.hit_eps;
      } else if (!banded) {
        s = entered[c] ? F[c] : F[c] - a.deltas[0];
        hit = false;
      } else {
        s = Fsum[c];
        hit = fabsf(Fsum[c]) < a.hit_eps;
        entered[c] = true;
      }
      bool done = hit;
      if (!hit) {
        p[c][0] += s * d[c][0];
        p[c][1] += s * d[c][1];
        p[c][2] += s * d[c][2];
        t[c] += s;
        steps[c] += 1;
        const float r2 = p[c][0] * p[c][0] + p[c][1] * p[c][1] + p[c][2] * p[c][2];
        // steps are signed: the ray left when it moves away from the centre
        const float outward = (p[c][0] * d[c][0] + p[c][1] * d[c][1] + p[c][2] * d[c][2]) * s;
        done = steps[c] >= a.max_steps || (r2 > R2 && outward > 0.f);
      }
      if (done) {
        if (lane < S::ROWS) {
          const int64_t id = rid[c];
          a.steps[id] = steps[c];
          if (hit) {
            a.depth[id] = t0[c] + t[c];
            a.hit_pos[3 * id + 0] = p[c][0];
            a.hit_pos[3 * id + 1] = p[c][1];
            a.hit_pos[3 * id + 2] = p[c][2];
          } else {
            a.depth[id] = -1.0f;
          }
        }
        rid[c] = -1;
      }
    }
  }
  // per-level evaluation counts (one atomic per lane that counted)
  if (lane < S::ROWS)
    for (int l = 0; l < a.n_levels; ++l)
      if (evals[l]) atomicAdd(&a.stats[l], evals[l]);
}

template <class S, int NC>
__global__ __launch_bounds__(256) void k_trace(TraceArgs a) {
  extern __shared__ uint4 lds_raw[];
  trace_rays<S, NC>(a, reinterpret_cast<typename S::Frag*>(reinterpret_cast<char*>(lds_raw) +
                                                          (threadIdx.x >> 6) * TRACE_HAND_BYTES));
}

// Every job in turn: a wave moves to the next job once the current one has no
// rays left for it, so tiles never mix networks and idle waves help with the
// jobs still running.
template <class S>
__global__ __launch_bounds__(256) void k_trace_jobs(const TraceArgs* __restrict__ jobs, int n_jobs) {
  extern __shared__ uint4 lds_raw[];
  typename S::Frag* hand = reinterpret_cast<typename S::Frag*>(reinterpret_cast<char*>(lds_raw) +
                                                               (threadIdx.x >> 6) * TRACE_HAND_BYTES);
  for (int j = 0; j < n_jobs; ++j) {
    // two-block tiles unless a level of the job is 512 wide (FP32 and
    // three-term BF16: always one)
    if constexpr (S::ROWS == 32 && !S::SPLIT) {
      if (jobs[j].cols == 2) {
        trace_rays<S, 2>(jobs[j], hand);
        continue;
      }
    }
    trace_rays<S, 1>(jobs[j], hand);
  }
}

// Compact the hit rays: ids and positions.  Each workgroup takes
// COLLECT_PER_BLOCK consecutive rays (COLLECT_K per thread, strided by the
// block size) and reserves its output range with one atomic: a reservation per
// wave on the one counter serialises at the L2.
static constexpr int COLLECT_K = 8;
static constexpr int COLLECT_PER_BLOCK = 256 * COLLECT_K;

__global__ __launch_bounds__(256) void k_collect_hits(const float* depth, const float* hit_pos, int64_t n,
                                                      uint32_t* ids, float* pos, unsigned long long* count) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t b0 = (int64_t)blockIdx.x * COLLECT_PER_BLOCK;
  uint64_t m[COLLECT_K];
  int wave_hits = 0;
#pragma unroll
  for (int k = 0; k < COLLECT_K; ++k) {
    const int64_t i = b0 + k * 256 + threadIdx.x;
    m[k] = __ballot(i < n && depth[i] >= 0.f);
    wave_hits += __popcll(m[k]);
  }
  __shared__ int wsum[4];
  __shared__ unsigned long long base;
  if (lane == 0) wsum[wave] = wave_hits;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int tot = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    base = tot ? atomicAdd(count, (unsigned long long)tot) : 0ull;
  }
  __syncthreads();
  unsigned long long k0 = base;
  for (int w = 0; w < wave; ++w) k0 += wsum[w];
  const uint64_t below_mask = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
#pragma unroll
  for (int k = 0; k < COLLECT_K; ++k) {
    if ((m[k] >> lane) & 1ull) {
      const int64_t i = b0 + k * 256 + threadIdx.x;
      const unsigned long long o = k0 + __popcll(m[k] & below_mask);
      ids[o] = (uint32_t)i;
      pos[3 * o + 0] = hit_pos[3 * i + 0];
      pos[3 * o + 1] = hit_pos[3 * i + 1];
      pos[3 * o + 2] = hit_pos[3 * i + 2];
    }
    k0 += __popcll(m[k]);
  }
}

// Background, then Lambert shading of the hits with the light at the eye
// (Shade_kernel's diffuse term, renderer/cuda/kernels.cu:122-275:
// ambient 0.2, diffuse 0.54, light = -ray direction).
__global__ void k_background(uint32_t* rgba, float* normals, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  rgba[i] = rgba8(1.f, 1.f, 1.f);
  if (normals) {
    normals[3 * i] = 0.f; normals[3 * i + 1] = 0.f; normals[3 * i + 2] = 0.f;
  }
}


// Per-job outputs of the batched tracer beside TraceArgs.
struct TraceOut {
  uint32_t* ids;              // [rays] hit ray ids, compacted
  float* cpos;                // [rays][3] hit positions, compacted
  float* grad;                // [rays][3] gradient at the hits
  unsigned long long* nhits;  // device count
};

// Background for every ray of every job, then the hits compacted per job.
__global__ void k_collect_jobs(const TraceArgs* __restrict__ jobs, const TraceOut* __restrict__ outs, int n_jobs) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int lane = threadIdx.x & 63;
  for (int j = 0; j < n_jobs; ++j) {
    const TraceArgs& a = jobs[j];
    const TraceOut& o = outs[j];
    const int64_t n = a.total_rays;
    // wave-uniform loop bound: every lane of a wave takes the ballots
    for (int64_t w = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63); w < n; w += stride) {
      const int64_t i = w + lane;
      bool hit = false;
      if (i < n) {
        a.rgba[i] = rgba8(1.f, 1.f, 1.f);
        if (a.normals) {
          a.normals[3 * i] = 0.f; a.normals[3 * i + 1] = 0.f; a.normals[3 * i + 2] = 0.f;
        }
        hit = a.depth[i] >= 0.f;
      }
      const uint64_t m = __ballot(hit);
      if (!m) continue;
      unsigned long long base = 0;
      if (lane == 0) base = atomicAdd(o.nhits, (unsigned long long)__popcll(m));
      base = __shfl(base, 0);
      if (hit) {
        const uint64_t below = (lane == 0) ? 0ull : (m & ((~0ull) >> (64 - lane)));
        const unsigned long long k = base + __popcll(below);
        o.ids[k] = (uint32_t)i;
        o.cpos[3 * k + 0] = a.hit_pos[3 * i + 0];
        o.cpos[3 * k + 1] = a.hit_pos[3 * i + 1];
        o.cpos[3 * k + 2] = a.hit_pos[3 * i + 2];
      }
    }
  }
}

__device__ inline void shade_one(const TraceArgs& a, const uint32_t* ids, const float* grad, int64_t k) {
  const int64_t id = ids[k];
  float o[3], d[3], t0;
  bool inside;
  ray_init(a, id, o, d, t0, inside);
  float nx = grad[3 * k], ny = grad[3 * k + 1], nz = grad[3 * k + 2];
  const float len = sqrtf(nx * nx + ny * ny + nz * nz);
  const float inv = len > 0.f ? 1.0f / len : 0.f;
  nx *= inv; ny *= inv; nz *= inv;
  const float lambert = fmaxf(-(nx * d[0] + ny * d[1] + nz * d[2]), 0.f);
  const float c = 0.2f + 0.54f * lambert;
  a.rgba[id] = rgba8(c, c, c);
  if (a.normals) {
    a.normals[3 * id] = nx; a.normals[3 * id + 1] = ny; a.normals[3 * id + 2] = nz;
  }
}

__global__ void k_shade(TraceArgs a, const uint32_t* ids, const float* grad, int64_t nh) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nh) shade_one(a, ids, grad, k);
}

__global__ void k_shade_jobs(const TraceArgs* __restrict__ jobs, const TraceOut* __restrict__ outs, int n_jobs) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int j = 0; j < n_jobs; ++j) {
    const int64_t nh = (int64_t)*outs[j].nhits;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nh; k += stride)
      shade_one(jobs[j], outs[j].ids, outs[j].grad, k);
  }
}

template <class S>
static int launch_trace_jobs(int blocks, size_t lds, hipStream_t s, const TraceArgs* jobs, int n_jobs) {
  static std::atomic<uint64_t> attr_devs{0};
  MP_HIP(set_lds_once(attr_devs, reinterpret_cast<const void*>(&k_trace_jobs<S>), (int)lds));
  hipLaunchKernelGGL((k_trace_jobs<S>), dim3(blocks), dim3(256), lds, s, jobs, n_jobs);
  MP_LAUNCHED("k_trace_jobs");
  return MP_OK;
}

template <class S, int NC>
static int launch_trace(int blocks, size_t lds, hipStream_t s, const TraceArgs& a) {
  static std::atomic<uint64_t> attr_devs{0};
  MP_HIP(set_lds_once(attr_devs, reinterpret_cast<const void*>(&k_trace<S, NC>), (int)lds));
  hipLaunchKernelGGL((k_trace<S, NC>), dim3(blocks), dim3(256), lds, s, a);
  MP_LAUNCHED("k_trace");
  return MP_OK;
}

// The tracer instantiates one level shape per stack: an MP_BF16X3F stack is
// traced when every level it runs in the mixed shape has that shape, i.e.
// every level up to width 256 has three hidden layers.
static bool mixed_levels_ok(const mp_stack* st) {
  if (st->shape != SH_32x32x16_BF16X3F) return true;
  for (int l = 0; l < st->n_levels; ++l)
    if (st->lv[l].width <= 256 && st->lv[l].shape != SH_32x32x16_BF16X3F) return false;
  return true;
}

}  // namespace mp

using namespace mp;

extern "C" {

int mp_trace(mp_stack* st, const float* cams, int n_cams, const mp_trace_params* prm, float* depth,
             float* normals, uint32_t* rgba, int32_t* steps, int64_t* stats, void* stream) {
  MP_REQUIRE(st != nullptr && cams != nullptr && prm != nullptr, MP_EINVAL, "mp_trace: NULL argument");
  MP_REQUIRE(n_cams >= 1 && prm->width >= 1 && prm->height >= 1, MP_EINVAL, "mp_trace: empty image");
  MP_REQUIRE(depth != nullptr && steps != nullptr && rgba != nullptr, MP_EINVAL,
             "mp_trace: depth, steps and rgba are required");
  MP_REQUIRE(prm->n_deltas == st->n_levels - 1, MP_EINVAL, "mp_trace: need %d deltas, got %d",
             st->n_levels - 1, prm->n_deltas);
  MP_REQUIRE(prm->max_steps >= 1 && prm->hit_eps > 0.f && prm->bound_radius > 0.f, MP_EINVAL,
             "mp_trace: bad parameters");
  MP_REQUIRE(mixed_levels_ok(st), MP_EUNSUPPORTED,
             "mp_trace: MP_BF16X3F traces stacks whose levels up to width 256 have 3 hidden layers; use MP_BF16X3");
  hipStream_t s = as_stream(stream);
  const int64_t total = (int64_t)n_cams * prm->width * prm->height;
  MP_REQUIRE(total < (1ll << 32), MP_EUNSUPPORTED, "mp_trace: too many rays");
  int maxw = 0;
  for (int l = 0; l < st->n_levels; ++l) maxw = st->lv[l].width > maxw ? st->lv[l].width : maxw;

  // scratch: counters | hit positions [total][3] | hit ids [total] | compact pos [total][3] | grad [total][3]
  const size_t off_pos = 256;
  const size_t off_ids = off_pos + (size_t)total * 12;
  const size_t off_cpos = (off_ids + (size_t)total * 4 + 255) & ~(size_t)255;
  const size_t off_grad = off_cpos + (size_t)total * 12;
  int rc = ensure_scratch(st, off_grad + (size_t)total * 12);
  if (rc != MP_OK) return rc;
  char* base = reinterpret_cast<char*>(st->scratch);
  unsigned long long* ctr = reinterpret_cast<unsigned long long*>(base);
  float* hit_pos = reinterpret_cast<float*>(base + off_pos);
  uint32_t* ids = reinterpret_cast<uint32_t*>(base + off_ids);
  float* cpos = reinterpret_cast<float*>(base + off_cpos);
  float* grad = reinterpret_cast<float*>(base + off_grad);

  TraceArgs a{};
  a.cams = cams;
  a.width = prm->width;
  a.height = prm->height;
  a.total_rays = total;
  a.max_steps = prm->max_steps;
  a.hit_eps = prm->hit_eps;
  a.bound_radius = prm->bound_radius;
  a.band_scale = prm->band_scale;
  a.pixel_center = prm->pixel_center;
  a.flip_y = prm->flip_y;
  a.n_levels = st->n_levels;
  for (int l = 0; l < st->n_levels; ++l) a.levels[l] = st->lv[l].dev;
  for (int i = 0; i < prm->n_deltas; ++i) a.deltas[i] = prm->deltas[i];
  a.next_ray = &ctr[0];
  a.stats = &ctr[1];
  a.depth = depth;
  a.steps = steps;
  a.hit_pos = hit_pos;
  a.rgba = rgba;
  a.normals = normals;

  hipLaunchKernelGGL(k_fill_ull, dim3(1), dim3(32), 0, s, ctr, 32, 0ull);
  MP_LAUNCHED("k_fill_ull");
  int dev = 0, cus = 0;
  MP_HIP(hipGetDevice(&dev));
  MP_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  // two column blocks per wave unless the stack asks for one (512-wide levels
  // then run one block at a time, only where a ray needs them); three-term
  // BF16 tiles hold one
  const int want = st->trace_tiles ? st->trace_tiles : st->nc;
  const int nc = want == 2 && st->shape != SH_32x32x16_BF16X3 && st->shape != SH_32x32x16_BF16X3F ? 2 : 1;
  // persistent grid: one workgroup of four waves per CU (the LDS handoff
  // regions take 128 KB of the CU's LDS)
  const int blocks = cus;
  const size_t lds = 4 * TRACE_HAND_BYTES;
  auto launch = [&](const TraceArgs& ta) -> int {
    switch (st->shape) {
      case SH_32x32x16_BF16:
        return nc == 2 ? launch_trace<S32x32x16_bf16, 2>(blocks, lds, s, ta)
                       : launch_trace<S32x32x16_bf16, 1>(blocks, lds, s, ta);
      case SH_32x32x16_F16:
        return nc == 2 ? launch_trace<S32x32x16_f16, 2>(blocks, lds, s, ta)
                       : launch_trace<S32x32x16_f16, 1>(blocks, lds, s, ta);
      case SH_16x16x4_F32:
        return launch_trace<S16x16x4_f32, 1>(blocks, lds, s, ta);
      case SH_32x32x16_BF16X3:
        return launch_trace<S32x32x16_bf16x3, 1>(blocks, lds, s, ta);
      case SH_32x32x16_BF16X3F:
        return launch_trace<S32x32x16_bf16x3f, 1>(blocks, lds, s, ta);
      default:
        set_error("mp_trace: unsupported fragment shape %d", st->shape);
        return MP_EUNSUPPORTED;
    }
  };
  if (st->n_levels == 1) {
    rc = launch(a);
    if (rc != MP_OK) return rc;
  } else {
    // two launches: the first traces the coarse level and parks every ray
    // where it enters the coarse band (before its step); the second traces
    // the parked rays on all levels.  The parked list (24 B per ray, each ray
    // parked at most once) takes the compacted-hit and gradient regions,
    // written only after tracing.
    static_assert(sizeof(ParkedRay) == 24, "parked rays must fit the 24 B per ray of cpos + grad");
    ParkedRay* parked = reinterpret_cast<ParkedRay*>(cpos);
    a.eval_levels = 1;
    a.park_in = nullptr;
    a.n_park_in = nullptr;
    a.park_out = parked;
    a.n_park_out = &ctr[9];
    a.next_ray = &ctr[0];
    rc = launch(a);
    if (rc != MP_OK) return rc;
    a.eval_levels = st->n_levels;
    a.park_in = parked;
    a.n_park_in = &ctr[9];
    a.park_out = nullptr;
    a.n_park_out = nullptr;
    a.next_ray = &ctr[17];
    rc = launch(a);
    if (rc != MP_OK) return rc;
  }
  hipLaunchKernelGGL(k_collect_hits, dim3((unsigned)((total + COLLECT_PER_BLOCK - 1) / COLLECT_PER_BLOCK)),
                     dim3(256), 0, s, depth,
                     hit_pos, total, ids, cpos, &ctr[8]);
  MP_LAUNCHED("k_collect_hits");
  hipLaunchKernelGGL(k_background, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, rgba,
                     normals, total);
  MP_LAUNCHED("k_background");
  unsigned long long h[9] = {0};
  MP_HIP(hipMemcpyAsync(h, ctr, sizeof(h), hipMemcpyDeviceToHost, s));
  MP_HIP(hipStreamSynchronize(s));
  const int64_t nh = (int64_t)h[8];
  if (nh > 0) {
    for (int l = 0; l < st->n_levels; ++l) {
      rc = launch_level_grad(st, l, cpos, nh, nullptr, grad, l > 0, s);
      if (rc != MP_OK) return rc;
    }
    hipLaunchKernelGGL(k_shade, dim3((unsigned)((nh + 255) / 256)), dim3(256), 0, s, a, ids, grad, nh);
    MP_LAUNCHED("k_shade");
  }
  if (stats) {
    for (int l = 0; l < st->n_levels; ++l) stats[l] = (int64_t)h[1 + l];
    stats[MP_MAX_LEVELS] = nh;
  }
  return MP_OK;
}

static constexpr int TB_CTR = 8;  // counters per job: next ray, evaluations [4], hits at [7]

void mp_trace_batch_destroy(mp_trace_batch* p) {
  if (!p) return;
  if (p->mem) (void)hipFree(p->mem);
  delete p;
}

int mp_trace_batch_plan(const mp_trace_job* jobs, int n_jobs, void* stream, mp_trace_batch** out) {
  MP_REQUIRE(out != nullptr, MP_EINVAL, "mp_trace_batch_plan: out is NULL");
  *out = nullptr;
  MP_REQUIRE(jobs != nullptr && n_jobs >= 1, MP_EINVAL, "mp_trace_batch_plan: no jobs");
  hipStream_t s = as_stream(stream);
  const int shape = jobs[0].stack ? jobs[0].stack->shape : -1;
  for (int j = 0; j < n_jobs; ++j) {
    const mp_trace_job& J = jobs[j];
    const mp_trace_params* prm = &J.params;
    MP_REQUIRE(J.stack != nullptr && J.cam != nullptr, MP_EINVAL, "mp_trace_batch_plan: job %d: NULL stack or camera", j);
    MP_REQUIRE(mixed_levels_ok(J.stack), MP_EUNSUPPORTED,
               "mp_trace_batch_plan: job %d: MP_BF16X3F traces stacks whose levels up to width 256 have 3 hidden "
               "layers; use MP_BF16X3", j);
    MP_REQUIRE(J.stack->shape == shape, MP_EINVAL, "mp_trace_batch_plan: job %d: every stack must have the same precision", j);
    MP_REQUIRE(J.depth != nullptr && J.steps != nullptr && J.rgba != nullptr, MP_EINVAL,
               "mp_trace_batch_plan: job %d: depth, steps and rgba are required", j);
    MP_REQUIRE(prm->width >= 1 && prm->height >= 1 && prm->max_steps >= 1 && prm->hit_eps > 0.f &&
                   prm->bound_radius > 0.f, MP_EINVAL, "mp_trace_batch_plan: job %d: bad parameters", j);
    MP_REQUIRE(prm->n_deltas == J.stack->n_levels - 1, MP_EINVAL, "mp_trace_batch_plan: job %d: need %d deltas, got %d",
               j, J.stack->n_levels - 1, prm->n_deltas);
    MP_REQUIRE((int64_t)prm->width * prm->height < (1ll << 32), MP_EUNSUPPORTED, "mp_trace_batch_plan: job %d: too many rays", j);
  }
  mp_trace_batch* p = new mp_trace_batch();
  p->shape = shape;
  p->n_jobs = n_jobs;
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  size_t off = 0;
  const size_t off_jobs = off; off += al(sizeof(TraceArgs) * n_jobs);
  const size_t off_outs = off; off += al(sizeof(TraceOut) * n_jobs);
  const size_t off_ctr = off; off += al(sizeof(unsigned long long) * TB_CTR * n_jobs);
  std::vector<size_t> o_pos(n_jobs), o_ids(n_jobs), o_cpos(n_jobs), o_grad(n_jobs);
  for (int j = 0; j < n_jobs; ++j) {
    const int64_t n = (int64_t)jobs[j].params.width * jobs[j].params.height;
    o_pos[j] = off; off += al((size_t)n * 12);
    o_ids[j] = off; off += al((size_t)n * 4);
    o_cpos[j] = off; off += al((size_t)n * 12);
    o_grad[j] = off; off += al((size_t)n * 12);
  }
  hipError_t e = hipMalloc(&p->mem, off);
  if (e != hipSuccess) {
    delete p;
    set_error("mp_trace_batch_plan: hipMalloc(%zu) failed: %s", off, hipGetErrorString(e));
    return MP_ENOMEM;
  }
  char* base = reinterpret_cast<char*>(p->mem);
  p->jobs = base + off_jobs;
  p->outs = base + off_outs;
  p->ctr = reinterpret_cast<unsigned long long*>(base + off_ctr);
  std::vector<TraceArgs> ta(n_jobs);
  std::vector<TraceOut> to(n_jobs);
  for (int j = 0; j < n_jobs; ++j) {
    const mp_trace_job& J = jobs[j];
    TraceArgs& a = ta[j];
    a = TraceArgs{};
    a.cams = J.cam;
    a.width = J.params.width;
    a.height = J.params.height;
    a.total_rays = (int64_t)a.width * a.height;
    a.max_steps = J.params.max_steps;
    a.hit_eps = J.params.hit_eps;
    a.bound_radius = J.params.bound_radius;
    a.band_scale = J.params.band_scale;
    a.pixel_center = J.params.pixel_center;
    a.flip_y = J.params.flip_y;
    a.n_levels = J.stack->n_levels;
    int jw = 0;
    for (int l = 0; l < a.n_levels; ++l) {
      a.levels[l] = J.stack->lv[l].dev;
      jw = J.stack->lv[l].width > jw ? J.stack->lv[l].width : jw;
    }
    const int want = J.stack->trace_tiles ? J.stack->trace_tiles : J.stack->nc;
    a.cols = (want == 2 && jw <= 256 && J.stack->shape != SH_32x32x16_BF16X3 &&
              J.stack->shape != SH_32x32x16_BF16X3F) ? 2 : 1;
    for (int i = 0; i < J.params.n_deltas; ++i) a.deltas[i] = J.params.deltas[i];
    unsigned long long* c = p->ctr + (size_t)j * TB_CTR;
    a.next_ray = &c[0];
    a.stats = &c[1];
    a.depth = J.depth;
    a.steps = J.steps;
    a.hit_pos = reinterpret_cast<float*>(base + o_pos[j]);
    a.rgba = J.rgba;
    a.normals = J.normals;
    to[j].ids = reinterpret_cast<uint32_t*>(base + o_ids[j]);
    to[j].cpos = reinterpret_cast<float*>(base + o_cpos[j]);
    to[j].grad = reinterpret_cast<float*>(base + o_grad[j]);
    to[j].nhits = &c[TB_CTR - 1];
    p->stacks.push_back(J.stack);
    p->rays.push_back(a.total_rays);
    p->cpos.push_back(to[j].cpos);
    p->grad.push_back(to[j].grad);
  }
  e = hipMemcpyAsync(p->jobs, ta.data(), sizeof(TraceArgs) * n_jobs, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(p->outs, to.data(), sizeof(TraceOut) * n_jobs, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);  // the host tables go out of scope
  int dev = 0, cus = 0;
  if (e == hipSuccess) e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) {
    mp_trace_batch_destroy(p);
    set_error("mp_trace_batch_plan: %s", hipGetErrorString(e));
    return MP_EHIP;
  }
  p->blocks = cus;
  *out = p;
  return MP_OK;
}

int mp_trace_batch_run(mp_trace_batch* p, void* stream) {
  MP_REQUIRE(p != nullptr, MP_EINVAL, "mp_trace_batch_run: plan is NULL");
  hipStream_t s = as_stream(stream);
  const int n_ctr = TB_CTR * p->n_jobs;
  hipLaunchKernelGGL(k_fill_ull, dim3((n_ctr + 255) / 256), dim3(256), 0, s, p->ctr, n_ctr, 0ull);
  MP_LAUNCHED("k_fill_ull");
  const TraceArgs* jobs = reinterpret_cast<const TraceArgs*>(p->jobs);
  const TraceOut* outs = reinterpret_cast<const TraceOut*>(p->outs);
  const size_t lds = 4 * TRACE_HAND_BYTES;
  int rc = MP_OK;
  switch (p->shape) {
    case SH_32x32x16_BF16: rc = launch_trace_jobs<S32x32x16_bf16>(p->blocks, lds, s, jobs, p->n_jobs); break;
    case SH_32x32x16_F16: rc = launch_trace_jobs<S32x32x16_f16>(p->blocks, lds, s, jobs, p->n_jobs); break;
    case SH_16x16x4_F32: rc = launch_trace_jobs<S16x16x4_f32>(p->blocks, lds, s, jobs, p->n_jobs); break;
    case SH_32x32x16_BF16X3: rc = launch_trace_jobs<S32x32x16_bf16x3>(p->blocks, lds, s, jobs, p->n_jobs); break;
    case SH_32x32x16_BF16X3F: rc = launch_trace_jobs<S32x32x16_bf16x3f>(p->blocks, lds, s, jobs, p->n_jobs); break;
    default:
      set_error("mp_trace_batch_run: unsupported fragment shape %d", p->shape);
      return MP_EUNSUPPORTED;
  }
  if (rc != MP_OK) return rc;
  hipLaunchKernelGGL(k_collect_jobs, dim3(p->blocks * 4), dim3(256), 0, s, jobs, outs, p->n_jobs);
  MP_LAUNCHED("k_collect_jobs");
  // analytic normals at the hits: the hit counts stay on the device
  for (int j = 0; j < p->n_jobs; ++j) {
    const unsigned long long* nh = p->ctr + (size_t)j * TB_CTR + TB_CTR - 1;
    for (int l = 0; l < p->stacks[j]->n_levels; ++l) {
      rc = launch_level_grad(p->stacks[j], l, p->cpos[j], p->rays[j], nullptr, p->grad[j], l > 0, s, nh);
      if (rc != MP_OK) return rc;
    }
  }
  hipLaunchKernelGGL(k_shade_jobs, dim3(p->blocks * 4), dim3(256), 0, s, jobs, outs, p->n_jobs);
  MP_LAUNCHED("k_shade_jobs");
  return MP_OK;
}

int mp_trace_batch_stats(const mp_trace_batch* p, int64_t* stats, void* stream) {
  MP_REQUIRE(p != nullptr && stats != nullptr, MP_EINVAL, "mp_trace_batch_stats: NULL argument");
  hipStream_t s = as_stream(stream);
  std::vector<unsigned long long> h((size_t)TB_CTR * p->n_jobs);
  MP_HIP(hipMemcpyAsync(h.data(), p->ctr, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
  MP_HIP(hipStreamSynchronize(s));
  for (int j = 0; j < p->n_jobs; ++j) {
    for (int l = 0; l < MP_MAX_LEVELS; ++l)
      stats[(size_t)j * (MP_MAX_LEVELS + 1) + l] = (int64_t)h[(size_t)j * TB_CTR + 1 + l];
    stats[(size_t)j * (MP_MAX_LEVELS + 1) + MP_MAX_LEVELS] = (int64_t)h[(size_t)j * TB_CTR + TB_CTR - 1];
  }
  return MP_OK;
}

}  // extern "C"
