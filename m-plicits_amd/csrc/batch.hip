// Batched band-culled extraction of many level stacks of uneven width and
// resolution, one persistent launch per stage, every launch free of host
// synchronisation, allocation and memset so the whole run can be captured in
// a hipGraph and replayed.
//
// Per model, the same semantics as mp_extract_volume (create_mesh_multistage's
// sampling loop, i3d/meshing.py:203-238, per-chunk break).  Stages:
//   reset     counters and chunk flags of every model (fill kernel)
//   level 0   every model's coarse level over its grid; candidates |s| < max delta
//   level l   band mark over all models' candidates, band select, then level l
//             of every model with more than l levels over its selection
// The level stages run k_level_jobs: persistent workgroups take blocks from a
// stage-wide counter and map them to (model, block) through the stage's job
// table; block counts of index-list jobs are read from device counters at run
// time, so the launches never wait for the host.
#include "eval_dispatch.h"
#include "level_core.h"

#include <vector>

struct mp_batch {
  int shape = 0;
  int n_jobs = 0;
  int max_levels = 0;
  int blocks = 0;                // persistent workgroups per stage
  void* mem = nullptr;           // tables, counters, flags, lists (one allocation)
  size_t mem_bytes = 0;
  // LevelJob tables per stage, one per tile shape (0: default, 1: half column blocks)
  void* level_tab[MP_MAX_LEVELS][2] = {};
  int n_level_jobs[MP_MAX_LEVELS][2] = {};
  void* band_tab[MP_MAX_LEVELS] = {};   // BandJob [n]
  int n_band_jobs[MP_MAX_LEVELS] = {};
  uint32_t* counters = nullptr;  // [MP_MAX_LEVELS][2] stage item counters, then [n_jobs][MP_MAX_LEVELS] counts
  int64_t n_counters = 0;
  uint8_t* flags = nullptr;      // chunk flags of every model
  int64_t n_flags = 0;
  std::vector<int64_t> grid_points;  // N^3 per job
  std::vector<int> levels;           // levels per job
  std::vector<int> culled;           // band culling per job
};

namespace mp {

struct LevelJob {
  LevelDev L;
  EvalTarget t;  // SRC_GRID or SRC_LIST (t.list_count on the device)
};

struct BandJob {
  const uint32_t* cand;
  const uint32_t* ccount;
  const float* vol;
  float delta;
  int64_t chunk, nchunks;
  uint8_t* flags;  // [levels - 1][nchunks] of this model
  int level;
  uint32_t* list;
  uint32_t* lcount;
};

__global__ void k_fill_words(uint32_t* p, int64_t n, uint32_t v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void k_fill_bytes(uint8_t* p, int64_t n, uint8_t v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// Column blocks per wave of the batched level kernel: the width's default
// (the B operand set fills 128 VGPRs) or, in the HALF kernel, half of it (a
// tile shape the autotuner may pick: half the registers and LDS hand-off per
// wave).  The two shapes are separate kernels, each with one instantiation
// per width.
template <class S>
constexpr int batch_cols(int W, bool half) {
  return half && default_cols<S>(W) > 1 ? default_cols<S>(W) / 2 : default_cols<S>(W);
}

template <class S, int W, bool HALF>
__device__ inline int64_t job_blocks(const EvalTarget& t) {
  constexpr int per_block = 4 * LevelCore<S, W, batch_cols<S>(W, HALF), true>::TILE;
  const int64_t n = (t.src == SRC_LIST) ? (int64_t)*t.list_count : t.n;
  return (n + per_block - 1) / per_block;
}

template <class S, bool HALF>
__device__ inline int64_t blocks_of(const LevelJob& j) {
  switch (j.L.width) {
    case 64: return job_blocks<S, 64, HALF>(j.t);
    case 128: return job_blocks<S, 128, HALF>(j.t);
    case 256: return job_blocks<S, 256, HALF>(j.t);
    default: return job_blocks<S, 512, HALF>(j.t);
  }
}

// Persistent level stage: every workgroup loops over blocks of the stage's jobs.
template <class S, bool HALF>
__global__ __launch_bounds__(256) void k_level_jobs(const LevelJob* __restrict__ jobs, int n_jobs,
                                                    uint32_t* item_ctr) {
  extern __shared__ uint4 lds_raw[];
  typename S::Frag* lds = reinterpret_cast<typename S::Frag*>(lds_raw);
  // The block index is broadcast through the first LDS word: between the
  // barrier that ends a block (all waves done, their ring copies waited for)
  // and the one after the read, nothing else uses the LDS.
  volatile uint32_t* s_item = reinterpret_cast<volatile uint32_t*>(lds_raw);
  while (true) {
    if (threadIdx.x == 0) *s_item = atomicAdd(item_ctr, 1u);
    __syncthreads();
    int64_t b = *s_item;
    __syncthreads();  // every wave has read it before the LDS is reused
    int j = 0;
    for (; j < n_jobs; ++j) {
      const int64_t nb = blocks_of<S, HALF>(jobs[j]);
      if (b < nb) break;
      b -= nb;
    }
    if (j == n_jobs) break;  // workgroup-uniform
    const LevelJob& job = jobs[j];
    const int64_t n = (job.t.src == SRC_LIST) ? (int64_t)*job.t.list_count : job.t.n;
    // one call site per width keeps one copy of each width's code
    switch (job.L.width) {
      case 64: eval_block<S, 64, batch_cols<S>(64, HALF)>(job.L, job.t, n, b, lds); break;
      case 128: eval_block<S, 128, batch_cols<S>(128, HALF)>(job.L, job.t, n, b, lds); break;
      case 256: eval_block<S, 256, batch_cols<S>(256, HALF)>(job.L, job.t, n, b, lds); break;
      default: eval_block<S, 512, batch_cols<S>(512, HALF)>(job.L, job.t, n, b, lds); break;
    }
    __syncthreads();  // the next block may reuse the LDS ring
  }
}

template <class S, bool HALF>
constexpr size_t jobs_lds_bytes() {
  size_t m = 0;
  const size_t b[4] = {LevelCore<S, 64, batch_cols<S>(64, HALF), true>::LDS_BYTES,
                       LevelCore<S, 128, batch_cols<S>(128, HALF), true>::LDS_BYTES,
                       LevelCore<S, 256, batch_cols<S>(256, HALF), true>::LDS_BYTES,
                       LevelCore<S, 512, batch_cols<S>(512, HALF), true>::LDS_BYTES};
  for (size_t v : b) m = v > m ? v : m;
  return m;
}

// flags[c] = 1 if some candidate of chunk c has |s| < delta, every job.
__global__ void k_band_mark_jobs(const BandJob* __restrict__ jobs, int n_jobs) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int j = 0; j < n_jobs; ++j) {
    const BandJob& b = jobs[j];
    const int64_t n = *b.ccount;
    uint8_t* flags = b.flags + (int64_t)(b.level - 1) * b.nchunks;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
      const uint32_t k = b.cand[i];
      if (fabsf(b.vol[k]) < b.delta) flags[k / b.chunk] = 1;
    }
  }
}

// Select each job's candidates for its level: each thread classifies 8
// candidates (strided by the block size), the block scans its threads' counts
// and reserves its output range with one atomic (a per-wave atomic on the
// job's one counter serialises).
__global__ __launch_bounds__(256) void k_band_select_jobs(const BandJob* __restrict__ jobs, int n_jobs) {
  constexpr int ITEMS = 8;
  __shared__ uint32_t wsum[4], wbase;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int j = 0; j < n_jobs; ++j) {
    const BandJob& b = jobs[j];
    const int64_t n = *b.ccount;
    // block-uniform loop: every thread reaches the barriers
    for (int64_t b0 = (int64_t)blockIdx.x * 256 * ITEMS; b0 < n; b0 += (int64_t)gridDim.x * 256 * ITEMS) {
      uint32_t keys[ITEMS];
      uint32_t mask = 0;
#pragma unroll
      for (int it = 0; it < ITEMS; ++it) {
        const int64_t i = b0 + it * 256 + threadIdx.x;
        keys[it] = 0;
        if (i < n) {
          const uint32_t k = b.cand[i];
          keys[it] = k;
          if (fabsf(b.vol[k]) < b.delta) {
            // grid indices are below 2^32: a 32-bit division (chunks past 2^32 hold everything)
            const uint32_t c = k / (uint32_t)(b.chunk < 0xFFFFFFFFll ? b.chunk : 0xFFFFFFFFll);
            bool alive = true;
            for (int l = 1; l <= b.level; ++l) alive = alive && b.flags[(int64_t)(l - 1) * b.nchunks + c];
            if (alive) mask |= 1u << it;
          }
        }
      }
      const uint32_t cnt = __popc(mask);
      uint32_t incl = cnt;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const uint32_t x = __shfl_up(incl, o);
        if (lane >= o) incl += x;
      }
      if (lane == 63) wsum[w] = incl;
      __syncthreads();
      uint32_t off = 0, tot = 0;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        off += k < w ? wsum[k] : 0u;
        tot += wsum[k];
      }
      if (threadIdx.x == 0) wbase = tot ? atomicAdd(b.lcount, tot) : 0u;
      __syncthreads();
      uint32_t pos = wbase + off + incl - cnt;
#pragma unroll
      for (int it = 0; it < ITEMS; ++it)
        if (mask & (1u << it)) b.list[pos++] = keys[it];
      __syncthreads();  // wsum / wbase are reused by the next iteration
    }
  }
}

template <class S, bool HALF>
static int run_level_stage(mp_batch* p, int stage, hipStream_t s) {
  if (p->n_level_jobs[stage][HALF] == 0) return MP_OK;
  constexpr size_t lds = jobs_lds_bytes<S, HALF>();
  static_assert(lds <= 160 * 1024, "LDS budget of one workgroup");
  static std::atomic<uint64_t> attr_devs{0};
  MP_HIP(set_lds_once(attr_devs, reinterpret_cast<const void*>(&k_level_jobs<S, HALF>), (int)lds));
  hipLaunchKernelGGL((k_level_jobs<S, HALF>), dim3(p->blocks), dim3(256), lds, s,
                     reinterpret_cast<const LevelJob*>(p->level_tab[stage][HALF]), p->n_level_jobs[stage][HALF],
                     &p->counters[2 * stage + HALF]);
  MP_LAUNCHED("k_level_jobs");
  return MP_OK;
}

// Level `stage` of every job: the default tile shape's jobs, then the halved one's.
template <class S>
static int run_level_both(mp_batch* p, int stage, hipStream_t s) {
  const int rc = run_level_stage<S, false>(p, stage, s);
  return rc != MP_OK ? rc : run_level_stage<S, true>(p, stage, s);
}

static int run_level(mp_batch* p, int stage, hipStream_t s) {
  switch (p->shape) {
    case SH_32x32x16_BF16: return run_level_both<S32x32x16_bf16>(p, stage, s);
    case SH_32x32x16_F16: return run_level_both<S32x32x16_f16>(p, stage, s);
    case SH_16x16x4_F32: return run_level_both<S16x16x4_f32>(p, stage, s);
  }
  set_error("mp_extract_batch_run: unsupported fragment shape %d", p->shape);
  return MP_EUNSUPPORTED;
}

static inline float voxel_h(int N) { return (float)(2.0 / (double)(N - 1)); }

}  // namespace mp

using namespace mp;

extern "C" {

void mp_extract_batch_destroy(mp_batch* p) {
  if (!p) return;
  if (p->mem) (void)hipFree(p->mem);
  delete p;
}

int mp_extract_batch_plan(const mp_extract_job* jobs, int n_jobs, void* stream, mp_batch** out) {
  MP_REQUIRE(out != nullptr, MP_EINVAL, "mp_extract_batch_plan: out is NULL");
  *out = nullptr;
  MP_REQUIRE(jobs != nullptr && n_jobs >= 1, MP_EINVAL, "mp_extract_batch_plan: no jobs");
  hipStream_t s = as_stream(stream);
  const int shape = jobs[0].stack ? jobs[0].stack->shape : -1;
  MP_REQUIRE(!jobs[0].stack || (jobs[0].stack->precision != MP_BF16X3 && jobs[0].stack->precision != MP_BF16X3F),
             MP_EUNSUPPORTED, "mp_extract_batch_plan: %s",
             "three-term BF16 (MP_BF16X3, MP_BF16X3F) is not served by the batched extraction");
  int max_levels = 0;
  for (int j = 0; j < n_jobs; ++j) {
    const mp_extract_job& J = jobs[j];
    MP_REQUIRE(J.stack != nullptr && J.vol != nullptr, MP_EINVAL, "mp_extract_batch_plan: job %d: NULL stack or vol", j);
    MP_REQUIRE(J.stack->shape == shape, MP_EINVAL,
               "mp_extract_batch_plan: job %d: every stack must have the same precision", j);
    MP_REQUIRE(J.N >= 2 && J.N <= 1625, MP_EINVAL, "mp_extract_batch_plan: job %d: N must be in [2, 1625]", j);
    MP_REQUIRE(J.grid_mode == MP_GRID_REFERENCE || J.grid_mode == MP_GRID_LATTICE, MP_EINVAL,
               "mp_extract_batch_plan: job %d: unknown grid mode %d", j, J.grid_mode);
    MP_REQUIRE(J.chunk >= 1, MP_EINVAL, "mp_extract_batch_plan: job %d: chunk must be >= 1", j);
    MP_REQUIRE(J.n_deltas == 0 || J.n_deltas == J.stack->n_levels - 1, MP_EINVAL,
               "mp_extract_batch_plan: job %d: need 0 or %d deltas, got %d", j, J.stack->n_levels - 1, J.n_deltas);
    max_levels = J.stack->n_levels > max_levels ? J.stack->n_levels : max_levels;
  }
  mp_batch* p = new mp_batch();
  p->shape = shape;
  p->n_jobs = n_jobs;
  p->max_levels = max_levels;

  // device layout: tables | counters | flags | per job: candidates [N^3], list [N^3]
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  size_t off = 0;
  size_t off_ltab[MP_MAX_LEVELS][2], off_btab[MP_MAX_LEVELS];
  for (int l = 0; l < max_levels; ++l) {
    off_ltab[l][0] = off; off += al(sizeof(LevelJob) * n_jobs);
    off_ltab[l][1] = off; off += al(sizeof(LevelJob) * n_jobs);
    off_btab[l] = off; off += al(sizeof(BandJob) * n_jobs);
  }
  const size_t off_ctr = off;
  p->n_counters = 2 * MP_MAX_LEVELS + (int64_t)n_jobs * MP_MAX_LEVELS;
  off += al(p->n_counters * 4);
  const size_t off_flags = off;
  std::vector<size_t> flag_off(n_jobs), cand_off(n_jobs), list_off(n_jobs);
  std::vector<int64_t> nchunks(n_jobs);
  for (int j = 0; j < n_jobs; ++j) {
    const int64_t n = (int64_t)jobs[j].N * jobs[j].N * jobs[j].N;
    nchunks[j] = (n + jobs[j].chunk - 1) / jobs[j].chunk;
    flag_off[j] = off - off_flags;
    off += (size_t)(jobs[j].stack->n_levels - 1) * nchunks[j];
  }
  p->n_flags = (int64_t)(off - off_flags);
  off = al(off);
  for (int j = 0; j < n_jobs; ++j) {
    const int64_t n = (int64_t)jobs[j].N * jobs[j].N * jobs[j].N;
    const bool culled = jobs[j].n_deltas > 0;
    cand_off[j] = off; off += culled ? al((size_t)n * 4) : 0;
    list_off[j] = off; off += culled ? al((size_t)n * 4) : 0;
  }
  hipError_t e = hipMalloc(&p->mem, off > 0 ? off : 256);
  if (e != hipSuccess) {
    delete p;
    set_error("mp_extract_batch_plan: hipMalloc(%zu) failed: %s", off, hipGetErrorString(e));
    return MP_ENOMEM;
  }
  p->mem_bytes = off;
  char* base = reinterpret_cast<char*>(p->mem);
  p->counters = reinterpret_cast<uint32_t*>(base + off_ctr);
  p->flags = reinterpret_cast<uint8_t*>(base + off_flags);
  uint32_t* counts = p->counters + 2 * MP_MAX_LEVELS;  // [j][0] candidates, [j][l] level-l selection

  std::vector<std::vector<LevelJob>> ltab[2] = {std::vector<std::vector<LevelJob>>(max_levels),
                                               std::vector<std::vector<LevelJob>>(max_levels)};
  std::vector<std::vector<BandJob>> btab(max_levels);
  for (int j = 0; j < n_jobs; ++j) {
    const mp_extract_job& J = jobs[j];
    const int L = J.stack->n_levels;
    const int64_t n = (int64_t)J.N * J.N * J.N;
    const bool culled = J.n_deltas > 0;
    p->grid_points.push_back(n);
    p->levels.push_back(L);
    p->culled.push_back(culled ? 1 : 0);
    float dmax = 0.f;
    for (int i = 0; i < J.n_deltas; ++i) dmax = J.deltas[i] > dmax ? J.deltas[i] : dmax;
    uint32_t* cand = reinterpret_cast<uint32_t*>(base + cand_off[j]);
    uint32_t* list = reinterpret_cast<uint32_t*>(base + list_off[j]);
    for (int l = 0; l < L; ++l) {
      LevelJob lj{};
      lj.L = J.stack->lv[l].dev;
      EvalTarget& t = lj.t;
      t.N = J.N;
      t.grid_mode = J.grid_mode;
      t.h = voxel_h(J.N);
      t.out = J.vol;
      t.accumulate = l > 0;
      if (l == 0 || !culled) {
        t.src = SRC_GRID;
        t.first = 0;
        t.n = n;
        if (l == 0 && culled) {
          t.band = dmax;
          t.cand = cand;
          t.cand_count = &counts[j * MP_MAX_LEVELS + 0];
        }
      } else {
        t.src = SRC_LIST;
        t.list = list;
        t.list_count = &counts[j * MP_MAX_LEVELS + l];
        t.n = n;
        BandJob bj{};
        bj.cand = cand;
        bj.ccount = &counts[j * MP_MAX_LEVELS + 0];
        bj.vol = J.vol;
        bj.delta = J.deltas[l - 1];
        bj.chunk = J.chunk;
        bj.nchunks = nchunks[j];
        bj.flags = p->flags + flag_off[j];
        bj.level = l;
        bj.list = list;
        bj.lcount = &counts[j * MP_MAX_LEVELS + l];
        btab[l].push_back(bj);
      }
      // the stack's tile shape for the batched kernel (1: default, 2: half)
      ltab[J.stack->batch_tiles == 2 ? 1 : 0][l].push_back(lj);
    }
  }
  for (int l = 0; l < max_levels; ++l) {
    p->band_tab[l] = base + off_btab[l];
    p->n_band_jobs[l] = (int)btab[l].size();
    e = hipSuccess;
    for (int h = 0; h < 2; ++h) {
      p->level_tab[l][h] = base + off_ltab[l][h];
      p->n_level_jobs[l][h] = (int)ltab[h][l].size();
      if (e == hipSuccess && !ltab[h][l].empty())
        e = hipMemcpyAsync(p->level_tab[l][h], ltab[h][l].data(), sizeof(LevelJob) * ltab[h][l].size(),
                           hipMemcpyHostToDevice, s);
    }
    int rc = MP_OK;
    if (p->n_level_jobs[l][0] + p->n_level_jobs[l][1] > 0) {
      if (e == hipSuccess && !btab[l].empty())
        e = hipMemcpyAsync(p->band_tab[l], btab[l].data(), sizeof(BandJob) * btab[l].size(),
                           hipMemcpyHostToDevice, s);
      if (e == hipSuccess) e = hipStreamSynchronize(s);  // the host tables go out of scope
      if (e != hipSuccess) {
        set_error("mp_extract_batch_plan: table upload failed: %s", hipGetErrorString(e));
        rc = MP_EHIP;
      }
    }
    if (rc != MP_OK) {
      mp_extract_batch_destroy(p);
      return rc;
    }
  }
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
    mp_extract_batch_destroy(p);
    set_error("mp_extract_batch_plan: device query failed");
    return MP_EHIP;
  }
  p->blocks = cus;  // one persistent workgroup per CU (the level stage fills the LDS)
  *out = p;
  return MP_OK;
}

int mp_extract_batch_run(mp_batch* p, void* stream) {
  MP_REQUIRE(p != nullptr, MP_EINVAL, "mp_extract_batch_run: plan is NULL");
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(k_fill_words, dim3(1), dim3(256), 0, s, p->counters, p->n_counters, 0u);
  MP_LAUNCHED("k_fill_words");
  if (p->n_flags > 0) {
    hipLaunchKernelGGL(k_fill_bytes, dim3(p->blocks), dim3(256), 0, s, p->flags, p->n_flags, (uint8_t)0);
    MP_LAUNCHED("k_fill_bytes");
  }
  int rc = run_level(p, 0, s);
  if (rc != MP_OK) return rc;
  for (int l = 1; l < p->max_levels; ++l) {
    if (p->n_band_jobs[l] > 0) {
      hipLaunchKernelGGL(k_band_mark_jobs, dim3(p->blocks * 4), dim3(256), 0, s,
                         reinterpret_cast<const BandJob*>(p->band_tab[l]), p->n_band_jobs[l]);
      MP_LAUNCHED("k_band_mark_jobs");
      hipLaunchKernelGGL(k_band_select_jobs, dim3(p->blocks * 4), dim3(256), 0, s,
                         reinterpret_cast<const BandJob*>(p->band_tab[l]), p->n_band_jobs[l]);
      MP_LAUNCHED("k_band_select_jobs");
    }
    if (p->n_level_jobs[l][0] + p->n_level_jobs[l][1] > 0) {
      rc = run_level(p, l, s);
      if (rc != MP_OK) return rc;
    }
  }
  return MP_OK;
}

int mp_extract_batch_counts(const mp_batch* p, int64_t* counts, void* stream) {
  MP_REQUIRE(p != nullptr && counts != nullptr, MP_EINVAL, "mp_extract_batch_counts: NULL argument");
  hipStream_t s = as_stream(stream);
  std::vector<uint32_t> h((size_t)p->n_jobs * MP_MAX_LEVELS);
  MP_HIP(hipMemcpyAsync(h.data(), p->counters + 2 * MP_MAX_LEVELS, h.size() * 4, hipMemcpyDeviceToHost, s));
  MP_HIP(hipStreamSynchronize(s));
  for (int j = 0; j < p->n_jobs; ++j) {
    for (int l = 0; l < MP_MAX_LEVELS; ++l) {
      int64_t v = 0;
      if (l < p->levels[j]) v = (l == 0 || !p->culled[j]) ? p->grid_points[j] : (int64_t)h[(size_t)j * MP_MAX_LEVELS + l];
      counts[(size_t)j * MP_MAX_LEVELS + l] = v;
    }
  }
  return MP_OK;
}

}  // extern "C"
