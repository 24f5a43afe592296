// Band-culled volume evaluation: the sampling loop of create_mesh_multistage
// (i3d/meshing.py:203-238) without the 2 GB coordinate tensor, the per-chunk
// host synchronisations or the masked gathers.
//
// Semantics kept from the reference loop, per chunk of `chunk` consecutive
// linear sample indices (max_batch, reconstruct.py:128):
//   s = f_0(x)
//   for level i >= 1:  m = |s| < delta_{i-1};  if no sample of the chunk has m:
//                      break (skip this and all finer levels for the chunk);
//                      s[m] += f_i(x[m])
// Pipeline:
//   1. level 0 over the whole grid; store s; append every index with
//      |s| < max(delta) to a candidate list (wave ballot + one atomic per wave).
//      Samples outside that band can never be in any level's mask.
//   2. per level i: over the candidates, mark the chunks that hold a sample
//      with |s| < delta_{i-1}; then select the candidates with
//      |s| < delta_{i-1} whose chunk is marked at every level 1..i;
//      evaluate f_i on the selection and add it in place.
#include "eval_dispatch.h"

namespace mp {

__global__ void k_fill_u32(uint32_t* p, int64_t n, uint32_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_u8(uint8_t* p, int64_t n, uint8_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// any[c] = 1 if some candidate of chunk c has |s| < delta
__global__ void k_band_mark(const uint32_t* cand, const uint32_t* ccount, const float* vol,
                            float delta, uint32_t chunk, uint8_t* any) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)*ccount) return;
  const uint32_t k = cand[i];
  if (fabsf(vol[k]) < delta) any[k / chunk] = 1;
}

// Select the candidates evaluated by level `level`.  Each thread classifies
// BAND_ITEMS candidates (strided by the block size, so the reads coalesce);
// the block scans its threads' counts and reserves its output range with a
// single atomic (one per BAND_ITEMS x 256 candidates: a per-wave atomic on the
// one counter serialises).  Grid indices are below 2^32 (N <= 1625).
constexpr int BAND_ITEMS = 8;

__global__ __launch_bounds__(256) void k_band_select(const uint32_t* __restrict__ cand,
                                                     const uint32_t* __restrict__ ccount,
                                                     const float* __restrict__ vol, float delta, uint32_t chunk,
                                                     const uint8_t* __restrict__ any, int64_t nchunks, int level,
                                                     uint32_t* __restrict__ list, uint32_t* __restrict__ lcount) {
  __shared__ uint32_t wsum[4], wbase;
  const int64_t n = (int64_t)*ccount;
  const int64_t b0 = (int64_t)blockIdx.x * 256 * BAND_ITEMS;
  if (b0 >= n) return;  // workgroup-uniform
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t keys[BAND_ITEMS];
  uint32_t mask = 0;
#pragma unroll
  for (int it = 0; it < BAND_ITEMS; ++it) {
    const int64_t i = b0 + it * 256 + threadIdx.x;
    keys[it] = 0;
    if (i < n) {
      const uint32_t k = cand[i];
      keys[it] = k;
      if (fabsf(vol[k]) < delta) {
        const uint32_t c = k / chunk;
        bool alive = true;
        for (int j = 1; j <= level; ++j) alive = alive && any[(int64_t)(j - 1) * nchunks + c];
        if (alive) mask |= 1u << it;
      }
    }
  }
  // exclusive scan of the thread counts over the block
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
  if (threadIdx.x == 0) wbase = tot ? atomicAdd(lcount, tot) : 0u;
  __syncthreads();
  uint32_t pos = wbase + off + incl - cnt;
#pragma unroll
  for (int it = 0; it < BAND_ITEMS; ++it)
    if (mask & (1u << it)) list[pos++] = keys[it];
}

__global__ void k_grid_coords(int N, int mode, float h, int64_t first, int64_t n, float* xyz) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float c[3];
  grid_coord(first + i, N, mode, h, c);
  xyz[3 * i + 0] = c[0];
  xyz[3 * i + 1] = c[1];
  xyz[3 * i + 2] = c[2];
}

static inline unsigned blocks_for(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

static int read_count(const uint32_t* d, int64_t* out, hipStream_t s) {
  uint32_t v = 0;
  MP_HIP(hipMemcpyAsync(&v, d, sizeof(v), hipMemcpyDeviceToHost, s));
  MP_HIP(hipStreamSynchronize(s));
  *out = v;
  return MP_OK;
}

// fl32(2 / (N - 1)): the Python float voxel size as torch multiplies it into a
// float32 tensor.
static inline float voxel_size(int N) { return (float)(2.0 / (double)(N - 1)); }

}  // namespace mp

using namespace mp;

extern "C" {

int mp_grid_coords(int N, int grid_mode, int64_t first, int64_t n, float* xyz, void* stream) {
  MP_REQUIRE(N >= 2, MP_EINVAL, "mp_grid_coords: N must be >= 2, got %d", N);
  MP_REQUIRE(grid_mode == MP_GRID_REFERENCE || grid_mode == MP_GRID_LATTICE, MP_EINVAL,
             "mp_grid_coords: unknown grid mode %d", grid_mode);
  const int64_t total = (int64_t)N * N * N;
  MP_REQUIRE(first >= 0 && n >= 0 && first + n <= total, MP_EINVAL,
             "mp_grid_coords: range [%lld, %lld) outside [0, %lld)", (long long)first,
             (long long)(first + n), (long long)total);
  if (n == 0) return MP_OK;
  MP_REQUIRE(xyz != nullptr, MP_EINVAL, "mp_grid_coords: xyz is NULL");
  hipLaunchKernelGGL(k_grid_coords, dim3(blocks_for(n, 256)), dim3(256), 0, as_stream(stream), N,
                     grid_mode, voxel_size(N), first, n, xyz);
  MP_LAUNCHED("k_grid_coords");
  return MP_OK;
}

int mp_extract_volume(mp_stack* st, int N, int grid_mode, int64_t chunk, const float* deltas,
                      int n_deltas, float* vol, int64_t* level_counts, void* stream) {
  return mp::extract_volume(st, N, grid_mode, chunk, deltas, n_deltas, vol, level_counts, stream, nullptr, 0,
                            "mp_extract_volume");
}

}  // extern "C"

namespace mp {

size_t extract_scratch_bytes(int n_levels, int64_t n, int64_t chunk) {
  const int64_t nchunks = (n + chunk - 1) / chunk;
  const size_t off_any = 256;
  const size_t off_cand = (off_any + (size_t)(n_levels - 1) * nchunks + 255) & ~(size_t)255;
  return off_cand + (size_t)n * 8;
}

int extract_volume(mp_stack* st, int N, int grid_mode, int64_t chunk, const float* deltas, int n_deltas,
                   float* vol, int64_t* level_counts, void* stream, void* scratch, size_t scratch_cap,
                   const char* who) {
  MP_REQUIRE(st != nullptr, MP_EINVAL, "%s: stack is NULL", who);
  MP_REQUIRE(N >= 2 && N <= 1625, MP_EINVAL, "%s: N must be in [2, 1625], got %d", who, N);
  MP_REQUIRE(grid_mode == MP_GRID_REFERENCE || grid_mode == MP_GRID_LATTICE, MP_EINVAL,
             "%s: unknown grid mode %d", who, grid_mode);
  MP_REQUIRE(chunk >= 1, MP_EINVAL, "%s: chunk must be >= 1", who);
  MP_REQUIRE(vol != nullptr, MP_EINVAL, "%s: vol is NULL", who);
  MP_REQUIRE(n_deltas == 0 || n_deltas == st->n_levels - 1, MP_EINVAL,
             "%s: need 0 or %d deltas, got %d", who, st->n_levels - 1, n_deltas);
  MP_REQUIRE(n_deltas == 0 || deltas != nullptr, MP_EINVAL, "%s: deltas is NULL", who);
  hipStream_t s = as_stream(stream);
  const int64_t n = (int64_t)N * N * N;
  const float h = voxel_size(N);
  const int L = st->n_levels;

  EvalTarget t{};
  t.src = SRC_GRID;
  t.N = N;
  t.grid_mode = grid_mode;
  t.h = h;
  t.first = 0;
  t.n = n;
  t.out = vol;

  if (n_deltas == 0) {
    // every level everywhere (i3d/meshing.py:232-233)
    for (int l = 0; l < L; ++l) {
      t.accumulate = l > 0;
      int rc = launch_level(st, l, t, s);
      if (rc != MP_OK) return rc;
      if (level_counts) level_counts[l] = n;
    }
    return MP_OK;
  }

  float dmax = 0.f;
  for (int i = 0; i < n_deltas; ++i) dmax = deltas[i] > dmax ? deltas[i] : dmax;
  const int64_t nchunks = (n + chunk - 1) / chunk;
  const int64_t chunk32 = chunk < n ? chunk : n;  // grid indices (< 2^32) divide by it in 32 bits
  // scratch: counters (256 B) | any flags [L-1][nchunks] | candidates [n] | list [n]
  const size_t off_any = 256;
  const size_t off_cand = (off_any + (size_t)(L - 1) * nchunks + 255) & ~(size_t)255;
  const size_t off_list = off_cand + (size_t)n * 4;
  const size_t need = extract_scratch_bytes(L, n, chunk);  // = off_list + n * 4
  char* base;
  if (scratch == nullptr) {
    int rc = ensure_scratch(st, need);
    if (rc != MP_OK) return rc;
    base = reinterpret_cast<char*>(st->scratch);
  } else {
    MP_REQUIRE(need <= scratch_cap, MP_EINVAL, "%s: needs %zu bytes of scratch, the caller's buffer holds %zu",
               who, need, scratch_cap);
    base = reinterpret_cast<char*>(scratch);
  }
  int rc;
  uint32_t* counters = reinterpret_cast<uint32_t*>(base);
  uint8_t* any = reinterpret_cast<uint8_t*>(base + off_any);
  uint32_t* cand = reinterpret_cast<uint32_t*>(base + off_cand);
  uint32_t* list = reinterpret_cast<uint32_t*>(base + off_list);

  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, counters, (int64_t)64, 0u);
  MP_LAUNCHED("k_fill_u32");
  hipLaunchKernelGGL(k_fill_u8, dim3(blocks_for((L - 1) * nchunks, 256)), dim3(256), 0, s, any,
                     (int64_t)(L - 1) * nchunks, (uint8_t)0);
  MP_LAUNCHED("k_fill_u8");

  // 1. coarse level over the grid, candidates appended in the epilogue
  t.accumulate = 0;
  t.band = dmax;
  t.cand = cand;
  t.cand_count = &counters[0];
  rc = launch_level(st, 0, t, s);
  if (rc != MP_OK) return rc;
  if (level_counts) level_counts[0] = n;
  int64_t ncand = 0;
  rc = read_count(&counters[0], &ncand, s);
  if (rc != MP_OK) return rc;

  // 2. residual levels
  for (int l = 1; l < L; ++l) {
    const float d = deltas[l - 1];
    uint8_t* any_l = any + (int64_t)(l - 1) * nchunks;
    uint32_t* lcount = &counters[l];
    int64_t nsel = 0;
    if (ncand > 0) {
      hipLaunchKernelGGL(k_band_mark, dim3(blocks_for(ncand, 256)), dim3(256), 0, s, cand,
                         &counters[0], vol, d, (uint32_t)chunk32, any_l);
      MP_LAUNCHED("k_band_mark");
      hipLaunchKernelGGL(k_band_select, dim3(blocks_for(ncand, 256 * BAND_ITEMS)), dim3(256), 0, s, cand,
                         &counters[0], vol, d, (uint32_t)chunk32, any, nchunks, l, list, lcount);
      MP_LAUNCHED("k_band_select");
      rc = read_count(lcount, &nsel, s);
      if (rc != MP_OK) return rc;
    }
    if (level_counts) level_counts[l] = nsel;
    if (nsel == 0) continue;
    EvalTarget tl{};
    tl.src = SRC_LIST;
    tl.N = N;
    tl.grid_mode = grid_mode;
    tl.h = h;
    tl.list = list;
    tl.list_count = lcount;
    tl.n = nsel;
    tl.out = vol;
    tl.accumulate = 1;
    rc = launch_level(st, l, tl, s);
    if (rc != MP_OK) return rc;
  }
  return MP_OK;
}

}  // namespace mp
