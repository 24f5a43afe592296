// mp_extract_volume_noalloc: the band-culled extraction of mp_extract_volume
// with its work lists in a fixed device array of this library instead of
// memory allocated at call time, for callers that cannot allocate (code under
// graph capture, serving loops, harnesses that forbid hipMalloc).
//
// `scratch` is a host-side stand-in for a per-device array: host code never
// passes its name to a kernel.  The first call on a device (with that device
// current) resolves the device's own copy with hipGetSymbolAddress and keeps
// the address; every later call on that device uses the kept address.
#include <mutex>

#include "eval_dispatch.h"

namespace mp {
namespace {

// counters (256 B) | chunk flags (at most 1 MiB) | candidate list | band list
constexpr size_t NOALLOC_SCRATCH_BYTES = MP_NOALLOC_SCRATCH_BYTES;
static_assert(NOALLOC_SCRATCH_BYTES >= 256 + (1u << 20) + (size_t)MP_NOALLOC_MAX_POINTS * 8,
              "scratch holds the counters, the flags and both lists at the maximum size");

__device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mutex;
void* g_scratch_ptr[MAX_DEVICES] = {};

// The device address of `scratch` on the current device (resolved once per device).
int scratch_on_current_device(void** out) {
  int dev = -1;
  MP_HIP(hipGetDevice(&dev));
  MP_REQUIRE(dev >= 0 && dev < MAX_DEVICES, MP_EUNSUPPORTED,
             "mp_extract_volume_noalloc: device id %d outside [0, %d)", dev, MAX_DEVICES);
  std::lock_guard<std::mutex> lock(g_mutex);
  if (g_scratch_ptr[dev] == nullptr) {
    void* ptr = nullptr;
    MP_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    MP_REQUIRE(ptr != nullptr, MP_EHIP, "mp_extract_volume_noalloc: no device address for the scratch array");
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return MP_OK;
}

}  // namespace
}  // namespace mp

extern "C" int mp_extract_volume_noalloc(mp_stack* st, int N, int grid_mode, int64_t chunk, const float* deltas,
                                         int n_deltas, float* vol, int64_t* level_counts, void* stream) {
  MP_REQUIRE(N >= 2 && (int64_t)N * N * N <= MP_NOALLOC_MAX_POINTS, MP_EINVAL,
             "mp_extract_volume_noalloc: N = %d exceeds the fixed scratch (at most %d grid points, N <= 256)", N,
             MP_NOALLOC_MAX_POINTS);
  MP_REQUIRE(st != nullptr, MP_EINVAL, "mp_extract_volume_noalloc: stack is NULL");
  MP_REQUIRE(chunk >= 1, MP_EINVAL, "mp_extract_volume_noalloc: chunk must be >= 1");
  const int64_t n = (int64_t)N * N * N;
  const int64_t nchunks = (n + chunk - 1) / chunk;
  MP_REQUIRE((int64_t)(st->n_levels - 1) * nchunks <= (1 << 20), MP_EINVAL,
             "mp_extract_volume_noalloc: %lld chunks x %d residual levels exceed the fixed scratch's 2^20 chunk "
             "flags (use a larger chunk)",
             (long long)nchunks, st->n_levels - 1);
  void* ptr = nullptr;
  int rc = mp::scratch_on_current_device(&ptr);
  if (rc != MP_OK) return rc;
  return mp::extract_volume(st, N, grid_mode, chunk, deltas, n_deltas, vol, level_counts, stream, ptr,
                            mp::NOALLOC_SCRATCH_BYTES, "mp_extract_volume_noalloc");
}
