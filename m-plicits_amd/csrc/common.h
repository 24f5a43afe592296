// Error plumbing shared by every translation unit of libmplicits_hip.so.
// All C entry points return MP_* status codes; the message of the last failure
// is kept per host thread and returned by mp_last_error().
#pragma once

#include <hip/hip_runtime.h>
#include <stdarg.h>

#include <atomic>
#include <stdint.h>
#include <stdio.h>

#include "../../include/mplicits.h"

namespace mp {

void set_error(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
const char* get_error();

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// hipFuncSetAttribute(MaxDynamicSharedMemorySize) once per kernel and device:
// `done` is the kernel's set of device ids (bit per id < 64) already configured.
inline hipError_t set_lds_once(std::atomic<uint64_t>& done, const void* fn, int bytes) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  const uint64_t bit = dev < 64 ? (1ull << dev) : 0ull;
  if (bit && (done.load(std::memory_order_relaxed) & bit)) return hipSuccess;
  e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  if (e == hipSuccess) done.fetch_or(bit);
  return e;
}

}  // namespace mp

// Evaluate a HIP call; on failure record the message and return MP_EHIP.
#define MP_HIP(call)                                                                   \
  do {                                                                                 \
    hipError_t mp_err_ = (call);                                                       \
    if (mp_err_ != hipSuccess) {                                                       \
      mp::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(mp_err_), __FILE__, \
                    __LINE__);                                                         \
      return MP_EHIP;                                                                  \
    }                                                                                  \
  } while (0)

// Check the launch that was just enqueued.
#define MP_LAUNCHED(name) MP_HIP(hipGetLastError())

#define MP_REQUIRE(cond, code, ...) \
  do {                              \
    if (!(cond)) {                  \
      mp::set_error(__VA_ARGS__);   \
      return (code);                \
    }                               \
  } while (0)
