// Per-level evaluation launches shared by mp_eval, the extraction pipeline and
// the training forward.
#pragma once

#include "stack.h"

namespace mp {

enum EvalSrc { SRC_POINTS = 0, SRC_GRID = 1, SRC_LIST = 2 };

// Where the points of one level launch come from and where f goes.
struct EvalTarget {
  int src;                       // EvalSrc
  const float* xyz;              // SRC_POINTS: [n][3]
  int N, grid_mode;              // SRC_GRID / SRC_LIST: grid size and coordinate mode
  float h;                       // grid spacing fl32(2 / (N - 1))
  int64_t first;                 // SRC_GRID: first grid index
  const uint32_t* list;          // SRC_LIST: grid indices
  const uint32_t* list_count;    // SRC_LIST: device count (n is a host upper bound)
  int64_t n;                     // points (or upper bound for SRC_LIST)
  float* out;                    // SRC_POINTS: out[i]; grid sources: out[grid index]
  int accumulate;                // out = out + f instead of out = f
  float band;                    // > 0: append the key of every point with |out| < band
  uint32_t* cand;
  uint32_t* cand_count;
};

int launch_level(const mp_stack* st, int level, const EvalTarget& t, hipStream_t stream);


// f_level and its input gradient at n points xyz [n][3]; f may be NULL;
// accumulate adds into f / g instead of overwriting.  With n_dev the point
// count is read on the device (n is then an upper bound: the launch size).
int launch_level_grad(const mp_stack* st, int level, const float* xyz, int64_t n, float* f, float* g,
                      int accumulate, hipStream_t stream, const unsigned long long* n_dev = nullptr);

// The same in forward mode (four MFMA columns per point: value and three
// tangents; level.hip), for widths whose reverse-mode kernel would spill.
int launch_level_tangent(int shape, const LevelDev& L, const float* xyz, int64_t n, float* f, float* g,
                         int accumulate, hipStream_t stream, const unsigned long long* n_dev = nullptr);

// Input gradient of sum_{l=lvl0}^{lvl1-1} f_l at n points (overwrites grad),
// and the sum itself into sdf unless it is NULL (the gradient kernels run each
// level's forward anyway).
int launch_grad(const mp_stack* st, int lvl0, int lvl1, const float* xyz, int64_t n, float* sdf, float* grad,
                hipStream_t stream);

// The band-culled extraction pipeline behind mp_extract_volume.  scratch ==
// NULL: work lists in the stack's own (grown on demand) scratch; otherwise the
// caller's device buffer of scratch_cap bytes (a call needing more is
// rejected before any launch).  `who` names the entry point in messages.
int extract_volume(mp_stack* st, int N, int grid_mode, int64_t chunk, const float* deltas, int n_deltas,
                   float* vol, int64_t* level_counts, void* stream, void* scratch, size_t scratch_cap,
                   const char* who);
// Scratch bytes extract_volume needs for a culled extraction of n grid points.
size_t extract_scratch_bytes(int n_levels, int64_t n, int64_t chunk);


}  // namespace mp
