// Arguments of the persistent sphere-tracing kernel.
#pragma once

#include "stack.h"

namespace mp {

// State of a ray parked by the coarse phase where it entered the coarse band
// (direction and entry distance are regenerated from the id).
struct ParkedRay {
  uint32_t rid;
  int32_t steps;
  float t;
  float p[3];
};

// Ray grouping of one k_trace launch.  TRACE_ALL: every ray with every level
// (single pass).  TRACE_COARSE: every ray with the coarse level only; a ray is
// parked, before its step, once its coarse value enters the coarse band.
// TRACE_BAND: the parked rays, with every level.  Coarse then band gives each
// ray the same evaluations and steps as the single pass; it only keeps the
// rays that need residual levels together in the tiles.
enum TracePhase { TRACE_ALL = 0, TRACE_COARSE = 1, TRACE_BAND = 2 };

struct TraceArgs {
  const float* cams;  // [n_cams][32]: inverse view then inverse projection, column-major 4x4
  int width, height;
  int64_t total_rays;
  int max_steps;
  float hit_eps, bound_radius, band_scale;
  int pixel_center, flip_y;
  int n_levels;
  int cols;                      // column blocks per wave tile (batched tracer: per job)
  LevelDev levels[MP_MAX_LEVELS];
  float deltas[MP_MAX_LEVELS];
  unsigned long long* next_ray;  // global ray queue head
  unsigned long long* stats;     // [n_levels] evaluations per level
  float* depth;                  // [rays] distance from the camera, -1 for a miss
  int32_t* steps;                // [rays]
  float* hit_pos;                // [rays][3] (scratch)
  uint32_t* rgba;                // [rays] RGBA8
  float* normals;                // [rays][3] or NULL
  int phase;                     // TracePhase
  ParkedRay* park;               // [rays] parked rays (TRACE_COARSE writes, TRACE_BAND reads)
  unsigned long long* n_parked;  // parked-ray count
};

}  // namespace mp
