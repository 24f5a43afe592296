// Arguments of the persistent sphere-tracing kernel.
#pragma once

#include "stack.h"

namespace mp {

// State of a ray parked by one launch for the next (direction and entry
// distance are regenerated from the id).
struct ParkedRay {
  uint32_t rid;
  int32_t steps;
  float t;
  float p[3];
};

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
  // Two-launch tracing (eval_levels > 0): the first launch (eval_levels 1)
  // evaluates the coarse level and parks, before its step, every ray that
  // enters the coarse band; the second (eval_levels = n_levels) traces the
  // parked rays on all levels.  Each ray gets the same evaluations and steps
  // as in the single pass (eval_levels 0: every level, the ray queue, no
  // parking); only the rays that need the residual levels share tiles.
  int eval_levels;
  const ParkedRay* park_in;               // rays of this launch (NULL: the ray queue)
  const unsigned long long* n_park_in;    // their count
  ParkedRay* park_out;                    // [rays] rays parked for the next launch, or NULL
  unsigned long long* n_park_out;         // their count
};

}  // namespace mp
