// Arguments of the persistent sphere-tracing kernel.
#pragma once

#include "stack.h"

namespace mp {

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
};

}  // namespace mp
