// Device helpers shared by the tracers: level evaluation on the points of a
// wave tile and ray generation (InitializeRays_kernel,
// renderer/cuda/kernels.cu:400-487).
#pragma once

#include "eval_dispatch.h"
#include "level_core.h"
#include "trace.h"

namespace mp {

// LDS handoff bytes per wave: every width the tracer runs (NC * width <= 512)
// fits this region.
constexpr size_t TRACE_HAND_BYTES = 32 * 1024;

// Shape of the 512-wide levels of an S stack: the same, except that a
// three-term BF16 stack packs them in FP16 (their operand pairs would not fit
// a wave's registers).
template <class S>
struct WideShape {
  typedef S type;
};
template <>
struct WideShape<S32x32x16_bf16x3> {
  typedef S32x32x16_f16 type;
};
template <>
struct WideShape<S32x32x16_bf16x3f> {
  typedef S32x32x16_f16 type;
};

// f_level at the NC points of this lane; hand: the wave's handoff region.
// A 512-wide level runs one column block at a time (its operands fill the
// registers for one), and only the blocks whose bit is set in cols (a
// wave-uniform mask; the others keep f unchanged).
template <class S, int NC>
__device__ __forceinline__ void eval_level_any(const LevelDev& L, const float (&x)[NC][3], float (&f)[NC],
                                      typename S::Frag* hand, unsigned cols = ~0u) {
  static_assert(!S::SPLIT || NC == 1, "three-term BF16 tiles hold one column block");
  typedef typename WideShape<S>::type SW;
  switch (L.width) {
    case 64: LevelCore<S, 64, NC, false>::eval(L, x, f, nullptr, hand); break;
    case 128: LevelCore<S, 128, NC, false>::eval(L, x, f, nullptr, hand); break;
    case 256: LevelCore<S, 256, NC, false>::eval(L, x, f, nullptr, hand); break;
    default:
      if constexpr (NC == 1) {
        LevelCore<SW, 512, 1, false>::eval(L, x, f, nullptr, reinterpret_cast<typename SW::Frag*>(hand));
      } else {
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          if (!((cols >> c) & 1u)) continue;
          const float xc[1][3] = {{x[c][0], x[c][1], x[c][2]}};
          float fc[1];
          LevelCore<S, 512, 1, false>::eval(L, xc, fc, nullptr, hand);
          f[c] = fc[0];
        }
      }
      break;
  }
}

// Ray rid = camera * width * height + j * width + i: origin o where the ray
// enters the sphere of radius R (t0 along the ray; the eye itself when the eye
// is inside), unit direction d; inside = false when the ray never meets the
// sphere (o is then the eye).
__device__ inline void ray_gen(const float* cams, int width, int height, int pixel_center, int flip_y, float R,
                               int64_t rid, float o[3], float d[3], float& t0, bool& inside) {
  const int64_t per_cam = (int64_t)width * height;
  const int cam = (int)(rid / per_cam);
  const int64_t pix = rid % per_cam;
  const int i = (int)(pix % width), j = (int)(pix / width);
  const float* iv = cams + cam * 32;
  const float* ip = iv + 16;
  const float off = pixel_center ? 0.5f : 0.0f;
  const float nx = (2.0f * (i + off)) / width - 1.0f;
  float ny = (2.0f * (j + off)) / height - 1.0f;
  if (flip_y) ny = -ny;
  const float clip[4] = {nx, ny, -1.0f, 1.0f};
  float eye[4];
  for (int k = 0; k < 4; ++k)
    eye[k] = ip[0 * 4 + k] * clip[0] + ip[1 * 4 + k] * clip[1] + ip[2 * 4 + k] * clip[2] + ip[3 * 4 + k] * clip[3];
  float world[3];
  for (int k = 0; k < 3; ++k)
    world[k] = iv[0 * 4 + k] * eye[0] + iv[1 * 4 + k] * eye[1] + iv[2 * 4 + k] * eye[2] + iv[12 + k];
  o[0] = iv[12]; o[1] = iv[13]; o[2] = iv[14];
  d[0] = world[0] - o[0]; d[1] = world[1] - o[1]; d[2] = world[2] - o[2];
  const float inv = 1.0f / sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  d[0] *= inv; d[1] *= inv; d[2] *= inv;
  // advance to the bounding sphere
  const float b = o[0] * d[0] + o[1] * d[1] + o[2] * d[2];
  const float c = o[0] * o[0] + o[1] * o[1] + o[2] * o[2] - R * R;
  t0 = 0.f;
  inside = true;
  if (c > 0.f) {
    const float disc = b * b - c;
    if (b < 0.f && disc > 0.f) {
      t0 = -b - sqrtf(disc);
      o[0] += t0 * d[0]; o[1] += t0 * d[1]; o[2] += t0 * d[2];
    } else {
      inside = false;  // the ray never meets the bounding sphere
    }
  }
}

__device__ inline void ray_init(const TraceArgs& a, int64_t rid, float o[3], float d[3], float& t0, bool& inside) {
  ray_gen(a.cams, a.width, a.height, a.pixel_center, a.flip_y, a.bound_radius, rid, o, d, t0, inside);
}

// RGBA8 with rounding to nearest (the persistent tracer's images).
__device__ inline uint32_t rgba8(float r, float g, float b) {
  auto q = [](float v) { return (uint32_t)(fminf(fmaxf(v, 0.f), 1.f) * 255.f + 0.5f); };
  return q(r) | (q(g) << 8) | (q(b) << 16) | (255u << 24);
}


}  // namespace mp
