// Training-data generation on the GPU: the unsigned point-to-mesh distance
// that open3d's RaycastingScene.compute_distance gives the trainers, the
// per-vertex tubular-neighbourhood search built on it (calc_deltas,
// experiment_scripts/utils.py:46-89) and the band sampler
// (create_training_data_on_neighborhood, utils.py:92-169).
//
// Distance: a bounding-volume hierarchy over the triangles, built once on the
// host (median split of the triangle centroids along the longest axis, at
// most MESH_LEAF triangles per leaf, depth-first node order so the left child
// of node i is i + 1), then one thread per query point walks it nearest child
// first with a stack of (node, squared box distance) and prunes every box no
// closer than the best triangle so far.  The point-triangle distance is the
// Voronoi-region method of Ericson, Real-Time Collision Detection (2005),
// section 5.1.5.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.h"

namespace mp {

constexpr int MESH_LEAF = 4;
constexpr int MESH_STACK = 48;  // traversal stack entries; the build checks the depth

struct BvhNode {
  float lo[3];
  int32_t first;  // leaf: first triangle; inner: index of the right child (left = this + 1)
  float hi[3];
  int32_t count;  // leaf: triangle count (> 0); inner: 0
};

}  // namespace mp

struct mp_mesh {
  int64_t n_tris = 0, n_nodes = 0;
  int depth = 0;
  void* mem = nullptr;            // nodes then triangles (one allocation)
  mp::BvhNode* nodes = nullptr;   // device
  float4* tris = nullptr;         // device, [n_tris][3] (x, y, z, 0) per corner, leaf order
};

namespace mp {

namespace {

struct Builder {
  const float* v;
  const int32_t* f;
  std::vector<int32_t> order;
  std::vector<float> cen;  // [F][3]
  std::vector<BvhNode> nodes;
  int depth = 0;

  void tri_bounds(int32_t t, float lo[3], float hi[3]) const {
    for (int k = 0; k < 3; ++k) {
      lo[k] = INFINITY;
      hi[k] = -INFINITY;
    }
    for (int c = 0; c < 3; ++c) {
      const float* p = v + 3 * (int64_t)f[3 * (int64_t)t + c];
      for (int k = 0; k < 3; ++k) {
        lo[k] = std::min(lo[k], p[k]);
        hi[k] = std::max(hi[k], p[k]);
      }
    }
  }

  int32_t build(int64_t begin, int64_t end, int level) {
    depth = std::max(depth, level + 1);
    const int32_t id = (int32_t)nodes.size();
    nodes.push_back(BvhNode{});
    BvhNode nd{};
    float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int k = 0; k < 3; ++k) {
      nd.lo[k] = INFINITY;
      nd.hi[k] = -INFINITY;
    }
    for (int64_t i = begin; i < end; ++i) {
      float lo[3], hi[3];
      tri_bounds(order[i], lo, hi);
      for (int k = 0; k < 3; ++k) {
        nd.lo[k] = std::min(nd.lo[k], lo[k]);
        nd.hi[k] = std::max(nd.hi[k], hi[k]);
        const float c = cen[3 * (int64_t)order[i] + k];
        clo[k] = std::min(clo[k], c);
        chi[k] = std::max(chi[k], c);
      }
    }
    int axis = 0;
    for (int k = 1; k < 3; ++k)
      if (chi[k] - clo[k] > chi[axis] - clo[axis]) axis = k;
    if (end - begin <= MESH_LEAF || !(chi[axis] > clo[axis])) {
      nd.first = (int32_t)begin;
      nd.count = (int32_t)(end - begin);
      nodes[id] = nd;
      return id;
    }
    const int64_t mid = (begin + end) / 2;
    std::nth_element(order.begin() + begin, order.begin() + mid, order.begin() + end,
                     [&](int32_t a, int32_t b) { return cen[3 * (int64_t)a + axis] < cen[3 * (int64_t)b + axis]; });
    build(begin, mid, level + 1);
    const int32_t right = build(mid, end, level + 1);
    nd.first = right;
    nd.count = 0;
    nodes[id] = nd;
    return id;
  }
};

}  // namespace

__device__ inline float dot3(float ax, float ay, float az, float bx, float by, float bz) {
  return ax * bx + ay * by + az * bz;
}

// Squared distance from p to triangle (a, b, c): Ericson 2005, section 5.1.5
// (the closest point by Voronoi region: vertex, edge or face).
__device__ inline float point_tri_d2(float px, float py, float pz, float4 a, float4 b, float4 c) {
  const float abx = b.x - a.x, aby = b.y - a.y, abz = b.z - a.z;
  const float acx = c.x - a.x, acy = c.y - a.y, acz = c.z - a.z;
  const float apx = px - a.x, apy = py - a.y, apz = pz - a.z;
  float qx, qy, qz;
  const float d1 = dot3(abx, aby, abz, apx, apy, apz), d2 = dot3(acx, acy, acz, apx, apy, apz);
  if (d1 <= 0.f && d2 <= 0.f) {
    qx = a.x; qy = a.y; qz = a.z;
  } else {
    const float bpx = px - b.x, bpy = py - b.y, bpz = pz - b.z;
    const float d3 = dot3(abx, aby, abz, bpx, bpy, bpz), d4 = dot3(acx, acy, acz, bpx, bpy, bpz);
    const float cpx = px - c.x, cpy = py - c.y, cpz = pz - c.z;
    const float d5 = dot3(abx, aby, abz, cpx, cpy, cpz), d6 = dot3(acx, acy, acz, cpx, cpy, cpz);
    const float vc = d1 * d4 - d3 * d2, vb = d5 * d2 - d1 * d6, va = d3 * d6 - d5 * d4;
    if (d3 >= 0.f && d4 <= d3) {
      qx = b.x; qy = b.y; qz = b.z;
    } else if (d6 >= 0.f && d5 <= d6) {
      qx = c.x; qy = c.y; qz = c.z;
    } else if (vc <= 0.f && d1 >= 0.f && d3 <= 0.f) {
      const float t = d1 / (d1 - d3);
      qx = a.x + t * abx; qy = a.y + t * aby; qz = a.z + t * abz;
    } else if (vb <= 0.f && d2 >= 0.f && d6 <= 0.f) {
      const float t = d2 / (d2 - d6);
      qx = a.x + t * acx; qy = a.y + t * acy; qz = a.z + t * acz;
    } else if (va <= 0.f && (d4 - d3) >= 0.f && (d5 - d6) >= 0.f) {
      const float t = (d4 - d3) / ((d4 - d3) + (d5 - d6));
      qx = b.x + t * (c.x - b.x); qy = b.y + t * (c.y - b.y); qz = b.z + t * (c.z - b.z);
    } else {
      const float den = 1.f / (va + vb + vc);
      const float s = vb * den, t = vc * den;
      qx = a.x + abx * s + acx * t; qy = a.y + aby * s + acy * t; qz = a.z + abz * s + acz * t;
    }
  }
  const float dx = px - qx, dy = py - qy, dz = pz - qz;
  return dx * dx + dy * dy + dz * dz;
}

__device__ inline float box_d2(const BvhNode& n, float px, float py, float pz) {
  const float dx = fmaxf(fmaxf(n.lo[0] - px, px - n.hi[0]), 0.f);
  const float dy = fmaxf(fmaxf(n.lo[1] - py, py - n.hi[1]), 0.f);
  const float dz = fmaxf(fmaxf(n.lo[2] - pz, pz - n.hi[2]), 0.f);
  return dx * dx + dy * dy + dz * dz;
}

// Unsigned distance from p to the mesh.
__device__ float mesh_distance(const BvhNode* __restrict__ nodes, const float4* __restrict__ tris, float px, float py,
                               float pz) {
  float best = INFINITY;
  int32_t st_node[MESH_STACK];
  float st_d2[MESH_STACK];
  int sp = 0;
  int32_t node = 0;
  float nd2 = box_d2(nodes[0], px, py, pz);
  while (true) {
    if (nd2 < best) {
      const BvhNode n = nodes[node];
      if (n.count > 0) {
        for (int32_t t = n.first; t < n.first + n.count; ++t)
          best = fminf(best, point_tri_d2(px, py, pz, tris[3 * (int64_t)t], tris[3 * (int64_t)t + 1],
                                          tris[3 * (int64_t)t + 2]));
      } else {
        const int32_t l = node + 1, r = n.first;
        const float dl = box_d2(nodes[l], px, py, pz), dr = box_d2(nodes[r], px, py, pz);
        const bool left_first = dl <= dr;
        const int32_t near_n = left_first ? l : r, far_n = left_first ? r : l;
        const float near_d = left_first ? dl : dr, far_d = left_first ? dr : dl;
        if (far_d < best) {
          st_node[sp] = far_n;
          st_d2[sp] = far_d;
          ++sp;  // the build bounds the depth, so sp < MESH_STACK
        }
        node = near_n;
        nd2 = near_d;
        continue;
      }
    }
    if (sp == 0) break;
    --sp;
    node = st_node[sp];
    nd2 = st_d2[sp];
  }
  return sqrtf(best);
}

__global__ void k_mesh_distance(const BvhNode* __restrict__ nodes, const float4* __restrict__ tris,
                                const float* __restrict__ pts, int64_t n, float* __restrict__ dist) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  dist[i] = mesh_distance(nodes, tris, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
}

// calc_deltas, first part: out [V][7] = (x, y, z, n / |n|, delta0) and the
// step of every vertex = delta0 (utils.py:56-70; float32 operations in the
// reference's order, no contraction).
__global__ void k_deltas_init(const float* __restrict__ vert, int cols, int64_t V, float step0, float* __restrict__ out,
                              float* __restrict__ step) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= V) return;
  const float* v = vert + i * cols;
  const float nx = v[3], ny = v[4], nz = v[5];
  const float len = __fsqrt_rn(__fadd_rn(__fadd_rn(__fmul_rn(nx, nx), __fmul_rn(ny, ny)), __fmul_rn(nz, nz)));
  float* o = out + i * 7;
  o[0] = v[0]; o[1] = v[1]; o[2] = v[2];
  o[3] = __fdiv_rn(nx, len); o[4] = __fdiv_rn(ny, len); o[5] = __fdiv_rn(nz, len);
  o[6] = step0;
  step[i] = step0;
}

__device__ inline unsigned int float_key(float f) {
  const unsigned int b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

// calc_deltas, one iteration (utils.py:72-86): the distance of
// p + delta n to the mesh; where it differs from delta by more than eps the
// step halves and delta moves back by it, elsewhere delta moves on by the step.
// max_key receives the order-preserving key of the largest new delta.
__global__ void k_deltas_step(const BvhNode* __restrict__ nodes, const float4* __restrict__ tris,
                              const float* __restrict__ vert, int cols, int64_t V, float eps, float* __restrict__ out,
                              float* __restrict__ step, unsigned int* __restrict__ max_key) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  float nd = -INFINITY;
  if (i < V) {
    const float* v = vert + i * cols;
    float* o = out + i * 7;
    const float d = o[6];
    const float px = __fadd_rn(v[0], __fmul_rn(d, o[3]));
    const float py = __fadd_rn(v[1], __fmul_rn(d, o[4]));
    const float pz = __fadd_rn(v[2], __fmul_rn(d, o[5]));
    const float sdf = mesh_distance(nodes, tris, px, py, pz);
    const float diff = fabsf(__fsub_rn(d, sdf));
    float s = step[i];
    if (diff > eps) {
      s = __fmul_rn(s, 0.5f);
      nd = __fsub_rn(d, s);
    } else {
      nd = __fadd_rn(d, s);
    }
    step[i] = s;
    o[6] = nd;
  }
  // wave maximum, then one atomic per wave
  for (int off = 32; off > 0; off >>= 1) nd = fmaxf(nd, __shfl_xor(nd, off));
  if (lane == 0 && nd > -INFINITY) atomicMax(max_key, float_key(nd));
}

__global__ void k_key_reset(unsigned int* k) {
  if (threadIdx.x == 0) *k = 0u;
}

// create_training_data_on_neighborhood (utils.py:92-169) given its random
// draws: sampled vertex ids (NULL: every vertex in order) and the uniform
// numbers rnd [n][3].  Rows [0, n): the vertices, sdf 0; [n, 2n): offsets
// along the normal by delta_n * rnd_z (delta_n = the vertex's last column, or
// delta_eps), sdf delta_n * rnd_z; [2n, 3n) (dithering): p + delta / 2 n +
// delta rnd, sdf -1.  float32 operations in the reference's order.
__global__ void k_neigh_samples(const float* __restrict__ vert, int cols, const int64_t* __restrict__ idx,
                                const float* __restrict__ rnd, int64_t n, float delta, float delta_eps, int use_neigh,
                                int dither, float* __restrict__ pts, float* __restrict__ nrm, float* __restrict__ sdf) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t vi = idx ? idx[i] : i;
  const float* v = vert + vi * cols;
  const float p[3] = {v[0], v[1], v[2]}, q[3] = {v[3], v[4], v[5]};
  const float dn = use_neigh ? v[cols - 1] : delta_eps;
  const float rz = rnd[3 * i + 2];
  const float dr = __fmul_rn(dn, rz);
  const float half = __fmul_rn(0.5f, delta);
  for (int k = 0; k < 3; ++k) {
    pts[3 * i + k] = p[k];
    nrm[3 * i + k] = q[k];
    pts[3 * (n + i) + k] = __fadd_rn(p[k], __fmul_rn(dr, q[k]));
    nrm[3 * (n + i) + k] = q[k];
    if (dither) {
      pts[3 * (2 * n + i) + k] = __fadd_rn(__fmul_rn(rnd[3 * i + k], delta), __fadd_rn(p[k], __fmul_rn(half, q[k])));
      nrm[3 * (2 * n + i) + k] = q[k];
    }
  }
  sdf[i] = 0.f;
  sdf[n + i] = dr;
  if (dither) sdf[2 * n + i] = -1.f;
}

}  // namespace mp

using namespace mp;

extern "C" {

int mp_mesh_create(const float* verts, int64_t n_verts, const int32_t* faces, int64_t n_faces, void* stream,
                   mp_mesh** out) {
  MP_REQUIRE(out != nullptr, MP_EINVAL, "mp_mesh_create: out is NULL");
  *out = nullptr;
  MP_REQUIRE(verts && faces && n_verts > 0 && n_faces > 0, MP_EINVAL, "mp_mesh_create: empty mesh");
  MP_REQUIRE(n_faces < (1ll << 30), MP_EUNSUPPORTED, "mp_mesh_create: %lld triangles (at most 2^30)",
             (long long)n_faces);
  for (int64_t i = 0; i < 3 * n_faces; ++i)
    MP_REQUIRE(faces[i] >= 0 && faces[i] < n_verts, MP_EINVAL, "mp_mesh_create: face %lld indexes vertex %d of %lld",
               (long long)(i / 3), faces[i], (long long)n_verts);
  Builder b{verts, faces, {}, {}, {}, 0};
  b.order.resize(n_faces);
  b.cen.resize(3 * n_faces);
  for (int64_t t = 0; t < n_faces; ++t) {
    b.order[t] = (int32_t)t;
    for (int k = 0; k < 3; ++k) {
      float s = 0.f;
      for (int c = 0; c < 3; ++c) s += verts[3 * (int64_t)faces[3 * t + c] + k];
      b.cen[3 * t + k] = s / 3.f;
    }
  }
  b.nodes.reserve(2 * (n_faces / MESH_LEAF + 1));
  b.build(0, n_faces, 0);
  // the traversal holds at most one far child per level on its stack
  MP_REQUIRE(b.depth <= MESH_STACK, MP_EUNSUPPORTED, "mp_mesh_create: hierarchy depth %d exceeds %d", b.depth,
             MESH_STACK);
  std::vector<float4> tris(3 * n_faces);
  for (int64_t k = 0; k < n_faces; ++k)
    for (int c = 0; c < 3; ++c) {
      const float* p = verts + 3 * (int64_t)faces[3 * (int64_t)b.order[k] + c];
      tris[3 * k + c] = make_float4(p[0], p[1], p[2], 0.f);
    }
  mp_mesh* m = new mp_mesh();
  m->n_tris = n_faces;
  m->n_nodes = (int64_t)b.nodes.size();
  m->depth = b.depth;
  const size_t nb = ((size_t)m->n_nodes * sizeof(BvhNode) + 255) & ~(size_t)255;
  const size_t tb = (size_t)n_faces * 3 * sizeof(float4);
  hipError_t e = hipMalloc(&m->mem, nb + tb);
  if (e != hipSuccess) {
    delete m;
    set_error("mp_mesh_create: hipMalloc(%zu) failed: %s", nb + tb, hipGetErrorString(e));
    return MP_ENOMEM;
  }
  m->nodes = reinterpret_cast<BvhNode*>(m->mem);
  m->tris = reinterpret_cast<float4*>(reinterpret_cast<char*>(m->mem) + nb);
  hipStream_t s = as_stream(stream);
  e = hipMemcpyAsync(m->nodes, b.nodes.data(), (size_t)m->n_nodes * sizeof(BvhNode), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(m->tris, tris.data(), tb, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);  // the host arrays go out of scope
  if (e != hipSuccess) {
    (void)hipFree(m->mem);
    delete m;
    set_error("mp_mesh_create: upload failed: %s", hipGetErrorString(e));
    return MP_EHIP;
  }
  *out = m;
  return MP_OK;
}

void mp_mesh_destroy(mp_mesh* m) {
  if (!m) return;
  if (m->mem) (void)hipFree(m->mem);
  delete m;
}

int mp_mesh_info(const mp_mesh* m, int64_t* n_tris, int64_t* n_nodes, int* depth) {
  MP_REQUIRE(m != nullptr, MP_EINVAL, "mp_mesh_info: mesh is NULL");
  if (n_tris) *n_tris = m->n_tris;
  if (n_nodes) *n_nodes = m->n_nodes;
  if (depth) *depth = m->depth;
  return MP_OK;
}

int mp_mesh_distance(const mp_mesh* m, const float* pts, int64_t n, float* dist, void* stream) {
  MP_REQUIRE(m != nullptr && pts != nullptr && dist != nullptr && n >= 0, MP_EINVAL, "mp_mesh_distance: bad arguments");
  if (n == 0) return MP_OK;
  hipLaunchKernelGGL(k_mesh_distance, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, as_stream(stream), m->nodes,
                     m->tris, pts, n, dist);
  MP_LAUNCHED("k_mesh_distance");
  return MP_OK;
}

int mp_calc_deltas(const mp_mesh* m, const float* vert, int64_t V, int cols, float max_delta, int n_iters, float* out,
                   int* iters_done, void* stream) {
  MP_REQUIRE(m != nullptr && vert != nullptr && out != nullptr && V > 0 && cols >= 6 && n_iters >= 0, MP_EINVAL,
             "mp_calc_deltas: bad arguments");
  hipStream_t s = as_stream(stream);
  float* step = nullptr;
  unsigned int* key = nullptr;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(&step), (size_t)V * 4 + 256);
  if (e != hipSuccess) {
    set_error("mp_calc_deltas: scratch allocation failed: %s", hipGetErrorString(e));
    return MP_ENOMEM;
  }
  key = reinterpret_cast<unsigned int*>(reinterpret_cast<char*>(step) + (((size_t)V * 4 + 127) & ~(size_t)127));
  const float step0 = max_delta / (float)n_iters;
  const unsigned grid = (unsigned)((V + 255) / 256);
  hipLaunchKernelGGL(k_deltas_init, dim3(grid), dim3(256), 0, s, vert, cols, V, step0, out, step);
  int rc = MP_OK;
  e = hipGetLastError();
  // the reference's loop: while i < n_iters and max(deltas) < max_delta
  int i = 0;
  float cur_max = step0;
  while (e == hipSuccess && i < n_iters && cur_max < max_delta) {
    hipLaunchKernelGGL(k_key_reset, dim3(1), dim3(64), 0, s, key);
    hipLaunchKernelGGL(k_deltas_step, dim3(grid), dim3(256), 0, s, m->nodes, m->tris, vert, cols, V, 2e-4f, out, step,
                       key);
    unsigned int h = 0;
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(&h, key, 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    const uint32_t bits = (h & 0x80000000u) ? (h & 0x7FFFFFFFu) : ~h;  // inverse of float_key
    std::memcpy(&cur_max, &bits, 4);
    ++i;
  }
  if (e != hipSuccess) {
    set_error("mp_calc_deltas: %s", hipGetErrorString(e));
    rc = MP_EHIP;
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(step);
  if (iters_done) *iters_done = i;
  return rc;
}

int mp_neighborhood_samples(const float* vert, int64_t V, int cols, const int64_t* idx, const float* rnd, int64_t n,
                            float delta, float delta_eps, int use_neigh_delta, int dithering, float* pts,
                            float* normals, float* sdfs, void* stream) {
  MP_REQUIRE(vert && rnd && pts && normals && sdfs && V > 0 && cols >= 6 && n >= 0, MP_EINVAL,
             "mp_neighborhood_samples: bad arguments");
  MP_REQUIRE(idx != nullptr || n <= V, MP_EINVAL, "mp_neighborhood_samples: %lld rows without ids from %lld vertices",
             (long long)n, (long long)V);
  MP_REQUIRE(!use_neigh_delta || cols >= 7, MP_EINVAL,
             "mp_neighborhood_samples: neighbourhood deltas need a seventh column (mp_calc_deltas)");
  if (n == 0) return MP_OK;
  hipLaunchKernelGGL(k_neigh_samples, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, as_stream(stream), vert, cols,
                     idx, rnd, n, delta, delta_eps, use_neigh_delta, dithering, pts, normals, sdfs);
  MP_LAUNCHED("k_neigh_samples");
  return MP_OK;
}

}  // extern "C"
