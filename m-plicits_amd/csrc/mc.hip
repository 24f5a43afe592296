// Marching cubes on the GPU with integer-exact, deterministic vertex and face
// indexing.  Replaces the CPU call the reference makes after sampling,
// skimage.measure.marching_cubes(volume, 0.0, spacing=[h]*3) (Lewiner method,
// gradient_direction='descent', allow_degenerate=True; i3d/meshing.py:296-298).
//
// Volume layout: [nz][ny][nx] float32, x fastest (the reference's
// sdf.reshape(N, N, N)).  Output vertices are in voxel units, (axis0, axis1,
// axis2) = (z, y, x) order, as scikit-image returns them before the spacing
// multiply; the caller applies spacing and origin.
//
// Vertex ownership: grid point (x, y, z) owns the intersected edges that leave
// it in +x, +y, +z (one vertex each) and the centre vertex of the cell whose
// lowest corner it is, when that cell's tiling uses one.  Passes:
//   1. classify: the active points (those that own vertices or triangles)
//      as bit masks, one read of the volume with float sign tests
//   2. scan the per-tile active counts
//   3. compact: the active points' indices in tile order
//   4. per active point (vertex count, triangle count); 5. scan them: vertex
//      and face ids are exclusive prefix sums over the active points
//   6. vertices -> vbase[point]; 7. faces: each cell's tiling, every entry
//      mapped to (owner point, axis) -> vertex id
// Corner classification, ambiguity tests and vertex placement follow the
// Lewiner algorithm with scikit-image's numerics: double arithmetic on
// v = volume - level, a corner is inside when v > 0, tests against 2^-52,
// an edge vertex = its two corners weighted by 1 / (2^-52 + |v|) (the linear
// interpolation point), the centre vertex = the eight corners weighted the
// same way, the interior test's undecided sub-cases 5 and 10 read as false.
// The double arithmetic is there to reproduce scikit-image's decisions on
// near-zero and near-tied corner values bit for bit (the triangle sets are
// compared with its output); it runs only on the cells the surface crosses.
#include "common.h"

#define MC_TABLE __device__ static const
#include "mc_tables.h"

namespace mp {

static constexpr double MC_EPS = 2.220446049250313e-16;
static constexpr int MC_BLOCK = 256;
// vertex counts below this keep a point's edge flags in the top bits of its
// vertex-base word
static constexpr uint32_t MC_PACK_LIMIT = 1u << 28;

// corner p -> (dx, dy, dz)
__device__ static const int8_t MC_CX[8] = {0, 1, 1, 0, 0, 1, 1, 0};
__device__ static const int8_t MC_CY[8] = {0, 0, 1, 1, 0, 0, 1, 1};
__device__ static const int8_t MC_CZ[8] = {0, 0, 0, 0, 1, 1, 1, 1};
// edge -> corners
__device__ static const int8_t MC_EA[12] = {0, 1, 2, 3, 4, 5, 6, 7, 0, 1, 2, 3};
__device__ static const int8_t MC_EB[12] = {1, 2, 3, 0, 5, 6, 7, 4, 4, 5, 6, 7};
// edge -> owner point offset (dx, dy, dz) and axis (0 x, 1 y, 2 z)
__device__ static const int8_t MC_OX[12] = {0, 1, 0, 0, 0, 1, 0, 0, 0, 1, 1, 0};
__device__ static const int8_t MC_OY[12] = {0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1, 1};
__device__ static const int8_t MC_OZ[12] = {0, 0, 0, 0, 1, 1, 1, 1, 0, 0, 0, 0};
__device__ static const int8_t MC_AX[12] = {0, 1, 0, 1, 0, 1, 0, 1, 2, 2, 2, 2};
// For the interior test through edge e: corners of the three parallel edges,
// in cyclic order around it (B0 B1 C0 C1 D0 D1).
__device__ static const int8_t MC_PAR[12][6] = {
    {3, 2, 7, 6, 4, 5}, {0, 3, 4, 7, 5, 6}, {1, 0, 5, 4, 6, 7}, {2, 1, 6, 5, 7, 4},
    {7, 6, 3, 2, 0, 1}, {4, 7, 0, 3, 1, 2}, {5, 4, 1, 0, 2, 3}, {6, 5, 2, 1, 3, 0},
    {3, 7, 2, 6, 1, 5}, {0, 4, 3, 7, 2, 6}, {1, 5, 0, 4, 3, 7}, {2, 6, 1, 5, 0, 4}};
// face -> corners (A, B, C, D)
__device__ static const int8_t MC_FACE[6][4] = {{0, 4, 5, 1}, {1, 5, 6, 2}, {2, 6, 7, 3},
                                               {3, 7, 4, 0}, {0, 3, 2, 1}, {4, 7, 6, 5}};

struct McVol {
  const float* vol;
  int nx, ny, nz;
  double level;
};

__device__ inline double mc_value(const McVol& V, int x, int y, int z) {
  return (double)V.vol[((int64_t)z * V.ny + y) * V.nx + x] - V.level;
}
__device__ inline bool mc_pos(double v) { return v > 0; }
__device__ inline double mc_weight(double v) { return 1.0 / (MC_EPS + fabs(v)); }

__device__ inline bool mc_face(const double* v, int face) {
  const int f = (face < 0 ? -face : face) - 1;
  const double A = v[MC_FACE[f][0]], B = v[MC_FACE[f][1]], C = v[MC_FACE[f][2]], D = v[MC_FACE[f][3]];
  const double det = A * C - B * D;
  if (fabs(det) < MC_EPS) return face >= 0;
  return face * A * det >= 0;
}

__device__ inline bool mc_interior(const double* v, int cas, int cfg, int sub, int s) {
  double At = 0, Bt = 0, Ct = 0, Dt = 0;
  if (cas == 4 || cas == 10) {
    const double a = (v[4] - v[0]) * (v[6] - v[2]) - (v[7] - v[3]) * (v[5] - v[1]);
    const double b = v[2] * (v[4] - v[0]) + v[0] * (v[6] - v[2]) - v[1] * (v[7] - v[3]) -
                     v[3] * (v[5] - v[1]);
    const double t = -b / (2 * a);
    if (t < 0 || t > 1) return s > 0;
    At = v[0] + (v[4] - v[0]) * t;
    Bt = v[3] + (v[7] - v[3]) * t;
    Ct = v[2] + (v[6] - v[2]) * t;
    Dt = v[1] + (v[5] - v[1]) * t;
  } else {
    int edge;
    if (cas == 6) edge = MC_TEST6[cfg][2];
    else if (cas == 7) edge = MC_TEST7[cfg][4];
    else if (cas == 12) edge = MC_TEST12[cfg][3];
    else edge = MC_TILING13_5_1[cfg][sub][0];
    const double t = v[MC_EA[edge]] / (v[MC_EA[edge]] - v[MC_EB[edge]]);
    const int8_t* e = MC_PAR[edge];
    Bt = v[e[0]] + (v[e[1]] - v[e[0]]) * t;
    Ct = v[e[2]] + (v[e[3]] - v[e[2]]) * t;
    Dt = v[e[4]] + (v[e[5]] - v[e[4]]) * t;
  }
  const int test = (At >= 0) | ((Bt >= 0) << 1) | ((Ct >= 0) << 2) | ((Dt >= 0) << 3);
  switch (test) {
    case 5: return (At * Ct - Bt * Dt < MC_EPS) ? s > 0 : false;
    case 10: return (At * Ct - Bt * Dt >= MC_EPS) ? s > 0 : false;
    case 7: case 11: case 13: case 14: case 15: return s < 0;
    default: return s > 0;
  }
}

// Tiling of one cube: pointer to its table row and the triangle count.
__device__ inline int mc_tiling(const double* v, int idx, const int8_t** tri) {
  const int cas = MC_CASES[idx][0], cfg = MC_CASES[idx][1];
  const int8_t* t = nullptr;
  int n = 0;
  switch (cas) {
    case 0: break;
    case 1: t = MC_TILING1[cfg]; n = 1; break;
    case 2: t = MC_TILING2[cfg]; n = 2; break;
    case 3:
      if (mc_face(v, MC_TEST3[cfg])) { t = MC_TILING3_2[cfg]; n = 4; }
      else { t = MC_TILING3_1[cfg]; n = 2; }
      break;
    case 4:
      if (mc_interior(v, 4, cfg, 0, MC_TEST4[cfg])) { t = MC_TILING4_1[cfg]; n = 2; }
      else { t = MC_TILING4_2[cfg]; n = 6; }
      break;
    case 5: t = MC_TILING5[cfg]; n = 3; break;
    case 6:
      if (mc_face(v, MC_TEST6[cfg][0])) { t = MC_TILING6_2[cfg]; n = 5; }
      else if (mc_interior(v, 6, cfg, 0, MC_TEST6[cfg][1])) { t = MC_TILING6_1_1[cfg]; n = 3; }
      else { t = MC_TILING6_1_2[cfg]; n = 9; }
      break;
    case 7: {
      const int sub = mc_face(v, MC_TEST7[cfg][0]) | (mc_face(v, MC_TEST7[cfg][1]) << 1) |
                      (mc_face(v, MC_TEST7[cfg][2]) << 2);
      switch (sub) {
        case 0: t = MC_TILING7_1[cfg]; n = 3; break;
        case 1: t = MC_TILING7_2[cfg][0]; n = 5; break;
        case 2: t = MC_TILING7_2[cfg][1]; n = 5; break;
        case 3: t = MC_TILING7_3[cfg][0]; n = 9; break;
        case 4: t = MC_TILING7_2[cfg][2]; n = 5; break;
        case 5: t = MC_TILING7_3[cfg][1]; n = 9; break;
        case 6: t = MC_TILING7_3[cfg][2]; n = 9; break;
        default:
          if (mc_interior(v, 7, cfg, 0, MC_TEST7[cfg][3])) { t = MC_TILING7_4_2[cfg]; n = 9; }
          else { t = MC_TILING7_4_1[cfg]; n = 5; }
      }
      break;
    }
    case 8: t = MC_TILING8[cfg]; n = 2; break;
    case 9: t = MC_TILING9[cfg]; n = 4; break;
    case 10:
      if (mc_face(v, MC_TEST10[cfg][0])) {
        if (mc_face(v, MC_TEST10[cfg][1])) { t = MC_TILING10_1_1_[cfg]; n = 4; }
        else { t = MC_TILING10_2[cfg]; n = 8; }
      } else if (mc_face(v, MC_TEST10[cfg][1])) { t = MC_TILING10_2_[cfg]; n = 8; }
      else if (mc_interior(v, 10, cfg, 0, MC_TEST10[cfg][2])) { t = MC_TILING10_1_1[cfg]; n = 4; }
      else { t = MC_TILING10_1_2[cfg]; n = 8; }
      break;
    case 11: t = MC_TILING11[cfg]; n = 4; break;
    case 12:
      if (mc_face(v, MC_TEST12[cfg][0])) {
        if (mc_face(v, MC_TEST12[cfg][1])) { t = MC_TILING12_1_1_[cfg]; n = 4; }
        else { t = MC_TILING12_2[cfg]; n = 8; }
      } else if (mc_face(v, MC_TEST12[cfg][1])) { t = MC_TILING12_2_[cfg]; n = 8; }
      else if (mc_interior(v, 12, cfg, 0, MC_TEST12[cfg][2])) { t = MC_TILING12_1_1[cfg]; n = 4; }
      else { t = MC_TILING12_1_2[cfg]; n = 8; }
      break;
    case 13: {
      int sub = 0;
      for (int k = 0; k < 6; ++k) sub |= mc_face(v, MC_TEST13[cfg][k]) << k;
      const int sc = MC_SUBCONFIG13[sub];
      if (sc == 0) { t = MC_TILING13_1[cfg]; n = 4; }
      else if (sc <= 6) { t = MC_TILING13_2[cfg][sc - 1]; n = 6; }
      else if (sc <= 18) { t = MC_TILING13_3[cfg][sc - 7]; n = 10; }
      else if (sc <= 22) { t = MC_TILING13_4[cfg][sc - 19]; n = 12; }
      else if (sc <= 26) {
        if (mc_interior(v, 13, cfg, sc - 23, MC_TEST13[cfg][6])) { t = MC_TILING13_5_1[cfg][sc - 23]; n = 6; }
        else { t = MC_TILING13_5_2[cfg][sc - 23]; n = 10; }
      } else if (sc <= 38) { t = MC_TILING13_3_[cfg][sc - 27]; n = 10; }
      else if (sc <= 44) { t = MC_TILING13_2_[cfg][sc - 39]; n = 6; }
      else if (sc == 45) { t = MC_TILING13_1_[cfg]; n = 4; }
      break;
    }
    case 14: t = MC_TILING14[cfg]; n = 4; break;
  }
  *tri = t;
  return n;
}

__device__ inline bool mc_uses_centre(const int8_t* t, int n) {
  for (int k = 0; k < 3 * n; ++k)
    if (t[k] == 12) return true;
  return false;
}

// Per point: owned-edge flags (bit 0 x, 1 y, 2 z) and the counts.
struct McPoint {
  int flags, nv, nt;
};

__device__ inline int mc_edge_flags(const McVol& V, int x, int y, int z) {
  const bool p0 = mc_pos(mc_value(V, x, y, z));
  int f = 0;
  if (x + 1 < V.nx && mc_pos(mc_value(V, x + 1, y, z)) != p0) f |= 1;
  if (y + 1 < V.ny && mc_pos(mc_value(V, x, y + 1, z)) != p0) f |= 2;
  if (z + 1 < V.nz && mc_pos(mc_value(V, x, y, z + 1)) != p0) f |= 4;
  return f;
}

// Counts of point (x, y, z): the cell corners v[8] = volume - level (raw,
// loaded where they exist) and the case index into *idx (0 without a cell).
__device__ inline McPoint mc_point(const McVol& V, int x, int y, int z, double* v, int* idx) {
  McPoint r{0, 0, 0};
  *idx = 0;
  if (x + 1 < V.nx && y + 1 < V.ny && z + 1 < V.nz) {
    int id = 0;
#pragma unroll
    for (int p = 0; p < 8; ++p) {
      v[p] = mc_value(V, x + MC_CX[p], y + MC_CY[p], z + MC_CZ[p]);
      id |= mc_pos(v[p]) << p;
    }
    // the owned edges' far ends are cell corners 1 (x), 3 (y) and 4 (z)
    const int p0 = id & 1;
    r.flags = (((id >> 1) & 1) != p0) | ((((id >> 3) & 1) != p0) << 1) | ((((id >> 4) & 1) != p0) << 2);
    r.nv = __popc(r.flags);
    if (id != 0 && id != 255) {
      const int8_t* t;
      r.nt = mc_tiling(v, id, &t);
      if (mc_uses_centre(t, r.nt)) r.nv += 1;
    }
    *idx = id;
  } else {
    // border point: owned edges only
    r.flags = mc_edge_flags(V, x, y, z);
    r.nv = __popc(r.flags);
  }
  return r;
}

// ---------------------------------------------------------------------------
// Pass 1, classification: a point is active when its cell (lowest corner at
// the point) has corners on both sides of the level, or, on the far faces of
// the grid where it has no cell, when one of its owned edges is crossed.
// Exactly the active points own vertices or triangles.  One wave classifies a
// tile of 64 columns (x) x MT_Y rows (y) x MT_Z planes (z): every lane reads
// its column at x and x + 1 for the MT_Y + 1 rows of each plane (one plane
// ahead of its use, so that a plane's loads are in flight while the one
// before it is classified), keeps the corner signs as bits and writes one
// 64-bit active mask per row and plane.  Only float comparisons: v > level
// is the sign test of the double arithmetic below (v - level > 0) exactly.
// ---------------------------------------------------------------------------
static constexpr int MT_Y = 8, MT_Z = 16;

struct McTiles {
  int nsx, nsy, nsz;
  int64_t n;
};
__host__ __device__ inline McTiles mc_tiles(int nx, int ny, int nz) {
  McTiles t;
  t.nsx = (nx + 63) / 64;
  t.nsy = (ny + MT_Y - 1) / MT_Y;
  t.nsz = (nz + MT_Z - 1) / MT_Z;
  t.n = (int64_t)t.nsx * t.nsy * t.nsz;
  return t;
}

// One plane of a tile for this lane: a[r] = v(x, y0 + r) and h[r] = v(x0 + 64,
// y0 + r) (the column after the tile's last, the same address in every lane),
// r = 0 .. MT_Y; rows and columns past the grid read clamped addresses, their
// signs are never used.
struct McPlane {
  float a[MT_Y + 1], h[MT_Y + 1];
};
// roff[r]: offset of row y0 + r (clamped to the grid) inside a plane
__device__ inline void mc_load_plane(const McVol& V, const int (&roff)[MT_Y + 1], int cx, int hx, int z, McPlane& p) {
  const float* pl = V.vol + (int64_t)z * V.ny * V.nx;
#pragma unroll
  for (int r = 0; r <= MT_Y; ++r) {
    p.a[r] = pl[roff[r] + cx];
    p.h[r] = pl[roff[r] + hx];
  }
}

// Sign bits of a plane for this lane: bit r of .x = v(x, y0 + r) > level,
// bit r of .y = the same at x + 1 (from the next lane; the last lane takes
// the column after the tile, h).
__device__ inline uint2 mc_plane_bits(const McPlane& p, float level, int lane) {
  uint32_t a = 0, h = 0;
#pragma unroll
  for (int r = 0; r <= MT_Y; ++r) {
    a |= (p.a[r] > level ? 1u : 0u) << r;
    h |= (p.h[r] > level ? 1u : 0u) << r;
  }
  const uint32_t n = (uint32_t)__shfl_down((int)a, 1);
  return make_uint2(a, lane == 63 ? h : n);
}

// wave-uniform tile origin
__device__ inline void mc_tile_origin(const McTiles& T, int64_t t, int& sx, int& y0, int& z0) {
  sx = (int)(t % T.nsx);
  const int64_t r = t / T.nsx;
  y0 = (int)(r % T.nsy) * MT_Y;
  z0 = (int)(r / T.nsy) * MT_Z;
}

// Workgroups are dealt to the eight XCDs in turn (workgroup w to XCD w % 8).
// The logical block of workgroup w gives every XCD one contiguous range of
// blocks, so that the halo rows and columns a tile shares with its neighbours
// in x and y are fetched into that XCD's L2 once.
__device__ inline int64_t mc_xcd_block(unsigned w, unsigned nw) {
  const unsigned k = w % 8u, per = nw / 8u, rem = nw % 8u;
  return (int64_t)k * per + (k < rem ? k : rem) + w / 8u;
}

// mask[(z * ny + y) * nsx + sx] = active bits of the 64 points of row (y, z)
// in x-segment sx; tile_tot[2 t] = active points of tile t (pair for the scan).
__global__ __launch_bounds__(MC_BLOCK) void k_mc_classify(McVol V, McTiles T, unsigned long long* __restrict__ mask,
                                                          uint32_t* __restrict__ tile_tot) {
  // the wave index through readfirstlane: the compiler then knows t (and
  // every tile quantity) to be wave-uniform and keeps them in scalar registers
  const int64_t t = mc_xcd_block(blockIdx.x, gridDim.x) * (MC_BLOCK / 64) +
                    __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (t >= T.n) return;  // wave-uniform
  const int lane = threadIdx.x & 63;
  int sx, y0, z0;
  mc_tile_origin(T, t, sx, y0, z0);
  const int x = 64 * sx + lane;
  const int cx = x < V.nx ? x : V.nx - 1, hx = 64 * sx + 64 < V.nx ? 64 * sx + 64 : V.nx - 1;
  const bool xe = x < V.nx, xv = x + 1 < V.nx;
  const float level = (float)V.level;
  const int zend = z0 + MT_Z < V.nz ? z0 + MT_Z : V.nz;  // points: planes [z0, zend)
  const int nrows = y0 + MT_Y < V.ny ? MT_Y : V.ny - y0;
  int roff[MT_Y + 1];
#pragma unroll
  for (int r = 0; r <= MT_Y; ++r) roff[r] = (y0 + r < V.ny ? y0 + r : V.ny - 1) * V.nx;
  // per-row bits (bit r = row y0 + r): rows of the grid, rows with a +y neighbour
  const uint32_t ROWS = (1u << nrows) - 1u;
  const uint32_t YV = y0 + MT_Y < V.ny ? ROWS : ROWS >> 1;
  // two planes in flight ahead of the one classified: pa holds the odd steps'
  // upper plane, pb the even steps'
  McPlane pa, pb;
  mc_load_plane(V, roff, cx, hx, z0, pa);
  uint2 lo = mc_plane_bits(pa, level, lane);
  if (z0 + 1 < V.nz) mc_load_plane(V, roff, cx, hx, z0 + 1, pa);
  if (z0 + 2 < V.nz && z0 + 2 <= zend) mc_load_plane(V, roff, cx, hx, z0 + 2, pb);
  uint32_t count = 0;
  // classify plane z with its upper plane in `pl`, then refill `pl` with plane z + 3
  auto step = [&](int z, McPlane& pl) {
    const bool zv = z + 1 < V.nz;
    const uint2 up = zv ? mc_plane_bits(pl, level, lane) : make_uint2(0u, 0u);
    if (z + 3 < V.nz && z + 3 <= zend) mc_load_plane(V, roff, cx, hx, z + 3, pl);
    // bit r: the cell of row r has corners on both sides (its eight corners
    // are rows r and r + 1 of x and x + 1 in both planes)
    const uint32_t L = lo.x, L1 = lo.y, U = up.x, U1 = up.y;
    const uint32_t any = L | L1 | (L >> 1) | (L1 >> 1) | U | U1 | (U >> 1) | (U1 >> 1);
    const uint32_t all = L & L1 & (L >> 1) & (L1 >> 1) & U & U1 & (U >> 1) & (U1 >> 1);
    // owned-edge crossings (x, y and z edges of row r's point)
    const uint32_t ZV = zv ? ROWS : 0u;
    const uint32_t ex = xv ? (L ^ L1) : 0u, ey = (L ^ (L >> 1)) & YV, ez = (L ^ U) & ZV;
    const uint32_t inner = xv ? (YV & ZV) : 0u;  // rows whose point has a cell
    const uint32_t act = xe ? (((any & ~all) & inner) | ((ex | ey | ez) & ~inner)) & ROWS : 0u;
    unsigned long long mine = 0;  // lane r: the mask of row r
#pragma unroll
    for (int r = 0; r < MT_Y; ++r) {
      const unsigned long long m = __ballot((act >> r) & 1u);
      count += (uint32_t)__popcll(m);
      if (lane == r) mine = m;
    }
    if (lane < nrows) mask[((int64_t)z * V.ny + y0 + lane) * T.nsx + sx] = mine;
    lo = up;
  };
  for (int z = z0; z < zend; z += 2) {
    step(z, pa);
    if (z + 1 < zend) step(z + 1, pb);
  }
  if (lane == 0) {
    tile_tot[2 * t] = count;
    tile_tot[2 * t + 1] = 0u;
  }
}

// Block-wide exclusive scan of two counters (256 threads = 4 waves).
__device__ inline void block_scan2(int a, int b, int* ea, int* eb, int* ta, int* tb) {
  __shared__ int sa[MC_BLOCK / 64], sb[MC_BLOCK / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int ia = a, ib = b;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int xa = __shfl_up(ia, o), xb = __shfl_up(ib, o);
    if (lane >= o) { ia += xa; ib += xb; }
  }
  if (lane == 63) { sa[w] = ia; sb[w] = ib; }
  __syncthreads();
  int oa = 0, ob = 0, sumA = 0, sumB = 0;
#pragma unroll
  for (int k = 0; k < MC_BLOCK / 64; ++k) {
    if (k < w) { oa += sa[k]; ob += sb[k]; }
    sumA += sa[k];
    sumB += sb[k];
  }
  *ea = oa + ia - a;
  *eb = ob + ib - b;
  *ta = sumA;
  *tb = sumB;
  __syncthreads();
}

// Exclusive scan of the block sums (pairs: vertices, triangles) in place, in
// three passes: segment sums (MC_SEG blocks per workgroup), one workgroup
// scanning the segment sums (totals into tot[0..1]), then every segment
// scanned from its offset.
static constexpr int MC_SEG = 8 * MC_BLOCK;

__global__ __launch_bounds__(MC_BLOCK) void k_mc_scan_reduce(const uint32_t* __restrict__ block_tot, int64_t nblocks,
                                                             unsigned long long* __restrict__ seg) {
  const int64_t s0 = (int64_t)blockIdx.x * MC_SEG;
  unsigned long long a = 0, b = 0;
  for (int k = threadIdx.x; k < MC_SEG; k += MC_BLOCK) {
    const int64_t j = s0 + k;
    if (j < nblocks) {
      a += block_tot[2 * j];
      b += block_tot[2 * j + 1];
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_xor(a, o);
    b += __shfl_xor(b, o);
  }
  __shared__ unsigned long long sa[MC_BLOCK / 64], sb[MC_BLOCK / 64];
  if ((threadIdx.x & 63) == 0) {
    sa[threadIdx.x >> 6] = a;
    sb[threadIdx.x >> 6] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long ta = 0, tb = 0;
    for (int w = 0; w < MC_BLOCK / 64; ++w) {
      ta += sa[w];
      tb += sb[w];
    }
    seg[2 * blockIdx.x] = ta;
    seg[2 * blockIdx.x + 1] = tb;
  }
}

__global__ __launch_bounds__(1024) void k_mc_scan_top(unsigned long long* seg, int64_t nseg,
                                                      unsigned long long* tot) {
  __shared__ unsigned long long s0[1024], s1[1024];
  const int t = threadIdx.x;
  const int64_t per = (nseg + 1023) / 1024;
  const int64_t lo = t * per, hi = (lo + per < nseg) ? lo + per : nseg;
  unsigned long long a = 0, b = 0;
  for (int64_t k = lo; k < hi; ++k) { a += seg[2 * k]; b += seg[2 * k + 1]; }
  s0[t] = a;
  s1[t] = b;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    unsigned long long xa = 0, xb = 0;
    if (t >= o) { xa = s0[t - o]; xb = s1[t - o]; }
    __syncthreads();
    s0[t] += xa;
    s1[t] += xb;
    __syncthreads();
  }
  unsigned long long oa = s0[t] - a, ob = s1[t] - b;
  for (int64_t k = lo; k < hi; ++k) {
    const unsigned long long ca = seg[2 * k], cb = seg[2 * k + 1];
    seg[2 * k] = oa;
    seg[2 * k + 1] = ob;
    oa += ca;
    ob += cb;
  }
  if (t == 1023) { tot[0] = s0[1023]; tot[1] = s1[1023]; }
}

// Segment blockIdx.x: each thread scans 8 consecutive block sums serially,
// the workgroup scans the thread totals, then every entry gets its offset.
__global__ __launch_bounds__(MC_BLOCK) void k_mc_scan_apply(uint32_t* __restrict__ block_tot, int64_t nblocks,
                                                            const unsigned long long* __restrict__ seg) {
  constexpr int PER = MC_SEG / MC_BLOCK;
  const int64_t j0 = (int64_t)blockIdx.x * MC_SEG + (int64_t)threadIdx.x * PER;
  uint32_t ca[PER], cb[PER];
  uint32_t a = 0, b = 0;
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int64_t j = j0 + k;
    ca[k] = j < nblocks ? block_tot[2 * j] : 0u;
    cb[k] = j < nblocks ? block_tot[2 * j + 1] : 0u;
    a += ca[k];
    b += cb[k];
  }
  int ea, eb, ta, tb;
  block_scan2((int)a, (int)b, &ea, &eb, &ta, &tb);
  unsigned long long oa = seg[2 * blockIdx.x] + (unsigned)ea, ob = seg[2 * blockIdx.x + 1] + (unsigned)eb;
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int64_t j = j0 + k;
    if (j < nblocks) {
      block_tot[2 * j] = (uint32_t)oa;
      block_tot[2 * j + 1] = (uint32_t)ob;
    }
    oa += ca[k];
    ob += cb[k];
  }
}

// Pass 3, compaction: the active points of every tile, in tile order and
// inside a tile by plane, row and column, from the tile's exclusive offset.
// The tile's MT_Z * MT_Y = 128 row masks are loaded at once (lane l: rows l
// and l + 64 of the tile, in plane-major order) and offset by a wave scan of
// their counts; only the non-empty rows are then written out.
static_assert(MT_Y * MT_Z == 128, "two row masks per lane");
__global__ __launch_bounds__(MC_BLOCK) void k_mc_compact(McVol V, McTiles T, const unsigned long long* __restrict__ mask,
                                                         const uint32_t* __restrict__ tile_off,
                                                         uint32_t* __restrict__ list) {
  // the wave index through readfirstlane: the compiler then knows t (and
  // every tile quantity) to be wave-uniform and keeps them in scalar registers
  const int64_t t = (int64_t)blockIdx.x * (MC_BLOCK / 64) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (t >= T.n) return;  // wave-uniform
  const int lane = threadIdx.x & 63;
  int sx, y0, z0;
  mc_tile_origin(T, t, sx, y0, z0);
  const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  unsigned long long m[2];
  int c[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int w = lane + 64 * h;
    const int z = z0 + w / MT_Y, y = y0 + w % MT_Y;
    m[h] = (z < V.nz && y < V.ny) ? mask[((int64_t)z * V.ny + y) * T.nsx + sx] : 0ull;
    c[h] = __popcll(m[h]);
  }
  // inclusive wave scans of the two halves' counts
  int s0 = c[0], s1 = c[1];
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int a = __shfl_up(s0, o), b = __shfl_up(s1, o);
    if (lane >= o) {
      s0 += a;
      s1 += b;
    }
  }
  const int half0 = __shfl(s0, 63);
  const uint32_t base = tile_off[2 * t];
  const uint32_t off[2] = {base + (uint32_t)(s0 - c[0]), base + (uint32_t)(half0 + s1 - c[1])};
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    unsigned long long rows = __ballot(m[h] != 0ull);
    while (rows) {
      const int j = __ffsll((long long)rows) - 1;  // wave-uniform
      rows &= rows - 1;
      const unsigned long long mj = __shfl(m[h], j);
      const uint32_t oj = __shfl(off[h], j);
      if ((mj >> lane) & 1ull) {
        const int w = j + 64 * h;
        const int z = z0 + w / MT_Y, y = y0 + w % MT_Y;
        list[oj + (uint32_t)__popcll(mj & below)] = (uint32_t)(((int64_t)z * V.ny + y) * V.nx + 64 * sx + lane);
      }
    }
  }
}

__device__ inline void mc_xyz_of(const McVol& V, uint32_t i, int& x, int& y, int& z) {
  const uint32_t row = i / (uint32_t)V.nx;
  x = (int)(i - row * (uint32_t)V.nx);
  y = (int)(row % (uint32_t)V.ny);
  z = (int)(row / (uint32_t)V.ny);
}

// Pass 4: (vertex count, triangle count) of every active point into cnt[2 k + 0..1].
__global__ __launch_bounds__(MC_BLOCK) void k_mc_entry_counts(McVol V, const uint32_t* __restrict__ list, int64_t na,
                                                              uint32_t* __restrict__ cnt) {
  const int64_t k = (int64_t)blockIdx.x * MC_BLOCK + threadIdx.x;
  if (k >= na) return;
  int x, y, z;
  mc_xyz_of(V, list[k], x, y, z);
  double v[8];
  int idx;
  const McPoint p = mc_point(V, x, y, z, v, &idx);
  cnt[2 * k] = (uint32_t)p.nv;
  cnt[2 * k + 1] = (uint32_t)p.nt;
}

// Pass 6: vertices of every active point (vertex ids from the scanned counts;
// the active count from the device, a grid-stride loop).
__global__ __launch_bounds__(MC_BLOCK) void k_mc_vertices(McVol V, const uint32_t* __restrict__ list,
                                                          const unsigned long long* __restrict__ na_dev,
                                                          const uint32_t* __restrict__ off, uint32_t* vbase,
                                                          float* verts, float* normals, float* values) {
  const int64_t na = (int64_t)*na_dev;
  const bool pack = na_dev[2] < MC_PACK_LIMIT;  // vertex count
  for (int64_t k = (int64_t)blockIdx.x * MC_BLOCK + threadIdx.x; k < na; k += (int64_t)gridDim.x * MC_BLOCK) {
    const uint32_t i = list[k];
    int x, y, z;
    mc_xyz_of(V, i, x, y, z);
    double raw[8];
    int idx;
    const McPoint p = mc_point(V, x, y, z, raw, &idx);
    if (p.nv == 0) continue;
    // only points that own vertices are ever looked up (an edge a triangle
    // uses has its owner's flag set; a centre vertex belongs to this point)
    const uint32_t base = off[2 * k];
    vbase[i] = pack ? base | ((uint32_t)p.flags << 28) : base;
    const double v0 = mc_value(V, x, y, z);
    uint32_t id = base;
    for (int ax = 0; ax < 3; ++ax) {
      if (!(p.flags & (1 << ax))) continue;
      const int x1 = x + (ax == 0), y1 = y + (ax == 1), z1 = z + (ax == 2);
      const double v1 = mc_value(V, x1, y1, z1);
      const double w0 = mc_weight(v0), w1 = mc_weight(v1);
      const double t = w1 / (w0 + w1);
      double pos[3] = {(double)x, (double)y, (double)z};
      pos[ax] += t;
      verts[3 * (int64_t)id + 0] = (float)pos[2];
      verts[3 * (int64_t)id + 1] = (float)pos[1];
      verts[3 * (int64_t)id + 2] = (float)pos[0];
      if (normals) {
        // -grad of the volume by central differences (one-sided at the border),
        // interpolated along the edge, unit length
        double g[3];
        for (int kk = 0; kk < 3; ++kk) {
          const int dx = kk == 0, dy = kk == 1, dz = kk == 2;
          auto cd = [&](int px, int py, int pz) {
            const int ax0 = px - dx >= 0 ? px - dx : px, ay0 = py - dy >= 0 ? py - dy : py,
                      az0 = pz - dz >= 0 ? pz - dz : pz;
            const int ax1 = px + dx < V.nx ? px + dx : px, ay1 = py + dy < V.ny ? py + dy : py,
                      az1 = pz + dz < V.nz ? pz + dz : pz;
            const double span = (double)((ax1 - ax0) + (ay1 - ay0) + (az1 - az0));
            return (mc_value(V, ax1, ay1, az1) - mc_value(V, ax0, ay0, az0)) / (span > 0 ? span : 1.0);
          };
          g[kk] = (1 - t) * cd(x, y, z) + t * cd(x1, y1, z1);
        }
        const double len = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
        const double sc = len > 0 ? -1.0 / len : 0.0;
        normals[3 * (int64_t)id + 0] = (float)(g[2] * sc);
        normals[3 * (int64_t)id + 1] = (float)(g[1] * sc);
        normals[3 * (int64_t)id + 2] = (float)(g[0] * sc);
      }
      if (values) values[id] = (float)fabs(v1 - v0);
      ++id;
    }
    if (id < base + (uint32_t)p.nv) {
      // centre vertex of the cell at this point
      double sp[3] = {0, 0, 0}, ws = 0, vmax = -1e300, vmin = 1e300;
      for (int q = 0; q < 8; ++q) {
        const int dx = q & 1, dy = (q >> 1) & 1, dz = (q >> 2) & 1;
        const int lp = dz * 4 + (dy ? (dx ? 2 : 3) : (dx ? 1 : 0));
        const double w = mc_weight(raw[lp]);
        sp[0] += w * dx;
        sp[1] += w * dy;
        sp[2] += w * dz;
        ws += w;
        vmax = raw[lp] > vmax ? raw[lp] : vmax;
        vmin = raw[lp] < vmin ? raw[lp] : vmin;
      }
      verts[3 * (int64_t)id + 0] = (float)(z + sp[2] / ws);
      verts[3 * (int64_t)id + 1] = (float)(y + sp[1] / ws);
      verts[3 * (int64_t)id + 2] = (float)(x + sp[0] / ws);
      if (normals) {
        // gradient of the trilinear interpolant at the cell centre
        const double gx = (raw[1] + raw[2] + raw[5] + raw[6] - raw[0] - raw[3] - raw[4] - raw[7]) * 0.25;
        const double gy = (raw[2] + raw[3] + raw[6] + raw[7] - raw[0] - raw[1] - raw[4] - raw[5]) * 0.25;
        const double gz = (raw[4] + raw[5] + raw[6] + raw[7] - raw[0] - raw[1] - raw[2] - raw[3]) * 0.25;
        const double len = sqrt(gx * gx + gy * gy + gz * gz);
        const double sc = len > 0 ? -1.0 / len : 0.0;
        normals[3 * (int64_t)id + 0] = (float)(gz * sc);
        normals[3 * (int64_t)id + 1] = (float)(gy * sc);
        normals[3 * (int64_t)id + 2] = (float)(gx * sc);
      }
      if (values) values[id] = (float)(vmax - vmin);
    }
  }
}

// Vertex id of entry e (edge 0..11, or 12: the cell's centre vertex, after
// the point's own edge vertices).  vbase[owner] holds the owner's first vertex
// id and, when the vertex count leaves room (pack), its edge flags in bits
// 28..30: then no volume value of the owner is read.
__device__ inline uint32_t mc_vertex_id(const McVol& V, const uint32_t* vbase, bool pack, int x, int y, int z, int e) {
  const int ox = e == 12 ? x : x + MC_OX[e], oy = e == 12 ? y : y + MC_OY[e], oz = e == 12 ? z : z + MC_OZ[e];
  const int ax = e == 12 ? 3 : MC_AX[e];
  const uint32_t w = vbase[((int64_t)oz * V.ny + oy) * V.nx + ox];
  const uint32_t base = pack ? (w & (MC_PACK_LIMIT - 1u)) : w;
  const int f = pack ? (int)(w >> 28) : mc_edge_flags(V, ox, oy, oz);
  return base + __popc(f & ((1 << ax) - 1));
}

// Pass 7: faces of every active point, vertex ids looked up in vbase.
__global__ __launch_bounds__(MC_BLOCK) void k_mc_faces(McVol V, const uint32_t* __restrict__ list,
                                                       const unsigned long long* __restrict__ na_dev,
                                                       const uint32_t* __restrict__ off, const uint32_t* vbase,
                                                       int32_t* faces) {
  const int64_t na = (int64_t)*na_dev;
  const bool pack = na_dev[2] < MC_PACK_LIMIT;  // vertex count
  for (int64_t k = (int64_t)blockIdx.x * MC_BLOCK + threadIdx.x; k < na; k += (int64_t)gridDim.x * MC_BLOCK) {
    int x, y, z;
    mc_xyz_of(V, list[k], x, y, z);
    if (x + 1 >= V.nx || y + 1 >= V.ny || z + 1 >= V.nz) continue;
    double v[8];
    int idx = 0;
#pragma unroll
    for (int p = 0; p < 8; ++p) {
      v[p] = mc_value(V, x + MC_CX[p], y + MC_CY[p], z + MC_CZ[p]);
      idx |= mc_pos(v[p]) << p;
    }
    if (idx == 0 || idx == 255) continue;
    const int8_t* tri;
    const int nt = mc_tiling(v, idx, &tri);
    const int64_t fb = off[2 * k + 1];
    for (int kk = 0; kk < nt; ++kk) {
      // 'descent': faces in reversed table order
      for (int j = 0; j < 3; ++j)
        faces[3 * (fb + kk) + j] = (int32_t)mc_vertex_id(V, vbase, pack, x, y, z, tri[3 * kk + 2 - j]);
    }
  }
}

static size_t mc_segs(size_t nb) { return (nb + MC_SEG - 1) / MC_SEG; }
static size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

// Workspace (n points, T tiles):
//   totals: [0] active points, [1] 0 | [2] vertices, [3] faces  (256 B)
//   tile pairs [T][2] u32 | tile segments [segs(T)][2] u64 | masks [nz][ny][nsx] u64 |
//   active list [n] u32 | entry pairs [n][2] u32 | entry segments [segs(n)][2] u64 | vbase [n] u32
struct McLayout {
  size_t tiles, tseg, mask, list, cnt, eseg, vbase, total;
};
static McLayout mc_layout(int nz, int ny, int nx) {
  const McTiles T = mc_tiles(nx, ny, nz);
  const size_t n = (size_t)nz * ny * nx;
  McLayout L;
  L.tiles = 256;
  L.tseg = L.tiles + al256((size_t)T.n * 8);
  L.mask = L.tseg + al256(mc_segs((size_t)T.n) * 16);
  L.list = L.mask + al256((size_t)nz * ny * T.nsx * 8);
  L.cnt = L.list + al256(n * 4);
  L.eseg = L.cnt + al256(n * 8);
  L.vbase = L.eseg + al256(mc_segs(n) * 16);
  L.total = L.vbase + al256(n * 4);
  return L;
}

// Exclusive scan of `nb` pairs in place, the totals into tot[0..1].
static int mc_scan_pairs(uint32_t* pairs, int64_t nb, unsigned long long* seg, unsigned long long* tot, hipStream_t s) {
  const size_t ns = mc_segs((size_t)nb);
  hipLaunchKernelGGL(k_mc_scan_reduce, dim3((unsigned)ns), dim3(MC_BLOCK), 0, s, pairs, nb, seg);
  MP_LAUNCHED("k_mc_scan_reduce");
  hipLaunchKernelGGL(k_mc_scan_top, dim3(1), dim3(1024), 0, s, seg, (int64_t)ns, tot);
  MP_LAUNCHED("k_mc_scan_top");
  hipLaunchKernelGGL(k_mc_scan_apply, dim3((unsigned)ns), dim3(MC_BLOCK), 0, s, pairs, nb, seg);
  MP_LAUNCHED("k_mc_scan_apply");
  return MP_OK;
}

static int mc_check(const float* vol, int nz, int ny, int nx, void* ws, const char* fn) {
  MP_REQUIRE(vol != nullptr && ws != nullptr, MP_EINVAL, "%s: vol and ws must be non-NULL", fn);
  MP_REQUIRE(nz >= 2 && ny >= 2 && nx >= 2, MP_EINVAL,
             "%s: the volume must be at least 2x2x2, got %dx%dx%d", fn, nz, ny, nx);
  MP_REQUIRE((int64_t)nz * ny * nx <= (1ll << 31), MP_EUNSUPPORTED, "%s: volume too large", fn);
  return MP_OK;
}

}  // namespace mp

using namespace mp;

extern "C" {

size_t mp_mc_workspace_size(int nz, int ny, int nx) {
  if (nz < 2 || ny < 2 || nx < 2) return 0;
  return mc_layout(nz, ny, nx).total;
}

int mp_mc_count(const float* vol, int nz, int ny, int nx, float level, int64_t* n_verts,
                int64_t* n_faces, void* ws, void* stream) {
  int rc = mc_check(vol, nz, ny, nx, ws, "mp_mc_count");
  if (rc != MP_OK) return rc;
  hipStream_t s = as_stream(stream);
  McVol V{vol, nx, ny, nz, (double)level};
  const McTiles T = mc_tiles(nx, ny, nz);
  const McLayout L = mc_layout(nz, ny, nx);
  char* w = reinterpret_cast<char*>(ws);
  unsigned long long* tot = reinterpret_cast<unsigned long long*>(w);
  uint32_t* tiles = reinterpret_cast<uint32_t*>(w + L.tiles);
  unsigned long long* mask = reinterpret_cast<unsigned long long*>(w + L.mask);
  uint32_t* list = reinterpret_cast<uint32_t*>(w + L.list);
  uint32_t* cnt = reinterpret_cast<uint32_t*>(w + L.cnt);
  const unsigned wgs = (unsigned)((T.n + MC_BLOCK / 64 - 1) / (MC_BLOCK / 64));
  hipLaunchKernelGGL(k_mc_classify, dim3(wgs), dim3(MC_BLOCK), 0, s, V, T, mask, tiles);
  MP_LAUNCHED("k_mc_classify");
  rc = mc_scan_pairs(tiles, T.n, reinterpret_cast<unsigned long long*>(w + L.tseg), tot, s);
  if (rc != MP_OK) return rc;
  unsigned long long na = 0;
  MP_HIP(hipMemcpyAsync(&na, tot, sizeof(na), hipMemcpyDeviceToHost, s));
  MP_HIP(hipStreamSynchronize(s));
  unsigned long long h[2] = {0, 0};
  if (na > 0) {
    hipLaunchKernelGGL(k_mc_compact, dim3(wgs), dim3(MC_BLOCK), 0, s, V, T, mask, tiles, list);
    MP_LAUNCHED("k_mc_compact");
    hipLaunchKernelGGL(k_mc_entry_counts, dim3((unsigned)((na + MC_BLOCK - 1) / MC_BLOCK)), dim3(MC_BLOCK), 0, s, V,
                       list, (int64_t)na, cnt);
    MP_LAUNCHED("k_mc_entry_counts");
    rc = mc_scan_pairs(cnt, (int64_t)na, reinterpret_cast<unsigned long long*>(w + L.eseg), tot + 2, s);
    if (rc != MP_OK) return rc;
    MP_HIP(hipMemcpyAsync(h, tot + 2, sizeof(h), hipMemcpyDeviceToHost, s));
    MP_HIP(hipStreamSynchronize(s));
  }
  MP_REQUIRE(h[0] < (1ull << 31) && h[1] < (1ull << 31), MP_EUNSUPPORTED,
             "mp_mc_count: %llu vertices / %llu faces exceed int32 indexing", h[0], h[1]);
  if (n_verts) *n_verts = (int64_t)h[0];
  if (n_faces) *n_faces = (int64_t)h[1];
  return MP_OK;
}

int mp_mc_emit(const float* vol, int nz, int ny, int nx, float level, float* verts, int32_t* faces,
               float* normals, float* values, void* ws, void* stream) {
  int rc = mc_check(vol, nz, ny, nx, ws, "mp_mc_emit");
  if (rc != MP_OK) return rc;
  // verts / faces may be NULL only when mp_mc_count returned zero counts
  hipStream_t s = as_stream(stream);
  McVol V{vol, nx, ny, nz, (double)level};
  const McLayout L = mc_layout(nz, ny, nx);
  char* w = reinterpret_cast<char*>(ws);
  const unsigned long long* tot = reinterpret_cast<const unsigned long long*>(w);
  const uint32_t* list = reinterpret_cast<const uint32_t*>(w + L.list);
  const uint32_t* cnt = reinterpret_cast<const uint32_t*>(w + L.cnt);
  uint32_t* vbase = reinterpret_cast<uint32_t*>(w + L.vbase);
  // the active count stays on the device: a fixed grid of grid-stride loops
  int dev = 0, cus = 0;
  MP_HIP(hipGetDevice(&dev));
  MP_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const int64_t n = (int64_t)nz * ny * nx;
  const int64_t most = (n + MC_BLOCK - 1) / MC_BLOCK;
  const unsigned grid = (unsigned)(most < 8ll * cus ? most : 8ll * cus);
  hipLaunchKernelGGL(k_mc_vertices, dim3(grid), dim3(MC_BLOCK), 0, s, V, list, tot, cnt, vbase, verts, normals,
                     values);
  MP_LAUNCHED("k_mc_vertices");
  hipLaunchKernelGGL(k_mc_faces, dim3(grid), dim3(MC_BLOCK), 0, s, V, list, tot, cnt, vbase, faces);
  MP_LAUNCHED("k_mc_faces");
  return MP_OK;
}

}  // extern "C"
