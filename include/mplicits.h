/*
 * mplicits.h - C ABI of the MI355X (gfx950) M-plicits hot path.
 *
 * The reference project has no plugin/FFI layer: its boundary is the Python
 * call surface of i3d (SIREN.forward, create_mesh_multistage, the losses) and
 * the renderer's CUDA entry points.  This library sits under a Python mirror of
 * that surface (package `mplicits`) and is bound with ctypes.  Every symbol
 * below names the reference interface it replaces.
 *
 * Conventions
 *  - Plain C types only.  Streams are `hipStream_t` passed as `void*`
 *    (NULL = the null stream).  Every call is stream-ordered and asynchronous
 *    unless documented as a host-sync point.
 *  - Tensor arguments are caller-owned device pointers (torch tensors).
 *  - Return 0 on success, a negative MP_E* code otherwise; mp_last_error()
 *    returns the thread-local message of the last failure.  Nothing exits or
 *    throws across the ABI.
 *  - A stack handle must not be used from two host threads at once.
 */
#ifndef MPLICITS_H
#define MPLICITS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MP_ABI_VERSION 1

#define MP_OK 0
#define MP_EINVAL (-1)
#define MP_ENOMEM (-2)
#define MP_EHIP (-3)
#define MP_EUNSUPPORTED (-4)

/* arithmetic of the hidden-layer GEMMs (first layer, sines, output layer and
 * the residual sum are always FP32) */
#define MP_FP32 0
#define MP_BF16 1
#define MP_FP16 2
/* three-term BF16: every weight and activation as a BF16 pair hi + lo, each
 * product as hi*hi + hi*lo + lo*hi on the BF16 matrix cores with FP32
 * accumulation (about 16 significant bits per operand).  Levels wider than 256
 * run FP16 (their operand pairs would not fit a wave's registers).  Served by
 * evaluation (gradients in forward mode), extraction, tracing and rendering;
 * the batched extraction and the training entry points return
 * MP_EUNSUPPORTED. */
#define MP_BF16X3 3
/* three-term BF16 for the first of a level's two hidden products, and the
 * last one on the FP16 matrix cores with FP16 weights against the
 * activations as an FP16 pair (a*hi + a*lo: two MFMAs per step instead of
 * three).  Applies to levels with exactly three hidden layers up to width 256
 * (other levels as MP_BF16X3).  Served by evaluation (gradients in forward
 * mode), extraction and tracing (when every level up to width 256 has three
 * hidden layers); rendering, the batched extraction and the training entry
 * points return MP_EUNSUPPORTED.  The error of such a level stays near 5e-5
 * on the fitted coarse level where MP_BF16X3's is near 3e-6, and grows with
 * its output weights. */
#define MP_BF16X3F 4

/* sample coordinates of the extraction grid */
#define MP_GRID_REFERENCE 0 /* i3d/meshing.py:42-63 as torch evaluates it (float32 true division) */
#define MP_GRID_LATTICE 1   /* the regular lattice x = a*h - 1 */

#define MP_MAX_LEVELS 4

typedef struct mp_stack mp_stack;

/* One SIREN level: i3d/model.py:76-105 (n_in -> hidden[0] -> ... -> n_out),
 * weights in the renderer's .bin order (renderer/scripts/export_experiment.py:28-43):
 * `weights` = row-major W_k [out_k][in_k] concatenated for k = 0..n_hidden,
 * `biases` = b_k concatenated.  The frequency is baked (evaluated with w0 = 1,
 * i3d/model.py:143-180).  Both are device pointers that the stack borrows. */
typedef struct mp_level_desc {
  int n_in;             /* must be 3 */
  int n_out;            /* must be 1 */
  int n_hidden;         /* number of sine layers, len(hidden_layer_config) >= 1 */
  const int* hidden;    /* host array [n_hidden] of widths, each 1..512 */
  const float* weights; /* device */
  const float* biases;  /* device */
} mp_level_desc;

int mp_abi_version(void);
const char* mp_last_error(void);

/* Build a level stack (coarse level first).  Replaces the per-level
 * i3d.util.from_pth(...).to(device) (i3d/util.py:79-142) and the renderer's
 * NeuralImplicit::init (renderer/cuda/gemms.cu:111-134): packs every level into
 * MFMA fragment order in `precision`.  Allocates device memory. */
int mp_stack_create(const mp_level_desc* levels, int n_levels, int precision, void* stream,
                    mp_stack** out);
/* Re-pack from the borrowed weight pointers (after an optimizer step). */
int mp_stack_update(mp_stack* stack, void* stream);
void mp_stack_destroy(mp_stack* stack);
int mp_stack_levels(const mp_stack* stack);

/* Points per wavefront tile of the level kernels (a tile shape picked by
 * mp_autotune): 1 or 2 column blocks of the MFMA shape (2 = the default
 * block count of each width).
 * No counterpart in the reference (its CUTLASS tile shapes are compile-time,
 * renderer/cuda/layer_gemms.cu:1612-1856). */
int mp_stack_set_tiles(mp_stack* stack, int column_blocks);

/* Tile shapes per kernel family (no counterpart in the reference):
 *   MP_TILES_EXTRACT  the level kernels of mp_eval / mp_extract_volume, as
 *                     mp_stack_set_tiles (1: one column block per wave in
 *                     eight-wave workgroups, two waves per SIMD; 2: default);
 *   MP_TILES_BATCH    the batched extraction kernel (1: each width's default
 *                     column blocks per wave, 2: half of them);
 *   MP_TILES_TRACE    the tracers (0: follow MP_TILES_EXTRACT, 1 or 2 column
 *                     blocks per wave where the level widths allow two).
 * The tile shape changes the schedule, never a value. */
#define MP_TILES_EXTRACT 0
#define MP_TILES_BATCH 1
#define MP_TILES_TRACE 2
int mp_stack_set_kernel_tiles(mp_stack* stack, int family, int tiles);

/* Autotuner of those knobs: measures every tile shape of `family` for a
 * seeded level 3 -> width x n_hidden -> 1 in `precision` (HIP events after a
 * warm-up) and writes the fastest to *tiles:
 *   MP_TILES_EXTRACT  mp_extract_volume over a 160^3 grid (1 or 2);
 *   MP_TILES_BATCH    a one-job batched extraction over a 160^3 grid (1 or 2);
 *   MP_TILES_TRACE    mp_trace of a 512^2 frame, at most 32 steps (1 or 2),
 *                     per level evaluation.
 * cache_path (may be NULL) is a text file of "family width n_hidden precision
 * tiles" lines: a cached entry is returned without measuring, a new one is
 * appended.  Allocates device memory; host-sync point; serialised by a
 * process-wide mutex.  mp_autotune is the MP_TILES_EXTRACT family. */
int mp_autotune_kernel(const char* cache_path, int family, int width, int n_hidden, int precision, int* tiles);
int mp_autotune(const char* cache_path, int width, int n_hidden, int precision, int* column_blocks);

/* Evaluate sum_{l=lvl0}^{lvl1-1} f_l(x) at n points xyz [n][3] into sdf [n].
 * Replaces SIREN.forward (i3d/model.py:107-129) summed over levels as
 * create_mesh_multistage does with deltas=[] (i3d/meshing.py:232-233).
 * If grad != NULL, also writes the input gradient [n][3] of that sum
 * (i3d/diff_operators.py:259-265; renderer Normals::inference,
 * renderer/cuda/gemms.cu:263-337); both then come from one gradient pass per
 * level, whose forward rounds like the value-only kernels but may
 * accumulate in another order (values equal to within the precision's
 * rounding, not bit for bit). */
int mp_eval(const mp_stack* stack, const float* xyz, int64_t n, int lvl0, int lvl1, float* sdf,
            float* grad, void* stream);

/* Band-culled volume evaluation: the sampling loop of create_mesh_multistage
 * (i3d/meshing.py:203-238) with the same per-chunk culling semantics
 * (chunk = max_batch, reconstruct.py:128 uses 32768).  n_deltas must be 0
 * (all levels everywhere) or n_levels-1.  vol is N^3 floats, device.
 * level_counts (host, n_levels entries, may be NULL) receives the number of
 * samples each level was evaluated on.  Host-sync point (the band counts are
 * read back to size the launches).  Allocates device scratch. */
int mp_extract_volume(mp_stack* stack, int N, int grid_mode, int64_t chunk, const float* deltas,
                      int n_deltas, float* vol, int64_t* level_counts, void* stream);

/* mp_extract_volume without any allocation: the same kernels and results, with
 * the work lists in a fixed device array of the library
 * (MP_NOALLOC_SCRATCH_BYTES, held on every GPU the library runs on until the
 * process ends) instead of the stack's scratch.  Limits: N^3 <=
 * MP_NOALLOC_MAX_POINTS (N <= 256) and (levels - 1) * ceil(N^3 / chunk) <= 2^20;
 * larger calls return MP_EINVAL before any launch.  Still a host-sync point
 * (band counts are read back).  Every call on one GPU shares that GPU's array:
 * calls that overlap on one GPU (two streams, two threads) race.
 * No counterpart in the reference: for callers that cannot allocate (code
 * under graph capture, serving loops, harnesses that forbid hipMalloc). */
#define MP_NOALLOC_MAX_POINTS (1 << 24)
#define MP_NOALLOC_SCRATCH_BYTES (256 + (1 << 20) + 8 * (size_t)MP_NOALLOC_MAX_POINTS)
int mp_extract_volume_noalloc(mp_stack* stack, int N, int grid_mode, int64_t chunk, const float* deltas,
                              int n_deltas, float* vol, int64_t* level_counts, void* stream);

/* Batched band-culled extraction of many stacks (uneven widths, levels and
 * resolutions; every stack in the same precision).  The reference extracts
 * one model per create_mesh_multistage call (i3d/meshing.py:135-264); this
 * entry point runs the same per-model semantics for all models with one
 * persistent launch per stage.  mp_extract_batch_plan allocates the device
 * work lists and uploads the job tables (host-sync point); mp_extract_batch_run
 * only launches kernels (no allocation, host synchronisation or memset), so a
 * caller may capture it in a hipGraph and replay it; mp_extract_batch_counts
 * reads back the per-job, per-level sample counts of the last run (host-sync
 * point, [n_jobs][MP_MAX_LEVELS]).  The plan keeps pointers to the stacks'
 * packed weights and to the volumes: both must outlive it. */
typedef struct mp_batch mp_batch;
typedef struct mp_extract_job {
  mp_stack* stack;
  int N;                /* grid resolution */
  int grid_mode;        /* MP_GRID_* */
  int64_t chunk;        /* culling chunk (max_batch) */
  int n_deltas;         /* 0 (no culling) or stack levels - 1 */
  float deltas[MP_MAX_LEVELS - 1];
  float* vol;           /* N^3 floats, device */
} mp_extract_job;

int mp_extract_batch_plan(const mp_extract_job* jobs, int n_jobs, void* stream, mp_batch** out);
int mp_extract_batch_run(mp_batch* plan, void* stream);
int mp_extract_batch_counts(const mp_batch* plan, int64_t* counts, void* stream);
void mp_extract_batch_destroy(mp_batch* plan);

/* Sample coordinates of grid index range [first, first+n) into xyz [n][3]
 * (gen_mc_coordinate_grid, i3d/meshing.py:11-69, columns 0..2). */
int mp_grid_coords(int N, int grid_mode, int64_t first, int64_t n, float* xyz, void* stream);

/* Multiscale sphere tracing of the whole stack.  Step rule per ray: the
 * renderer's lod-0 phase (step by F_0 - delta_0) until the ray is first inside
 * the coarse band F_0 < band_scale * delta_0; from then on, inside the band
 * step by the sum of all levels (a hit when |sum| < hit_eps), outside it by
 * F_0 (IterateSphereTracing_kernel, kernels.cu:315-398, with its fixed-length
 * phases switched per ray).  A single level steps by F_0.
 * Replaces the renderer's
 * MIPplicit::SdfInference tracing loop and shading (renderer/cuda/state.cu:816-1055;
 * InitializeRays_kernel, IterateSphereTracing_kernel and Shade_kernel in
 * renderer/cuda/kernels.cu:122-487) and the Python multiscale_sphere_tracing
 * (renderer/python/siren-mips.py:126-248).
 * cams: device [n_cams][32] floats, per camera the inverse view matrix then the
 * inverse projection matrix, each column-major 4x4 (glm layout, as the renderer
 * uploads them).  Pixel (i, j) of camera k is ray k*W*H + j*W + i.
 * Outputs (device, one entry per ray): depth (distance from the camera, -1 for a
 * miss), steps, rgba (RGBA8: Lambert shading of the analytic normal, white
 * background), normals [rays][3] (unit, zero for misses; may be NULL).
 * stats (host, MP_MAX_LEVELS + 1 entries, may be NULL): SDF evaluations per
 * level, then the number of hits.  Host-sync point.  Allocates device scratch
 * owned by the stack and kept for later calls: 40 B per ray, 64 B per ray for
 * stacks of three or more levels (the level-by-level launches park rays). */
typedef struct mp_trace_params {
  int width, height;
  int max_steps;       /* e.g. 128 */
  float hit_eps;       /* |F| below this is a hit, e.g. 1e-4 */
  float bound_radius;  /* rays start on / end outside this sphere (1.2 in the renderer) */
  float band_scale;    /* residual levels are used where F_0 < band_scale * delta_0 (1.4) */
  int pixel_center;    /* 1: NDC from pixel centres; 0: the renderer's 2i/W - 1 */
  int flip_y;
  int n_deltas;        /* n_levels - 1 */
  float deltas[MP_MAX_LEVELS]; /* the band widths of levels 0 .. n_levels - 2; the step rule reads
                                  deltas[0] only, as the renderer reads delta_lod_0 only */
} mp_trace_params;

int mp_trace(mp_stack* stack, const float* cams, int n_cams, const mp_trace_params* params,
             float* depth, float* normals, uint32_t* rgba, int32_t* steps, int64_t* stats,
             void* stream);

/* The renderer's fixed-iteration frame (MIPplicit::SdfInference,
 * renderer/cuda/state.cu:816-1055, with InitializeRays_kernel,
 * IterateSphereTracing_kernel, sum_residual_normals and Shade_kernel of
 * renderer/cuda/kernels.cu:28-487; its -iters / -delta / -threshold /
 * -normal_lod / -shading / -skip_lod0 / -no_residual options,
 * renderer/cuda/renderer.cu:754-790).  Every pixel takes iters[0] steps on
 * level 0 (by f0 - delta), iters[1] on levels 0 + 1 and iters[2] on levels
 * 0 (+ 1) + 2 (by the residual sum where f0 < band_scale * delta, by f0
 * elsewhere); pixels whose last evaluated distance is below threshold and
 * whose position is inside [-1, 1]^3 are shaded with the normal of levels
 * 0..normal_lod (one level for standalone models), the rest are white.
 * Pixel (i, j) of camera k is k*W*H + j*W + i; NDC 2i/W - 1 (no pixel centre).
 * Outputs (device): rgba [rays] RGBA8 (truncated as RgbaFloatToInt does),
 * pos [rays][3] final positions, dist [rays] shading distances, normals
 * [rays][3] (unit where shaded, else 0); pos, dist and normals may be NULL.
 * stats (host, may be NULL; then a host-sync point): level evaluations
 * performed [3], shaded pixels.  With stats NULL the call only launches
 * kernels (device scratch grows on the first call of a size), so a frame can
 * be captured in a hipGraph. */
#define MP_SHADE_NORMALS 0
#define MP_SHADE_PHONG 1
typedef struct mp_render_params {
  int width, height;
  int iters[3];         /* steps of the level-0, level-1 and level-2 phases (20/0/0 default, 20/5/5 full) */
  float delta;          /* lod_0_delta: the band of level 0 */
  float threshold;      /* distance_threshold (0.05) */
  int normal_lod;       /* lod_to_show: levels whose normal shades the pixel */
  int shading;          /* MP_SHADE_PHONG or MP_SHADE_NORMALS */
  int residual;         /* 1: finer levels are residuals (the default); 0: standalone SDFs */
  int skip_lod0;        /* 1: no level-0 phase */
  int flip_y;
  int lod2_reads_lod0;  /* 1: level 2's distances read level 0's output, as state.cu:834 binds them */
  float bound_radius;   /* 1.2 */
  float band_scale;     /* 1.4 */
} mp_render_params;

int mp_render(mp_stack* stack, const float* cams, int n_cams, const mp_render_params* params, uint32_t* rgba,
              float* pos, float* dist, float* normals, int64_t* stats, void* stream);

/* Batched tracing of many stacks, one camera each (uneven widths, levels and
 * image sizes; every stack in the same precision): the persistent tracer walks
 * the jobs in order, a wave moving on to the next job once the current one has
 * no rays left for it.  Plan / run / stats as for mp_extract_batch_*: run only
 * launches kernels (hit counts stay on the device; the normal kernels read
 * them), so it can be captured in a hipGraph.  stats: [n_jobs][MP_MAX_LEVELS
 * + 1] evaluations per level, then hits (host-sync point).  No counterpart in
 * the reference, whose renderer draws one model at a time
 * (renderer/cuda/state.cu:816-1055). */
typedef struct mp_trace_batch mp_trace_batch;
typedef struct mp_trace_job {
  mp_stack* stack;
  const float* cam;     /* device [32]: inverse view, inverse projection (column-major) */
  mp_trace_params params;
  float* depth;         /* device [height][width] */
  float* normals;       /* device [height][width][3] or NULL */
  uint32_t* rgba;       /* device [height][width] */
  int32_t* steps;       /* device [height][width] */
} mp_trace_job;

int mp_trace_batch_plan(const mp_trace_job* jobs, int n_jobs, void* stream, mp_trace_batch** out);
int mp_trace_batch_run(mp_trace_batch* plan, void* stream);
int mp_trace_batch_stats(const mp_trace_batch* plan, int64_t* stats, void* stream);
void mp_trace_batch_destroy(mp_trace_batch* plan);

/* The training step of one level (the trainable residual level of
 * experiment_scripts/train_sdf_on_neighborhood_fine.py:220-252) in two calls,
 * replacing the forward, the create_graph input gradient
 * (i3d/diff_operators.py:259-265) and the double backward that autograd runs
 * for loss_add_detail_fine (i3d/loss_functions.py:327-394).  Every product
 * runs on the library's matrix-core GEMMs in the stack's precision (FP32,
 * BF16 or FP16 operands, FP32 accumulation; MP_BF16X3 is refused); the level
 * needs >= 2 sine layers of width <= 512.
 *   mp_train_fwd: f [n] and grad [n][3] (either may be NULL) of level `level`
 *     at points x [n][3] (device).  keep != 0: the layer activations stay in
 *     the stack's training workspace for the next mp_train_bwd of the same
 *     level, x pointer and n (dropped by mp_stack_update or another
 *     mp_train_fwd).
 *   mp_train_bwd: dparams (device, mp_train_param_count floats: every layer's
 *     weights [out][in] in order, then every bias, the .bin order of
 *     mp_level_desc) = the gradient of sum_p [dL_df[p] f_p + dL_dgrad[p] .
 *     grad f_p] with respect to the level's frequency-baked weights and
 *     biases; dL_df [n] or dL_dgrad [n][3] may be NULL (zero).  The layer
 *     activations are recomputed unless kept by mp_train_fwd.  In FP16 the
 *     seeds are scaled on the device by the power of two that brings
 *     max |seed| into [2^10, 2^11) and the gradients back by its inverse (exact),
 *     so that the FP16 adjoints neither flush to zero nor lose precision.
 * Both allocate the training workspace on first use (about 16 n W (2 H - 1)
 * bytes for H sine layers of width W in 16 bit, up to twice that in FP32) and
 * are otherwise asynchronous on the stream. */
int64_t mp_train_param_count(const mp_stack* stack, int level);
int mp_train_fwd(mp_stack* stack, int level, const float* x, int64_t n, float* f, float* grad, int keep,
                 void* stream);
int mp_train_bwd(mp_stack* stack, int level, const float* x, int64_t n, const float* dL_df, const float* dL_dgrad,
                 float* dparams, void* stream);

/* Training-data generation (the step before the training kernels).
 * mp_mesh: a triangle mesh for unsigned point-to-mesh distances, replacing
 * open3d's RaycastingScene (add_triangles / compute_distance) as the trainers
 * use it (experiment_scripts/train_sdf_on_neighborhood_fine.py:145-146,
 * experiment_scripts/utils.py:77-78).  verts [n_verts][3] float32 and faces
 * [n_faces][3] int32 are HOST arrays; the bounding-volume hierarchy is built
 * on the host and uploaded (host-sync point, allocates). */
typedef struct mp_mesh mp_mesh;
int mp_mesh_create(const float* verts, int64_t n_verts, const int32_t* faces, int64_t n_faces, void* stream,
                   mp_mesh** out);
void mp_mesh_destroy(mp_mesh* mesh);
int mp_mesh_info(const mp_mesh* mesh, int64_t* n_tris, int64_t* n_nodes, int* depth);
/* dist [n] = unsigned distance of device points pts [n][3] to the mesh. */
int mp_mesh_distance(const mp_mesh* mesh, const float* pts, int64_t n, float* dist, void* stream);
/* calc_deltas (experiment_scripts/utils.py:46-89): vert [V][cols] device
 * (x, y, z, nx, ny, nz, ...); out [V][7] = (x, y, z, unit normal, delta)
 * where delta is searched along the normal (start max_delta / n_iters, at
 * most n_iters steps while max(delta) < max_delta, eps 2e-4) for the offset
 * whose distance to the mesh equals it.  iters_done (host, may be NULL):
 * iterations run.  Host-sync point (one per iteration, as the reference's
 * loop test), allocates scratch. */
int mp_calc_deltas(const mp_mesh* mesh, const float* vert, int64_t V, int cols, float max_delta, int n_iters,
                   float* out, int* iters_done, void* stream);
/* create_training_data_on_neighborhood (experiment_scripts/utils.py:92-169)
 * given its random draws: idx [n] int64 sampled vertex rows (NULL: rows 0..n-1,
 * the reference's n == V case), rnd [n][3] its uniform numbers.  Writes
 * pts / normals [m][3] and sdfs [m], m = 3n with dithering, else 2n: the
 * vertices (sdf 0), offsets along the normal by delta_n * rnd_z (sdf the same;
 * delta_n = the vertex's last column if use_neigh_delta, else delta_eps) and
 * the dithered band points p + delta / 2 n + delta rnd (sdf -1). */
int mp_neighborhood_samples(const float* vert, int64_t V, int cols, const int64_t* idx, const float* rnd, int64_t n,
                            float delta, float delta_eps, int use_neigh_delta, int dithering, float* pts,
                            float* normals, float* sdfs, void* stream);

/* Marching cubes of a [nz][ny][nx] float32 device volume at `level`: replaces
 * skimage.measure.marching_cubes(volume, level) as convert_sdf_samples_to_ply
 * calls it (i3d/meshing.py:296-298; Lewiner method, 'descent' winding,
 * degenerate triangles kept).  Two calls: mp_mc_count (host-sync point) returns
 * the vertex and face counts, the caller allocates, mp_mc_emit writes
 *   verts   [V][3] float32, voxel units, (axis0, axis1, axis2) order,
 *   faces   [F][3] int32 vertex ids,
 *   normals [V][3] float32 unit -grad(volume) (may be NULL),
 *   values  [V] float32 local value range (may be NULL).
 * Vertex and face ids are deterministic: exclusive prefix sums over the
 * active points (those owning vertices or triangles) taken in a fixed order,
 * tiles of 64 x 8 x 16 points (x, y, z) in grid order, and inside a tile by
 * plane, row, then column.  The triangle set equals scikit-image's; only the
 * numbering differs.
 * verts and faces may be NULL only when the counts are zero.
 * `ws` is a device workspace of mp_mc_workspace_size(nz, ny, nx) bytes, kept
 * between the two calls. */
size_t mp_mc_workspace_size(int nz, int ny, int nx);
int mp_mc_count(const float* vol, int nz, int ny, int nx, float level, int64_t* n_verts,
                int64_t* n_faces, void* ws, void* stream);
int mp_mc_emit(const float* vol, int nz, int ny, int nx, float level, float* verts, int32_t* faces,
               float* normals, float* values, void* ws, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* MPLICITS_H */
