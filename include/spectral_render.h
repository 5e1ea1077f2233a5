/* spectral_render.h -- C ABI of the MI355X spectral path tracer.
 *
 * Drop-in boundary for the hot path of Xyzonox/taichi-spectral-render.  The
 * reference has no FFI: its boundary is the Python scene dict consumed by
 * setup_scene (main.py:509-631) and the kernel-call protocol of its main
 * loop (main.py:1826-1837: cast_rays per band, average_and_reset_block and
 * color_rays per block).  Each entry point below names the reference code it
 * replaces.
 *
 * Conventions
 *   - Every function returns 0 on success or a negative SR_E* code, and
 *     records a message readable with sr_last_error() (thread-local).
 *   - Handles are opaque, bound to the HIP device current at sr_create, and
 *     not thread-safe: use each from one host thread.
 *   - Host arrays are borrowed for the duration of the call and copied.
 *   - Arrays named dev_* are device pointers on the handle's device (for
 *     example a torch tensor's data_ptr()).  Work is ordered on the handle's
 *     stream (sr_set_stream); calls return once work is enqueued, except
 *     those documented as blocking.
 *   - Accumulator layout is the reference's field layout [W][H][B] (pixel
 *     (i, j) with j pointing up, main.py:365), f32.
 */
#ifndef SPECTRAL_RENDER_H
#define SPECTRAL_RENDER_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sr_handle sr_handle;

#define SR_OK 0
#define SR_EINVAL -1   /* bad argument */
#define SR_ESTATE -2   /* call out of order (no scene / no resolution) */
#define SR_EHIP -3     /* HIP runtime error */
#define SR_ENOMEM -4   /* allocation failed */
#define SR_EOVERFLOW -5 /* traversal stack overflow detected */

/* sr_trace_closest flags */
#define SR_DEVICE_PTRS 1

/* sr_render_adjoint flags: leave out one family of gradient terms (for
 * diagnosing estimator variance; 0 computes the full gradient). */
#define SR_ADJ_SKIP_HG_SCORE 1     /* d log p / d g of HG direction sampling */
#define SR_ADJ_SKIP_G_NEE 2        /* d / d g of the phase function in NEE */
#define SR_ADJ_SKIP_SIG_SCORE 4    /* free-flight and absorb/scatter choice scores */
#define SR_ADJ_SKIP_SIG_FACTOR 8   /* sigma_a in estimators and shadow transmittance */
#define SR_ADJ_SKIP_ROUGH_SCORE 16 /* GGX half-vector sampling score */

/* Last error message of this thread ("" when none). */
const char* sr_last_error(void);

/* ABI version (major * 100 + minor). */
int sr_version(void);

/* Creates a renderer for `bands` spectral bands on [wl_min, wl_max] nm
 * (SPECTRAL_BANDS / WAVELENGTH_MIN / WAVELENGTH_MAX, main.py:26-28;
 * band k is wl_min + k * (wl_max - wl_min) / (bands - 1), main.py:34). */
int sr_create(int bands, float wl_min, float wl_max, sr_handle** out);
int sr_destroy(sr_handle* h);

/* Orders all work of `h` on `stream` (a hipStream_t; NULL = the handle's own
 * default-priority stream).  Work already queued on the previous stream
 * completes before work queued on the new one (an event wait, no host
 * synchronization). */
int sr_set_stream(sr_handle* h, void* stream);

/* Replaces setup_scene (main.py:509-631) and the field allocation of
 * main.py:1683-1709.
 *   tris_v  [n_tris][3][3] f32  vertex positions per face
 *   tris_n  [n_tris][3][3] f32  per-vertex shading normals (the face normal
 *                               repeated for meshes without vertex_normals)
 *   tri_mat [n_tris]       i32  material id per face
 *   mat_f32 [n_mats][6][B] f32  refractive_index, extinction_coefficient,
 *                               roughness, scattering_coefficient, emission,
 *                               anistropy_factor (main.py:40-45)
 *   mat_i32 [n_mats][2][B] i32  is_true_volume, surface_emission (main.py:49-50)
 *   mat_eem [n_mats]       i32  eem_id (0 = none)
 *   eem     [n_eem][B][B]  f32  reemission_matrix (main.py:372); row 0 unused
 *   cam     [9]            f32  origin, lookat, up (main.py:242-246)
 * Builds face normals/areas (main.py:526-536), the per-band area-CDF light
 * lists (main.py:590-627) and the 4-wide BVH, and uploads everything. */
int sr_upload_scene(sr_handle* h, const float* tris_v, const float* tris_n, const int* tri_mat, int n_tris,
                    const float* mat_f32, const int* mat_i32, const int* mat_eem, int n_mats, const float* eem,
                    int n_eem, const float* cam);

/* setup_scene (main.py:509-631, the same records as sr_upload_scene) built on
 * the GPU: dev_tris_v / dev_tris_n [n_tris][3][3] f32 and dev_tri_mat
 * [n_tris] i32 are device pointers on the handle's device (e.g. torch
 * tensors); materials, EEM and camera are host arrays as for
 * sr_upload_scene.  Every compiled array -- face records, light lists, the
 * 4-wide SAH BVH and the leaf-order triangle records -- is byte-identical
 * to the host build (sr_scene_array reads them back); about 27x faster
 * than the host build for 1M faces.
 * SR_EINVAL for a material id outside [0, n_mats), SR_EOVERFLOW if the tree
 * is deeper than the traversal stack allows (sr_upload_scene then). */
int sr_upload_scene_gpu(sr_handle* h, const float* dev_tris_v, const float* dev_tris_n, const int* dev_tri_mat,
                        int n_tris, const float* mat_f32, const int* mat_i32, const int* mat_eem, int n_mats,
                        const float* eem, int n_eem, const float* cam);

/* update_camera (main.py:633-640): new origin/lookat/up for the next render. */
int sr_set_camera(sr_handle* h, const float* cam);

/* Material table update without rebuilding geometry (same layout as upload).
 * SR_EINVAL if any (material, band) surface_emission flag differs from the
 * uploaded scene's: the light lists, their CDFs, the emitting areas and the
 * light tree were built from those flags (upload the scene again). */
int sr_update_materials(sr_handle* h, const float* mat_f32, const int* mat_i32, const int* mat_eem, int n_mats,
                        const float* eem, int n_eem);

/* Scene statistics: BVH node count, 4-wide depth, max lights per band,
 * device bytes held by the scene. */
int sr_scene_info(sr_handle* h, int* n_nodes, int* bvh_depth, int* max_lights, long long* device_bytes);

/* Copies one compiled scene array to host memory, as the reference's
 * fields can be read back (triangles, bvh_nodes and the light tables,
 * main.py:1700-1709; Taichi's field.to_numpy()).  which: SR_ARRAY_*.
 * *bytes receives the array's size; with dst NULL only the size is
 * returned.  SR_EINVAL if capacity < size, SR_ESTATE without a scene.
 * Blocks until the copy is done. */
#define SR_ARRAY_NODES 0       /* Node4[n_nodes] */
#define SR_ARRAY_TRIS 1        /* TriRec[n_tris + 3], leaf order */
#define SR_ARRAY_SHADE 2       /* 4 x float4 per face, face order */
#define SR_ARRAY_LIGHTS 3      /* LightRec[bands][max(max_lights, 1)] */
#define SR_ARRAY_LVERTS 4      /* 3 x float4 per light slot */
#define SR_ARRAY_LIGHT_COUNT 5 /* int[bands] */
#define SR_ARRAY_LIGHT_AREA 6  /* float[bands] */
#define SR_ARRAY_LIGHT_TREE 7  /* LightNode[light tree nodes] (64 B each: sphere, cone, links), sibling pairs */
#define SR_ARRAY_LIGHT_POWER 8 /* float[bands][light tree nodes] */
#define SR_ARRAY_LIGHT_KEYS 9  /* {face, leaf}[emissive faces], sorted by face */
int sr_scene_array(sr_handle* h, int which, void* dst, size_t capacity, size_t* bytes);

/* Closest hit for n rays: trace (main.py:831-864) as a batch.  out_t = 1e9
 * and out_prim = -1 on a miss; out_prim is the ORIGINAL face id (the order
 * of tris_v at upload); ties in t resolve to the smallest face id.  With
 * SR_DEVICE_PTRS all arrays are device pointers and the call is
 * asynchronous; otherwise they are host arrays and the call blocks. */
int sr_trace_closest(sr_handle* h, int n, const float* orig, const float* dir, float* out_t, int* out_prim,
                     float* out_uv, int flags);

/* Renders spectral blocks [first_block, first_block + n_blocks): for each
 * block, each band k of 0..B-1 and each pixel, one single-band path of
 * cast_rays (main.py:1502-1590) with at most max_depth segments, then folds
 * the blocks into the running mean dev_accum [W][H][B] exactly as
 * average_and_reset_block (main.py:1592-1607) does once per block.
 * Randomness: Philox-4x32-10 keyed by seed, counter (i*H + j, block,
 * band | stream << 8, call/4).  The work items of a call -- W * H *
 * n_blocks * ceil(B / 8), each up to 8 bands of one pixel and block -- must
 * number fewer than 2^31.  flags: 0 or any of SR_RENDER_LIGHT_TREE,
 * SR_RENDER_SUBSURFACE, SR_RENDER_END_DEAD_PATHS. */
#define SR_RENDER_LIGHT_TREE 1 /* NEE picks its light through the light tree (importance by distance and
                                  orientation; the TODO of main.py:541) instead of the area CDF
                                  (main.py:1306-1313): same expected image, lower variance with many
                                  lights.  Default 0 reproduces the reference's choice exactly. */
#define SR_RENDER_SUBSURFACE 2 /* opaque materials with scattering_coefficient > 0 scatter by a random walk
                                  inside the mesh (sigma_s, 4 pi k / lambda, HG g of the band; Lambertian
                                  entry and exit) instead of the single-scatter Oren-Nayar lobe: the TODO
                                  of main.py:1066.  Meshes must be closed (a walk that leaves an open
                                  mesh ends).  Not in the reference; off by default. */
#define SR_RENDER_END_DEAD_PATHS 4 /* a path ends once its throughput is exactly 0 (a GGX reflection below
                                      the horizon, a refraction into a conductor, a medium that absorbs
                                      everything) instead of running to max_depth as cast_rays does
                                      (main.py:1534): every later contribution of such a path is 0, so the
                                      image is the same (the non-finite products 0 * inf the kernel
                                      drops are not formed); fewer segments are executed and counted.
                                      Off by default: sr_stats then counts the reference's segments. */
int sr_render(sr_handle* h, int width, int height, uint64_t seed, uint32_t stream, int first_block, int n_blocks,
              int max_depth, int flags, float* dev_accum);

/* sr_render without any device allocation -- no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
 * sr_render allocates the handle's block-sum buffer on its first call and
 * whenever the resolution grows; this entry point never calls hipMalloc,
 * hipFree or any allocating routine.  Its block sums, job table, work
 * counters and overflow flag live in one fixed device array per GPU
 * (SR_NOALLOC_MAX_ELEMS floats of block sums, 16 MiB, plus a few hundred
 * bytes, resident while the library is loaded).  Same arguments, kernels
 * and results as sr_render; width * height * bands must not exceed
 * SR_NOALLOC_MAX_ELEMS (SR_EINVAL otherwise).  The array is shared by every
 * call on that GPU and keeps the last call's contents (the block sums are
 * cleared at the start of each call): calls that overlap on one GPU --
 * from two host threads, or on two streams without ordering -- race.
 * sr_stats reports this call's counters until another sr_render_noalloc
 * call on the GPU overwrites them. */
#define SR_NOALLOC_MAX_ELEMS (1u << 22)
int sr_render_noalloc(sr_handle* h, int width, int height, uint64_t seed, uint32_t stream, int first_block,
                      int n_blocks, int max_depth, int flags, float* dev_accum);

/* Path-replay adjoint of sr_render (no reference counterpart: the reference
 * has no autodiff).  Replays blocks [first_block, first_block + n_blocks)
 * with the same seed/stream as the forward call and ADDS to dev_grad
 * [n_params] the gradient of L = sum(dev_d_accum * accum), where accum is
 * the running mean after those blocks (so each block weighs
 * 1 / (first_block + n_blocks)).  Parameters:
 *   albedo  per band, for every material m with alb_off[m] >= 0 (host array
 *           [n_mats]); param index alb_off[m] + band; the albedo is
 *           s / (s + 4 pi k / lambda + 1e-6) of main.py:1070, 1395;
 *   roughness of material rough_mat (-1: none) at param index rough_idx;
 *   anistropy_factor g of material g_mat at g_idx;
 *   sigma_a = 4 pi k / lambda per band of material sig_mat at sig_off + band.
 * n_params <= 1024.  flags: SR_ADJ_SKIP_* mask, 0 for the full gradient. */
int sr_render_adjoint(sr_handle* h, int width, int height, uint64_t seed, uint32_t stream, int first_block,
                      int n_blocks, int max_depth, const float* dev_d_accum, const int* alb_off, int rough_mat,
                      int rough_idx, int g_mat, int g_idx, int sig_mat, int sig_off, int n_params, int flags,
                      float* dev_grad);

/* sr_render that also records each path's total for the adjoint (no
 * reference counterpart).  dev_path_totals [n_blocks][W][H][B] f32 (device)
 * receives, for the path of block first_block + k, pixel (i, j) and band b,
 * the sum of its contributions -- or NaN when fluorescence sent some of
 * them to another band.  flags: 0 or SR_RENDER_END_DEAD_PATHS (the adjoint
 * replays the area-CDF light pick without the subsurface walk). */
int sr_render_record(sr_handle* h, int width, int height, uint64_t seed, uint32_t stream, int first_block,
                     int n_blocks, int max_depth, int flags, float* dev_accum, float* dev_path_totals);

/* sr_render_adjoint given the totals an sr_render_record call with the same
 * scene, parameters, seed, stream and blocks wrote: a path with a finite
 * total is replayed once instead of twice (its first pass only sums
 * dL/dacc x contribution over the path).  Same arguments and gradient as
 * sr_render_adjoint otherwise. */
int sr_render_adjoint_record(sr_handle* h, int width, int height, uint64_t seed, uint32_t stream, int first_block,
                             int n_blocks, int max_depth, const float* dev_d_accum, const int* alb_off,
                             int rough_mat, int rough_idx, int g_mat, int g_idx, int sig_mat, int sig_off,
                             int n_params, int flags, const float* dev_path_totals, float* dev_grad);

/* color_rays (main.py:1609-1629): dev_accum [W][H][B] -> dev_rgb
 * [out_w][out_h][3] with nearest upsampling.  cmf [B][3] (host),
 * xyz2rgb [3][3] row-major (host), norm = sum(cmf_y) * step. */
int sr_resolve(sr_handle* h, const float* dev_accum, int width, int height, const float* cmf, const float* xyz2rgb,
               float norm, float exposure, float gamma, int out_w, int out_h, float* dev_rgb);

/* visualize_normals (main.py:1632-1660), the geometry display mode: a
 * pixel-centre camera ray per render pixel (width x height grid), colour =
 * shading normal * 0.5 + 0.5, black where the ray misses, nearest-upsampled
 * into dev_rgb [out_w][out_h][3] f32 as color_rays does. */
int sr_render_normals(sr_handle* h, int width, int height, int out_w, int out_h, float* dev_rgb);

/* The display conversion of a [n_pixels][3] f32 image in [0, 1] (what the
 * reference hands to its GUI, main.py:1845) to [n_pixels] RGBA8: clamped to
 * [0, 1], x * 255 rounded to nearest, alpha 255, NaN shown as 0. */
int sr_to_rgba8(sr_handle* h, const float* dev_rgb, int64_t n_pixels, void* dev_rgba);

/* Counters of the last sr_render (blocking): path segments (extension
 * traversals, one per executed path-loop iteration), all traversals
 * (extension + shadow), dropped non-finite contributions. */
int sr_stats(sr_handle* h, uint64_t* segments, uint64_t* traversals, uint64_t* nonfinite);

/* Device time of the last sr_render's (or sr_render_adjoint's) path kernel in ms (blocking; HIP
 * events recorded on the handle's stream around the launch). */
int sr_last_kernel_ms(sr_handle* h, float* ms);

/* Blocks until all work of h is done. */
int sr_sync(sr_handle* h);

/* ---- Batched multi-scene launch (no reference counterpart: the reference
 * renders one scene held in global fields, main.py:361-394, 1683-1709, with
 * one kernel launch per band, main.py:1831).  A batch holds jobs -- each an
 * sr_render call on its own scene handle -- and runs them all in ONE
 * persistent launch fed from per-job device work queues, optionally
 * captured once as a hipGraph and replayed. */
typedef struct sr_batch sr_batch;

#define SR_BATCH_GRAPH 1 /* sr_batch_launch: capture on first launch, replay after */

int sr_batch_create(sr_batch** out);
int sr_batch_destroy(sr_batch* b);
/* Adds a job with sr_render's arguments (flags: SR_RENDER_*; a batch with a
 * SR_RENDER_SUBSURFACE job runs the walk kernel, which walks only in the
 * jobs that ask for it).  A handle may appear once. */
int sr_batch_submit(sr_batch* b, sr_handle* h, int width, int height, uint64_t seed, uint32_t stream,
                    int first_block, int n_blocks, int max_depth, int flags, float* dev_accum);
int sr_batch_clear(sr_batch* b);
/* Launch shape: workgroup size (64/128/256), grid as a percentage of the
 * occupancy-limited resident grid, items claimed per queue atomic. */
int sr_batch_config(sr_batch* b, int block, int grid_pct, int chunk);
/* Enqueues the batch on the batch's own stream (asynchronous): every job's
 * submitted block range.  A job whose scene handle changed since the last
 * launch (upload, camera, materials, a larger block-sum buffer) has its
 * job-table entry re-filled (and a captured graph captured again).  The
 * queue counter of a job must not wrap: SR_EINVAL when items + (waves + 1) x
 * chunk reaches 2^32. */
int sr_batch_launch(sr_batch* b, int flags);
/* Blocks; total device time, and per job (arrays of length = jobs, may be
 * NULL): completion latency from the launch's first wave (device wall
 * clock), path segments and traversals. */
int sr_batch_wait(sr_batch* b, float* total_ms, float* job_ms, uint64_t* job_segments, uint64_t* job_traversals);

#ifdef __cplusplus
}
#endif

#endif /* SPECTRAL_RENDER_H */
