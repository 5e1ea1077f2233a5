// sr_kernels.h -- kernel argument blocks and host-side launchers.
#pragma once
#include <vector>
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/spectral_render.h"
#include "sr_device.h"

#define SR_MAX_PARAMS 1024

// Bands per work item of the path kernel: a lane runs that many single-band
// paths of one pixel and block back to back (the same pixel's rays stay
// close to each other), and a block of W x H pixels is W * H * ceil(B / this)
// items, small enough for the persistent grid to balance at a launch's end.
#ifndef SR_BANDS_PER_ITEM
#define SR_BANDS_PER_ITEM 8
#endif
static inline uint32_t sr_items_per_pixel(int bands) { return (uint32_t)((bands + SR_BANDS_PER_ITEM - 1) / SR_BANDS_PER_ITEM); }

// Scene records built on the device by sr_build_scene_device (sr_build.hip);
// every pointer is a hipMalloc allocation the caller takes over (or frees
// with sr_free_scene_out).
struct DeviceSceneOut {
  void* nodes = nullptr;
  size_t nodes_bytes = 0;
  int n_nodes = 0, depth = 0;
  void* tris = nullptr;
  size_t tris_bytes = 0;
  void* shade = nullptr;
  size_t shade_bytes = 0;
  void* lights = nullptr;
  size_t lights_bytes = 0;
  int max_lights = 0;
  void* lverts = nullptr;
  size_t lverts_bytes = 0;
  int n_slots = 0;
  void* light_count = nullptr;
  size_t light_count_bytes = 0;
  void* light_area = nullptr;
  size_t light_area_bytes = 0;
  std::vector<int> h_light_count;
  std::vector<float> h_light_area;
  // host copies for the light tree build: slot faces, their materials, lverts
  std::vector<int> h_slots, h_slot_mat;
  std::vector<float> h_lverts;  // [n_slots][3][4]
};
// tris_v / tris_n [n][3][3] f32 and tri_mat [n] in device memory; mat_i32
// [n_mats][2][B] on the host.  Returns 0 or an SR_E* code with `err` set.
int sr_build_scene_device(const float* tris_v, const float* tris_n, const int* tri_mat, int n, const int* mat_i32,
                          int n_mats, int B, hipStream_t st, DeviceSceneOut* out, char* err, size_t errlen);
void sr_free_scene_out(DeviceSceneOut* o);

// Differentiable parameters of the adjoint (sr_render_adjoint): per-band
// albedo of any material with alb_off[m] >= 0 (param index alb_off[m] + band),
// the roughness of one material, the HG g of one material and the per-band
// sigma_a of one material (param index sig_off + band).
struct AdjMap {
  const float* d_accum;  // [W][H][B] dL / d accum
  float* grad;           // [n_params], accumulated into
  const int* alb_off;    // [M] device array
  int rough_mat, rough_idx, g_mat, g_idx, sig_mat, sig_off, n_params;
  int flags;             // SR_ADJ_SKIP_* term mask (diagnostics); 0 = all terms
  float scale;           // 1 / (first_block + n_blocks)
};

// Where a launch keeps a job's small scene tables -- the ones the path loop
// reads at every segment: materials, per-band wavelengths, light counts,
// areas and CDFs, light vertices, each material's EEM id -- in LDS, after
// the traversal stacks.  Byte offsets from the region's start, all
// multiples of 16; bytes = 0 leaves every table in global memory (tables
// larger than the budget).
struct LdsTables {
  uint32_t bytes;
  uint32_t mats, lights, lverts, lam_m, lam_nm, light_area, light_count, mat_eem;
  uint32_t n_mats, n_lights, n_lverts;  // records copied: MatRec, LightRec, float4
};
// LDS the tables may take beside a 256-lane block's stacks so that three
// blocks (12 waves) still share a CU's 160 KB; a smaller block gets its
// share of it, and the adjoint's gradient table comes out of it.
#define SR_LDS_TABLE_BUDGET 13312

// Work queues per job, one per XCD.  The blocks of a launch are dealt to
// the 8 XCDs in turn (block b runs on XCD b % 8) and each XCD has its own
// L2: queue x holds the x-th eighth of every spectral block's tiles, so the
// waves of one XCD work on one screen region at a time.  A wave drains its
// XCD's queue first, then the others in turn.
#define SR_QUEUES 8

// One scene render: the unit of work of the persistent path kernel.  A
// single render is a table of one job; a batch is a table of many.
struct Job {
  SceneDev s;
  float* sum;        // [W][H][B] block sums of this launch (zero on entry)
  uint32_t n_items;  // W * H * n_blocks * items per pixel
  int W, H, depth, first_block, n_blocks, tiled;
  uint32_t seed_lo, seed_hi, stream;
  int nee_tree;      // 1: NEE picks lights through the light tree (SR_RENDER_LIGHT_TREE)
  int sss;           // 1: subsurface random walk (SR_RENDER_SUBSURFACE; the walk kernel only)
  int end_dead;      // 1: a path ends once its throughput is exactly 0 (SR_RENDER_END_DEAD_PATHS)
  uint32_t per_block;  // items per spectral block: W * H * items per pixel
  // Path totals [n_blocks][W][H][B] (null: none).  The forward writes each
  // path's summed contributions there, or NaN when some of them went to
  // another band (fluorescence); the adjoint takes a finite total as its
  // first pass and replays only the other paths twice.
  float* rec;
  // queue x: items [q_off[x], q_off[x] + q_len[x]) of every spectral block
  // (whole 64-item tile groups when tiled), q_len[x] * n_blocks in all
  uint32_t q_off[SR_QUEUES], q_len[SR_QUEUES];
  LdsTables tab;     // this job's tables in the launch's LDS region (PathArgs::tab_bytes > 0)
};

struct PathArgs {
  const Job* jobs;                // [n_jobs] device table
  int n_jobs;
  uint32_t* queues;               // [n_jobs][SR_QUEUES] next item of each queue (zero on entry)
  uint32_t* done_items;           // [n_jobs] completed items (zero on entry)
  unsigned long long* clocks;     // [n_jobs + 1] wall clock: [0] launch start (min), [1 + j] job j done
  unsigned long long* counters;   // [n_jobs][4]: segments, traversals, non-finite drops
  uint32_t* overflow;             // traversal stack overflow flag
  uint32_t chunk;                 // items claimed per queue atomic (per wave)
  uint32_t tab_bytes;             // LDS region for the jobs' tables (every job's tab.bytes fits); 0: global
  uint32_t record;                // 1: the kernel built to write Job::rec (a single recording job)
  AdjMap adj;                     // adjoint only (single job)
};

// The LDS layout of a scene's small tables (n_lights = B * max(max_lights, 1)
// light records, n_slots light slots of three vertices), or bytes = 0 when
// they do not fit the budget of a `block`-lane block.
LdsTables sr_lds_tables(int n_mats, int B, int n_lights, int n_slots, int block, bool adjoint);
size_t sr_lds_table_budget(int block, bool adjoint);

struct ResolveParams {
  float cmf[SR_MAX_BANDS * 3];
  float T[9];
  float step, norm, exposure, gamma;
};

int sr_launch_trace(const SceneDev& s, int n, const float* o, const float* d, float* t, int* prim, float* uv,
                    uint32_t* overflow, hipStream_t st);
int sr_path_grid(int* grid_out);
int sr_adjoint_grid(int* grid_out);
int sr_path_grid_bs(int block, int* grid_out);
size_t sr_path_lds_bytes(int block, bool adjoint, uint32_t table_bytes = 0);
int sr_launch_path(const PathArgs& a, int grid, int block, hipStream_t st);
int sr_launch_adjoint(const PathArgs& a, int grid, int block, hipStream_t st);
int sr_path_sss_grid(int* grid_out);
int sr_path_sss_grid_bs(int block, int* grid_out);
int sr_launch_path_sss(const PathArgs& a, int grid, int block, hipStream_t st);
int sr_launch_fill_u32(uint32_t* p, size_t n_words, uint32_t v, hipStream_t st);
int sr_launch_finalize(float* mean, float* sum, size_t n, int first, int nb, hipStream_t st);
int sr_launch_fill(float* p, size_t n, float v, hipStream_t st);
// tmp: W * H * 3 floats of scratch for an upsampled display (OW, OH != W, H),
// or null (then each display pixel resolves its render pixel itself).
int sr_launch_resolve(const float* acc, int W, int H, int B, const ResolveParams& p, float* out, int OW, int OH,
                      float* tmp, hipStream_t st);
// tmp: W * H * 3 floats of scratch for an upsampled display, or null.
int sr_launch_normals(const SceneDev& s, int W, int H, int OW, int OH, float* out, float* tmp, uint32_t* overflow,
                      hipStream_t st);
int sr_launch_rgba8(const float* rgb, size_t n, void* out, hipStream_t st);
