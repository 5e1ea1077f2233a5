// sr_kernels.h -- kernel argument blocks and host-side launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "sr_device.h"

struct PathArgs {
  SceneDev s;
  float* sum;                     // [W][H][B] block sums of this launch (zero on entry)
  unsigned long long* counters;   // segments, traversals, non-finite drops
  uint32_t* queue;                // next work item (zero on entry)
  uint32_t* overflow;             // traversal stack overflow flag
  uint32_t n_items;               // W * H * n_blocks
  int W, H, depth, first_block, n_blocks, tiled;
  uint32_t seed_lo, seed_hi, stream;
};

struct ResolveParams {
  float cmf[SR_MAX_BANDS * 3];
  float T[9];
  float step, norm, exposure, gamma;
};

int sr_launch_trace(const SceneDev& s, int n, const float* o, const float* d, float* t, int* prim, float* uv,
                    uint32_t* overflow, hipStream_t st);
int sr_path_grid(int* grid_out);
int sr_launch_path(const PathArgs& a, int grid, hipStream_t st);
int sr_launch_finalize(float* mean, float* sum, size_t n, int first, int nb, hipStream_t st);
int sr_launch_fill(float* p, size_t n, float v, hipStream_t st);
int sr_launch_resolve(const float* acc, int W, int H, int B, const ResolveParams& p, float* out, int OW, int OH,
                      hipStream_t st);
