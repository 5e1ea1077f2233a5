// sr_kernels.hip -- HIP kernels for gfx950: closest-hit batch trace, the
// persistent spectral path loop, the block-mean update and the CMF resolve.
#include <hip/hip_runtime.h>

#include "sr_device.h"
#include "sr_kernels.h"

// ====================================================================== trace
// sr_trace_closest: one lane per ray (the parity entry point).
__global__ __launch_bounds__(256) void sr_trace_kernel(SceneDev s, int n, const float* __restrict__ orig,
                                                       const float* __restrict__ dir, float* __restrict__ out_t,
                                                       int* __restrict__ out_prim, float* __restrict__ out_uv,
                                                       uint32_t* overflow) {
  __shared__ int lds_stack[SR_STACK_SIZE * 256];
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  f3 o = mk(orig[r * 3], orig[r * 3 + 1], orig[r * 3 + 2]);
  f3 d = mk(dir[r * 3], dir[r * 3 + 1], dir[r * 3 + 2]);
  // walked in suspended pieces, as the path kernels walk theirs: a piece
  // ends once fewer than SR_TRACE_KEEP lanes of the wave still walk, and the
  // next resumes the unfinished walks (the hit equals an uninterrupted walk's)
  TraceState ts;
  trace_begin(ts, 1e9f, INFINITY);
  bool done = false;
  while (__ballot(!done))
    if (trace_resume(s.nodes, s.tris, o, d, lds_stack + threadIdx.x, blockDim.x, overflow, ts, !done, SR_TRACE_KEEP))
      done = true;
  const HitD h = ts.h;
  out_t[r] = h.t;
  out_prim[r] = h.prim;
  out_uv[r * 2] = h.prim >= 0 ? h.u : 0.0f;
  out_uv[r * 2 + 1] = h.prim >= 0 ? h.v : 0.0f;
}

// ====================================================================== block mean
// average_and_reset_block (main.py:1592-1607) for n blocks at once:
// mean = (mean * first + sum) / (first + n); sum = 0.
__global__ void sr_finalize_kernel(float* __restrict__ mean, float* __restrict__ sum, size_t n, float first,
                                   float total) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t step = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += step) {
    mean[i] = (mean[i] * first + sum[i]) / total;
    sum[i] = 0.0f;
  }
}

__global__ void sr_fill_kernel(float* __restrict__ p, size_t n, float v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t step = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += step) p[i] = v;
}

// ====================================================================== resolve
// color_rays, main.py:1609-1629: CMF projection, XYZ -> RGB, Reinhard,
// gamma, nearest upsample from the render grid to the display grid.
__device__ __forceinline__ void resolve_pixel(const float* a, int B, const ResolveParams& p, float* o) {
  float x = 0.f, y = 0.f, z = 0.f;
  for (int k = 0; k < B; ++k) {
    float v = a[k];
    x += v * p.cmf[k * 3 + 0];
    y += v * p.cmf[k * 3 + 1];
    z += v * p.cmf[k * 3 + 2];
  }
  const float sc = p.step / p.norm;
  x *= sc; y *= sc; z *= sc;
  float rgb[3];
  for (int r = 0; r < 3; ++r) {
    float c = p.T[r * 3 + 0] * x + p.T[r * 3 + 1] * y + p.T[r * 3 + 2] * z;
    float e = p.exposure * c;
    float m = e / (e + 1.0f);
    rgb[r] = powf(m, 1.0f / p.gamma);
  }
  o[0] = rgb[0]; o[1] = rgb[1]; o[2] = rgb[2];
}

// Display grid != render grid: one thread per display pixel.
__global__ void sr_resolve_kernel(const float* __restrict__ acc, int W, int H, int B, ResolveParams p,
                                  float* __restrict__ out, int OW, int OH) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= OW * OH) return;
  const int i = idx / OH, j = idx % OH;
  const int ri = (int)((float)i / (float)OW * (float)W);
  const int rj = (int)((float)j / (float)OH * (float)H);
  resolve_pixel(acc + ((size_t)ri * H + rj) * B, B, p, out + (size_t)idx * 3);
}

// Display grid == render grid: the 256 pixels of a workgroup own a
// contiguous slice of the [W][H][B] accumulator; it is staged in LDS with
// 16-byte loads (coalesced), then each thread reads its pixel's B values
// there (B odd or not, consecutive pixels start B words apart).
__global__ __launch_bounds__(256) void sr_resolve_tile_kernel(const float* __restrict__ acc, int npix, int B,
                                                              ResolveParams p, float* __restrict__ out) {
  extern __shared__ float tile[];
  const int p0 = blockIdx.x * 256;
  const int np = min(256, npix - p0);
  const size_t base = (size_t)p0 * B;  // a multiple of 256 B floats: 16-byte aligned
  const int nf = np * B;
  const float4* a4 = reinterpret_cast<const float4*>(acc + base);
  float4* t4 = reinterpret_cast<float4*>(tile);
  for (int t = threadIdx.x; t < nf / 4; t += 256) t4[t] = a4[t];
  for (int t = (nf & ~3) + threadIdx.x; t < nf; t += 256) tile[t] = acc[base + t];
  __syncthreads();
  if ((int)threadIdx.x < np) resolve_pixel(tile + threadIdx.x * B, B, p, out + (size_t)(p0 + threadIdx.x) * 3);
}

// Nearest-neighbour display upsample of a resolved [W][H][3] image, with the
// render pixel chosen exactly as sr_resolve_kernel chooses it
// (main.py:1614-1615).
__global__ void sr_upsample_rgb_kernel(const float* __restrict__ rgb, int W, int H, float* __restrict__ out, int OW,
                                       int OH) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= OW * OH) return;
  const int i = idx / OH, j = idx % OH;
  const int ri = (int)((float)i / (float)OW * (float)W);
  const int rj = (int)((float)j / (float)OH * (float)H);
  const float* s = rgb + ((size_t)ri * H + rj) * 3;
  float* d = out + (size_t)idx * 3;
  d[0] = s[0];
  d[1] = s[1];
  d[2] = s[2];
}

// ====================================================================== display
// visualize_normals, main.py:1632-1660: a pixel-centre camera ray per render
// pixel, colour = shading normal * 0.5 + 0.5 (black on a miss), then the
// same nearest upsample to the display grid as color_rays.  One thread per
// display pixel (the launcher runs it on the render grid and gathers when
// the display is larger).
__global__ __launch_bounds__(256) void sr_normals_kernel(SceneDev s, int W, int H, int OW, int OH,
                                                         float* __restrict__ out, uint32_t* overflow) {
  __shared__ int lds_stack[SR_STACK_SIZE * 256];
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= OW * OH) return;
  const int i = idx / OH, j = idx % OH;
  const int ri = (int)((float)i / (float)OW * (float)W);
  const int rj = (int)((float)j / (float)OH * (float)H);
  const float u = ((float)ri + 0.5f) / (float)W;
  const float v = ((float)rj + 0.5f) / (float)H;
  const float aspect = (float)W / (float)H;
  const f3 cu = mk(s.cam_u[0], s.cam_u[1], s.cam_u[2]);
  const f3 cv = mk(s.cam_v[0], s.cam_v[1], s.cam_v[2]);
  const f3 cw = mk(s.cam_w[0], s.cam_w[1], s.cam_w[2]);
  const f3 d = normalize(cu * (u * 2.0f - 1.0f) * aspect * 1.0f + cv * (v * 2.0f - 1.0f) * 1.0f - cw * 2.0f);
  const f3 o = mk(s.cam_o[0], s.cam_o[1], s.cam_o[2]);
  const HitD h = trace_closest(s.nodes, s.tris, o, d, lds_stack + threadIdx.x, blockDim.x, overflow);
  f3 c = mk(0.0f, 0.0f, 0.0f);
  if (h.prim >= 0) c = surface_of(s, h, d).sn * 0.5f + mk(0.5f, 0.5f, 0.5f);
  float* q = out + (size_t)idx * 3;
  q[0] = c.x;
  q[1] = c.y;
  q[2] = c.z;
}

// Display image [n][3] f32 in [0, 1] -> [n][4] RGBA8 (clamped, rounded to
// nearest, alpha 255).
__global__ void sr_rgba8_kernel(const float* __restrict__ rgb, size_t n, uchar4* __restrict__ out) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  unsigned char c[3];
  for (int r = 0; r < 3; ++r) {
    float x = rgb[k * 3 + r];
    x = x != x ? 0.0f : fminf(fmaxf(x, 0.0f), 1.0f);  // NaN -> 0
    c[r] = (unsigned char)(x * 255.0f + 0.5f);
  }
  out[k] = make_uchar4(c[0], c[1], c[2], 255);
}

// ====================================================================== launchers
int sr_launch_normals(const SceneDev& s, int W, int H, int OW, int OH, float* out, float* tmp, uint32_t* overflow,
                      hipStream_t st) {
  const long long n = (long long)OW * OH;
  if (n <= 0) return 0;
  if ((OW != W || OH != H) && tmp) {
    // upsampled display: one ray per render pixel (pixels_geometry,
    // main.py:1633-1653), then the nearest gather (main.py:1655-1658)
    const long long nr = (long long)W * H;
    hipLaunchKernelGGL(sr_normals_kernel, dim3((unsigned)((nr + 255) / 256)), dim3(256), 0, st, s, W, H, W, H, tmp,
                       overflow);
    hipLaunchKernelGGL(sr_upsample_rgb_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, tmp, W, H, out,
                       OW, OH);
  } else {
    hipLaunchKernelGGL(sr_normals_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, s, W, H, OW, OH, out,
                       overflow);
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int sr_launch_rgba8(const float* rgb, size_t n, void* out, hipStream_t st) {
  if (n == 0) return 0;
  hipLaunchKernelGGL(sr_rgba8_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, rgb, n, (uchar4*)out);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int sr_launch_trace(const SceneDev& s, int n, const float* o, const float* d, float* t, int* prim, float* uv,
                    uint32_t* overflow, hipStream_t st) {
  if (n <= 0) return 0;
  const int bs = 256;
  hipLaunchKernelGGL(sr_trace_kernel, dim3((n + bs - 1) / bs), dim3(bs), 0, st, s, n, o, d, t, prim, uv, overflow);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int sr_launch_finalize(float* mean, float* sum, size_t n, int first, int nb, hipStream_t st) {
  if (n == 0) return 0;
  size_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(sr_finalize_kernel, dim3((unsigned)blocks), dim3(256), 0, st, mean, sum, n, (float)first,
                     (float)(first + nb));
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int sr_launch_fill(float* p, size_t n, float v, hipStream_t st) {
  if (n == 0) return 0;
  size_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(sr_fill_kernel, dim3((unsigned)blocks), dim3(256), 0, st, p, n, v);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int sr_launch_resolve(const float* acc, int W, int H, int B, const ResolveParams& p, float* out, int OW, int OH,
                      float* tmp, hipStream_t st) {
  int n = OW * OH;
  if (n <= 0) return 0;
  const bool aligned = ((uintptr_t)acc & 15u) == 0;
  const int nr = W * H;
  if (OW == W && OH == H && aligned) {
    hipLaunchKernelGGL(sr_resolve_tile_kernel, dim3((n + 255) / 256), dim3(256), (size_t)256 * B * sizeof(float), st,
                       acc, n, B, p, out);
  } else if (tmp && aligned) {
    // upsampled display: resolve each render pixel once, then gather
    hipLaunchKernelGGL(sr_resolve_tile_kernel, dim3((nr + 255) / 256), dim3(256), (size_t)256 * B * sizeof(float), st,
                       acc, nr, B, p, tmp);
    hipLaunchKernelGGL(sr_upsample_rgb_kernel, dim3((n + 255) / 256), dim3(256), 0, st, tmp, W, H, out, OW, OH);
  } else {
    hipLaunchKernelGGL(sr_resolve_kernel, dim3((n + 255) / 256), dim3(256), 0, st, acc, W, H, B, p, out, OW, OH);
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
