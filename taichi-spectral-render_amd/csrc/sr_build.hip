// sr_build.hip -- scene compile on the device: the records sr_upload_scene
// builds on the host (setup_scene, main.py:509-631), built from triangle
// arrays already in device memory.
//
//   faces   normal and area per face in f32 with the host's operation order
//           (main.py:526, 535-536), shading records in original face order;
//   BVH     the host builder's tree (bvh4_build.cpp): binned SAH with 16
//           bins per axis into a binary tree with leaves of <= SR_LEAF_MAX
//           faces and object-median splits where the depth budget runs
//           short, built breadth-first -- chunked level kernels for large
//           ranges, one wave per range of <= 64 faces -- then collapsed to
//           the 4-wide Node4 layout in the host's preorder (greedy
//           largest-area slot filling, uniform fallback past the depth
//           bound).  Same arithmetic, stable partitions and sorts, exact
//           min/max: the node array and the leaf order of the triangle
//           records equal the host build's byte for byte;
//   lights  per-band light lists in face order with the f64 CDF of area
//           fractions (main.py:590-627, last entry forced to 1), slots in
//           first-appearance order -- the same records as the host build.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

#include "sr_kernels.h"

namespace {

constexpr int kMaxLevels = SR_MAX_BVH4_DEPTH;

__device__ __forceinline__ unsigned ord_f(float f) {  // order-preserving float -> uint
  unsigned u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float unord_f(unsigned u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

// bounds[0..2] centroid min, [3..5] centroid max (ordered uints), [6] max |coord|.
// Grid-stride over the faces with at most kFaceBlocks workgroups: the
// centroid bounds are reduced per wave, then per workgroup in LDS, so the
// 7 global atomics are issued once per workgroup (issued once per wave,
// the shared addresses serialise thousands of them).  Min and max are
// exact, so the result does not depend on the reduction order.
constexpr int kFaceBlocks = 1024;
__global__ __launch_bounds__(256) void k_faces(const float* __restrict__ v, const float* __restrict__ nrm,
                                               const int* __restrict__ mat, int n, int n_mats,
                                               float4* __restrict__ shade, float* __restrict__ area,
                                               float* __restrict__ box, float* __restrict__ cen,
                                               unsigned* __restrict__ bounds, unsigned* __restrict__ bad) {
#pragma clang fp contract(off)
  float cmin[3] = {3.0e38f, 3.0e38f, 3.0e38f}, cmax[3] = {-3.0e38f, -3.0e38f, -3.0e38f}, mabs = 1e-3f;
  for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < n; f += gridDim.x * blockDim.x) {
    const float* p = v + (size_t)f * 9;
    const float* q = nrm + (size_t)f * 9;
    const float e1x = p[3] - p[0], e1y = p[4] - p[1], e1z = p[5] - p[2];
    const float e2x = p[6] - p[0], e2y = p[7] - p[1], e2z = p[8] - p[2];
    float nx = e1y * e2z - e1z * e2y, ny = e1z * e2x - e1x * e2z, nz = e1x * e2y - e1y * e2x;
    const float a = sqrtf(nx * nx + ny * ny + nz * nz) / 2.0f;
    if (a > 1e-6f) {
      nx = nx / (a * 2.0f);
      ny = ny / (a * 2.0f);
      nz = nz / (a * 2.0f);
    }
    area[f] = a;
    const int m = mat[f];
    if (m < 0 || m >= n_mats) atomicAdd(bad, 1u);
    shade[(size_t)f * 4 + 0] = make_float4(nx, ny, nz, a);
    shade[(size_t)f * 4 + 1] = make_float4(q[0], q[1], q[2], __int_as_float(m));
    shade[(size_t)f * 4 + 2] = make_float4(q[3], q[4], q[5], 0.f);
    shade[(size_t)f * 4 + 3] = make_float4(q[6], q[7], q[8], 0.f);
    for (int ax = 0; ax < 3; ++ax) {
      const float lo = fminf(fminf(p[ax], p[3 + ax]), p[6 + ax]);
      const float hi = fmaxf(fmaxf(p[ax], p[3 + ax]), p[6 + ax]);
      box[(size_t)f * 6 + ax] = lo;
      box[(size_t)f * 6 + 3 + ax] = hi;
      const float c = (p[ax] + p[3 + ax] + p[6 + ax]) * (1.0f / 3.0f);
      cen[(size_t)f * 3 + ax] = c;
      cmin[ax] = fminf(cmin[ax], c);
      cmax[ax] = fmaxf(cmax[ax], c);
      mabs = fmaxf(mabs, fmaxf(fabsf(lo), fabsf(hi)));
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    for (int ax = 0; ax < 3; ++ax) {
      cmin[ax] = fminf(cmin[ax], __shfl_xor(cmin[ax], off));
      cmax[ax] = fmaxf(cmax[ax], __shfl_xor(cmax[ax], off));
    }
    mabs = fmaxf(mabs, __shfl_xor(mabs, off));
  }
  __shared__ float red[4][7];
  const int w = threadIdx.x >> 6;
  if (__lane_id() == 0) {
    for (int ax = 0; ax < 3; ++ax) {
      red[w][ax] = cmin[ax];
      red[w][3 + ax] = cmax[ax];
    }
    red[w][6] = mabs;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = (int)(blockDim.x + 63) >> 6;
    for (int k = 1; k < nw; ++k) {
      for (int ax = 0; ax < 3; ++ax) {
        cmin[ax] = fminf(cmin[ax], red[k][ax]);
        cmax[ax] = fmaxf(cmax[ax], red[k][3 + ax]);
      }
      mabs = fmaxf(mabs, red[k][6]);
    }
    for (int ax = 0; ax < 3; ++ax) {
      atomicMin(bounds + ax, ord_f(cmin[ax]));
      atomicMax(bounds + 3 + ax, ord_f(cmax[ax]));
    }
    atomicMax(bounds + 6, ord_f(mabs));
  }
}

// ------------------------------------------------------------------ binary SAH tree
// The host builder's recursion (bvh4_build.cpp, Build2::build) run level by
// level: every task is a contiguous range [lo, hi) of the face permutation
// `idx` and a binary node id.  Tasks of more than kSmall faces are split by
// the chunked level kernels (bounds -> bins -> split -> stable partition or
// segmented sort for the median); a task of <= kSmall faces builds its whole
// subtree in one wave (k_small).  The arithmetic is the host's, operation for
// operation (contraction off), and min/max are exact, so each split is the
// host's split.

constexpr int kBins = 16;
constexpr int kSmall = 64;    // faces per wave-built subtree
constexpr int kChunk = 2048;  // faces per workgroup in the level kernels (256 threads x 8)
constexpr int kDepthLimit = 2 * SR_MAX_BVH4_DEPTH;  // binary depth budget, as the host builder
constexpr int kBinWords = 7;  // count, lo xyz, hi xyz (ordered uints)

struct BNode {  // binary node: leaf when count > 0 (faces idx[first, first + count))
  float lo[3], hi[3];
  int left, right, first, count;
};
struct Task {
  int node, lo, hi, depth;
};
struct Split {  // kind 0: SAH, faces with bin <= bin go left; kind 1: median on axis
  int kind, axis, bin, mid;
  float clo, scale;
};

__device__ __forceinline__ int ceil_log2_dev(long long n) {
  int l = 0;
  long long v = 1;
  while (v < n) { v <<= 1; ++l; }
  return l;
}
// the host builder forces object-median splits when the subtree would run
// out of binary depth budget
__device__ __forceinline__ bool force_median(int n, int depth) {
  const int need = ceil_log2_dev(((long long)n + SR_LEAF_MAX - 1) / SR_LEAF_MAX);
  return depth + need + 1 >= kDepthLimit;
}
__device__ __forceinline__ float box_area(const float* lo, const float* hi) {
#pragma clang fp contract(off)
  const float d0 = hi[0] - lo[0], d1 = hi[1] - lo[1], d2 = hi[2] - lo[2];
  if (!(d0 >= 0.f)) return 0.f;
  return 2.0f * (d0 * d1 + d1 * d2 + d2 * d0);
}
__device__ __forceinline__ int bin_of(float c, float clo, float scale) {
#pragma clang fp contract(off)
  const int b = (int)((c - clo) * scale);
  return min(max(b, 0), kBins - 1);
}

struct BestSplit {
  int axis, bin, nleft;
};
// The host's SAH scan: per axis with a positive centroid extent, suffix
// areas/counts, then prefix candidates; the first strictly smaller cost wins.
// bin(a, b, lo, hi) returns the count of bin b on axis a and its box.
template <class BinFn>
__device__ BestSplit sah_choose(const float* ext, BinFn bin) {
#pragma clang fp contract(off)
  float best = INFINITY;
  BestSplit r{-1, -1, 0};
  for (int a = 0; a < 3; ++a) {
    if (!(ext[a] > 0.f)) continue;
    float rA[kBins];
    int rC[kBins];
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    float blo[3], bhi[3];
    int cnt = 0;
    for (int b = kBins - 1; b > 0; --b) {
      const int c = bin(a, b, blo, bhi);
      if (c > 0)
        for (int q = 0; q < 3; ++q) { lo[q] = fminf(lo[q], blo[q]); hi[q] = fmaxf(hi[q], bhi[q]); }
      cnt += c;
      rA[b] = box_area(lo, hi);
      rC[b] = cnt;
    }
    for (int q = 0; q < 3; ++q) { lo[q] = INFINITY; hi[q] = -INFINITY; }
    cnt = 0;
    for (int b = 0; b < kBins - 1; ++b) {
      const int c = bin(a, b, blo, bhi);
      if (c > 0)
        for (int q = 0; q < 3; ++q) { lo[q] = fminf(lo[q], blo[q]); hi[q] = fmaxf(hi[q], bhi[q]); }
      cnt += c;
      if (cnt == 0 || rC[b + 1] == 0) continue;
      const float cost = box_area(lo, hi) * (float)cnt + rA[b + 1] * (float)rC[b + 1];
      if (cost < best) {
        best = cost;
        r = BestSplit{a, b, cnt};
      }
    }
  }
  return r;
}
// median axis: the first axis of largest centroid extent
__device__ __forceinline__ int widest_axis(const float* ext) {
  int a = 0;
  float e = -1.f;
  for (int k = 0; k < 3; ++k)
    if (ext[k] > e) { e = ext[k]; a = k; }
  return a;
}

__global__ void k_iota(unsigned* __restrict__ idx, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) idx[i] = (unsigned)i;
}

__global__ void k_root(Task* __restrict__ big, Task* __restrict__ small, int* __restrict__ ctr, int n) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  ctr[0] = 1;  // node 0 is the root
  const Task t{0, 0, n, 0};
  if (n > kSmall) big[0] = t;
  else { small[0] = t; ctr[1] = 1; }
}

// per-task bounds (box lo, centroid lo as ordered-uint minima; box hi,
// centroid hi as maxima), bins, and chunk counts of this level's tasks
__global__ void k_level_init(const Task* __restrict__ tasks, int T, unsigned* __restrict__ tb,
                             unsigned* __restrict__ bins, int* __restrict__ cc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < T * 12) tb[i] = (i % 12) < 6 ? 0xFFFFFFFFu : 0u;
  if (i < T * 3 * kBins * kBinWords) {
    const int w = i % kBinWords;
    bins[i] = w == 0 ? 0u : (w <= 3 ? 0xFFFFFFFFu : 0u);
  }
  if (i < T) cc[i] = (tasks[i].hi - tasks[i].lo + kChunk - 1) / kChunk;
  if (i == T) cc[i] = 0;
}

// the task whose chunk range holds chunk b (coff: exclusive prefix of chunk counts, each >= 1)
__device__ __forceinline__ int find_task(const int* coff, int T, int b) {
  int lo = 0, hi = T - 1;
  while (lo < hi) {
    const int m = (lo + hi + 1) >> 1;
    if (coff[m] <= b) lo = m;
    else hi = m - 1;
  }
  return lo;
}

struct Chunk {
  int t, s, e;
};
__device__ __forceinline__ bool chunk_of(const int* coff, int T, const Task* tasks, Chunk& c) {
  const int b = blockIdx.x;
  if (b >= coff[T]) return false;
  c.t = find_task(coff, T, b);
  const Task k = tasks[c.t];
  c.s = k.lo + (b - coff[c.t]) * kChunk;
  c.e = min(c.s + kChunk, k.hi);
  return true;
}

__global__ void __launch_bounds__(256) k_bounds(const Task* __restrict__ tasks, int T, const int* __restrict__ coff,
                                                const unsigned* __restrict__ idx, const float* __restrict__ fbox,
                                                const float* __restrict__ cen, unsigned* __restrict__ tb) {
  __shared__ float red[4][12];
  Chunk c;
  if (!chunk_of(coff, T, tasks, c)) return;
  float v[12];
  for (int i = 0; i < 6; ++i) { v[i] = INFINITY; v[6 + i] = -INFINITY; }
  for (int p = c.s + (int)threadIdx.x; p < c.e; p += 256) {
    const unsigned f = idx[p];
    const float* bx = fbox + (size_t)f * 6;
    const float* ce = cen + (size_t)f * 3;
    for (int a = 0; a < 3; ++a) {
      v[a] = fminf(v[a], bx[a]);
      v[3 + a] = fminf(v[3 + a], ce[a]);
      v[6 + a] = fmaxf(v[6 + a], bx[3 + a]);
      v[9 + a] = fmaxf(v[9 + a], ce[a]);
    }
  }
  for (int off = 32; off > 0; off >>= 1)
    for (int i = 0; i < 12; ++i) {
      const float o = __shfl_xor(v[i], off);
      v[i] = i < 6 ? fminf(v[i], o) : fmaxf(v[i], o);
    }
  if ((threadIdx.x & 63) == 0)
    for (int i = 0; i < 12; ++i) red[threadIdx.x >> 6][i] = v[i];
  __syncthreads();
  if (threadIdx.x < 12) {
    const int i = threadIdx.x;
    float r = red[0][i];
    for (int q = 1; q < 4; ++q) r = i < 6 ? fminf(r, red[q][i]) : fmaxf(r, red[q][i]);
    if (i < 6) atomicMin(tb + (size_t)c.t * 12 + i, ord_f(r));
    else atomicMax(tb + (size_t)c.t * 12 + i, ord_f(r));
  }
}

// centroid lo and the binning scale of axis a from the task bounds
__device__ __forceinline__ void task_axes(const unsigned* tb, float* clo, float* ext) {
#pragma clang fp contract(off)
  for (int a = 0; a < 3; ++a) {
    clo[a] = unord_f(tb[3 + a]);
    ext[a] = unord_f(tb[9 + a]) - clo[a];
  }
}

__global__ void __launch_bounds__(256) k_bins(const Task* __restrict__ tasks, int T, const int* __restrict__ coff,
                                              const unsigned* __restrict__ idx, const float* __restrict__ fbox,
                                              const float* __restrict__ cen, const unsigned* __restrict__ tb,
                                              unsigned* __restrict__ bins) {
  __shared__ unsigned sb[3 * kBins * kBinWords];
  Chunk c;
  if (!chunk_of(coff, T, tasks, c)) return;
  const Task k = tasks[c.t];
  if (force_median(k.hi - k.lo, k.depth)) return;
  float clo[3], ext[3], scale[3];
  task_axes(tb + (size_t)c.t * 12, clo, ext);
  for (int a = 0; a < 3; ++a) scale[a] = ext[a] > 0.f ? kBins / ext[a] : 0.f;
  for (int i = threadIdx.x; i < 3 * kBins * kBinWords; i += 256) {
    const int w = i % kBinWords;
    sb[i] = w == 0 ? 0u : (w <= 3 ? 0xFFFFFFFFu : 0u);
  }
  __syncthreads();
  for (int p = c.s + (int)threadIdx.x; p < c.e; p += 256) {
    const unsigned f = idx[p];
    const float* bx = fbox + (size_t)f * 6;
    const float* ce = cen + (size_t)f * 3;
    for (int a = 0; a < 3; ++a) {
      if (!(ext[a] > 0.f)) continue;
      unsigned* q = sb + (a * kBins + bin_of(ce[a], clo[a], scale[a])) * kBinWords;
      atomicAdd(q, 1u);
      for (int j = 0; j < 3; ++j) {
        atomicMin(q + 1 + j, ord_f(bx[j]));
        atomicMax(q + 4 + j, ord_f(bx[3 + j]));
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 3 * kBins; i += 256) {
    const unsigned* q = sb + i * kBinWords;
    if (q[0] == 0u) continue;
    unsigned* g = bins + ((size_t)c.t * 3 * kBins + i) * kBinWords;
    atomicAdd(g, q[0]);
    for (int j = 0; j < 3; ++j) {
      atomicMin(g + 1 + j, q[1 + j]);
      atomicMax(g + 4 + j, q[4 + j]);
    }
  }
}

// ctr: [0] binary nodes, [1] small tasks, [2 + 2L] tasks of level L + 1, [3 + 2L] median tasks of level L
__device__ __forceinline__ void emit_task(const Task& t, Task* next, Task* small, int* ctr, int L) {
  if (t.hi - t.lo > kSmall) next[atomicAdd(ctr + 2 + 2 * L, 1)] = t;
  else small[atomicAdd(ctr + 1, 1)] = t;
}

__global__ void k_split(const Task* __restrict__ tasks, int T, const unsigned* __restrict__ tb,
                        const unsigned* __restrict__ bins, BNode* __restrict__ bn, Split* __restrict__ sp,
                        Task* __restrict__ next, Task* __restrict__ small, int* __restrict__ ctr, int L,
                        int* __restrict__ mbeg, int* __restrict__ mend) {
#pragma clang fp contract(off)
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T) return;
  const Task k = tasks[t];
  const int n = k.hi - k.lo;
  const unsigned* b4 = tb + (size_t)t * 12;
  float clo[3], ext[3];
  task_axes(b4, clo, ext);
  BestSplit bs{-1, -1, 0};
  if (!force_median(n, k.depth)) {
    const unsigned* g = bins + (size_t)t * 3 * kBins * kBinWords;
    bs = sah_choose(ext, [&](int a, int b, float* lo, float* hi) {
      const unsigned* q = g + (a * kBins + b) * kBinWords;
      if (q[0] == 0u) return 0;
      for (int j = 0; j < 3; ++j) { lo[j] = unord_f(q[1 + j]); hi[j] = unord_f(q[4 + j]); }
      return (int)q[0];
    });
  }
  Split s;
  if (bs.axis >= 0) {
    s.kind = 0;
    s.axis = bs.axis;
    s.bin = bs.bin;
    s.mid = k.lo + bs.nleft;
    s.clo = clo[bs.axis];
    s.scale = kBins / ext[bs.axis];
  } else {
    s.kind = 1;
    s.axis = widest_axis(ext);
    s.bin = -1;
    s.mid = k.lo + n / 2;
    s.clo = 0.f;
    s.scale = 0.f;
    const int m = atomicAdd(ctr + 3 + 2 * L, 1);
    mbeg[m] = k.lo;
    mend[m] = k.hi;
  }
  sp[t] = s;
  const int c0 = atomicAdd(ctr, 2);
  BNode nd;
  for (int j = 0; j < 3; ++j) { nd.lo[j] = unord_f(b4[j]); nd.hi[j] = unord_f(b4[6 + j]); }
  nd.left = c0;
  nd.right = c0 + 1;
  nd.first = k.lo;
  nd.count = 0;
  bn[k.node] = nd;
  emit_task(Task{c0, k.lo, s.mid, k.depth + 1}, next, small, ctr, L);
  emit_task(Task{c0 + 1, s.mid, k.hi, k.depth + 1}, next, small, ctr, L);
}

// median tasks: 64-bit keys (ordered centroid on the axis, -0 taken as +0,
// then the position in the range), so the radix sort reproduces the host's
// stable sort
__global__ void __launch_bounds__(256) k_mkeys(const Task* __restrict__ tasks, int T, const int* __restrict__ coff,
                                               const Split* __restrict__ sp, const unsigned* __restrict__ idx,
                                               const float* __restrict__ cen,
                                               unsigned long long* __restrict__ keys) {
  Chunk c;
  if (!chunk_of(coff, T, tasks, c)) return;
  const Split s = sp[c.t];
  if (s.kind != 1) return;
  const int lo = tasks[c.t].lo;
  for (int p = c.s + (int)threadIdx.x; p < c.e; p += 256) {
    const float x = cen[(size_t)idx[p] * 3 + s.axis] + 0.0f;
    keys[p] = ((unsigned long long)ord_f(x) << 32) | (unsigned)(p - lo);
  }
}

// stable partition of the SAH tasks: faces left of the split keep their
// order in [lo, mid), the others in [mid, hi)
constexpr int kPer = kChunk / 256;
__device__ __forceinline__ bool goes_left(const Split& s, const float* cen, unsigned f) {
  return bin_of(cen[(size_t)f * 3 + s.axis], s.clo, s.scale) <= s.bin;
}

__global__ void __launch_bounds__(256) k_pcount(const Task* __restrict__ tasks, int T, const int* __restrict__ coff,
                                                const Split* __restrict__ sp, const unsigned* __restrict__ idx,
                                                const float* __restrict__ cen, int* __restrict__ cl) {
  __shared__ int red[4];
  Chunk c;
  if (!chunk_of(coff, T, tasks, c)) {
    if (threadIdx.x == 0) cl[blockIdx.x] = 0;
    return;
  }
  const Split s = sp[c.t];
  if (s.kind != 0) {
    if (threadIdx.x == 0) cl[blockIdx.x] = 0;
    return;
  }
  int n = 0;
  for (int p = c.s + (int)threadIdx.x; p < c.e; p += 256) n += goes_left(s, cen, idx[p]) ? 1 : 0;
  for (int off = 32; off > 0; off >>= 1) n += __shfl_xor(n, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = n;
  __syncthreads();
  if (threadIdx.x == 0) cl[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ void __launch_bounds__(256) k_pscatter(const Task* __restrict__ tasks, int T, const int* __restrict__ coff,
                                                  const Split* __restrict__ sp, const unsigned* __restrict__ idx,
                                                  const float* __restrict__ cen, const int* __restrict__ cpre,
                                                  unsigned* __restrict__ out, int* __restrict__ err) {
  __shared__ int part[256];
  Chunk c;
  if (!chunk_of(coff, T, tasks, c)) return;
  const Split s = sp[c.t];
  if (s.kind != 0) return;
  const int lo = tasks[c.t].lo, hi = tasks[c.t].hi;
  const int left_before = cpre[blockIdx.x] - cpre[coff[c.t]];  // left faces of earlier chunks of this task
  const int right_before = (c.s - lo) - left_before;
  // thread i owns chunk elements [i * kPer, i * kPer + kPer)
  const int q0 = (int)threadIdx.x * kPer;
  unsigned f[kPer];
  bool fl[kPer];
  int mine = 0;
  for (int j = 0; j < kPer; ++j) {
    const int p = c.s + q0 + j;
    fl[j] = false;
    f[j] = 0u;
    if (p < c.e) {
      f[j] = idx[p];
      fl[j] = goes_left(s, cen, f[j]);
      mine += fl[j] ? 1 : 0;
    }
  }
  part[threadIdx.x] = mine;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {  // inclusive scan of the per-thread counts
    const int v = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  int L = part[threadIdx.x] - mine;  // left elements of this chunk before mine
  for (int j = 0; j < kPer; ++j) {
    const int q = q0 + j;
    if (c.s + q >= c.e) break;
    // the bins counted the same left faces, so every target lies in its half
    const int dst = fl[j] ? lo + left_before + L++ : s.mid + right_before + (q - L);
    if (dst < (fl[j] ? s.mid : hi) && dst >= (fl[j] ? lo : s.mid)) out[dst] = f[j];
    else atomicAdd(err, 1);
  }
}

__global__ void __launch_bounds__(256) k_copyback(const Task* __restrict__ tasks, int T, const int* __restrict__ coff,
                                                  const unsigned* __restrict__ src, unsigned* __restrict__ dst) {
  Chunk c;
  if (!chunk_of(coff, T, tasks, c)) return;
  for (int p = c.s + (int)threadIdx.x; p < c.e; p += 256) dst[p] = src[p];
}

// One wave builds the subtree of a task of <= kSmall faces: lane i holds
// the face at position i of the task's range; bins are lane-parallel (lane
// a * 16 + b owns bin b of axis a), the partition is a ballot prefix and the
// median a rank count with ties by position, all stable like the host.
__global__ void __launch_bounds__(64) k_small(const Task* __restrict__ tasks, int S, unsigned* __restrict__ idx,
                                              const float* __restrict__ fbox, const float* __restrict__ cen,
                                              BNode* __restrict__ bn, int* __restrict__ ctr, int* __restrict__ err) {
#pragma clang fp contract(off)
  __shared__ unsigned s_id[kSmall];
  __shared__ float s_box[kSmall][6];
  __shared__ float s_cen[kSmall][3];
  __shared__ int s_cnt[3 * kBins];
  __shared__ float s_blo[3 * kBins][3], s_bhi[3 * kBins][3];
  __shared__ int4 s_stack[kDepthLimit + 2];
  if ((int)blockIdx.x >= S) return;
  const Task k = tasks[blockIdx.x];
  const int m = k.hi - k.lo;
  const int lane = threadIdx.x;
  if (lane < m) {
    const unsigned f = idx[k.lo + lane];
    s_id[lane] = f;
    for (int j = 0; j < 6; ++j) s_box[lane][j] = fbox[(size_t)f * 6 + j];
    for (int j = 0; j < 3; ++j) s_cen[lane][j] = cen[(size_t)f * 3 + j];
  }
  __syncthreads();
  int node = k.node, a = 0, b = m, depth = k.depth, sp = 0;
  while (true) {
    const bool in = lane >= a && lane < b;
    unsigned f = 0u;
    float bx[6], ce[3];
    for (int j = 0; j < 6; ++j) bx[j] = j < 3 ? INFINITY : -INFINITY;
    for (int j = 0; j < 3; ++j) ce[j] = 0.f;
    float v[12];
    for (int j = 0; j < 6; ++j) { v[j] = INFINITY; v[6 + j] = -INFINITY; }
    if (in) {
      f = s_id[lane];
      for (int j = 0; j < 6; ++j) bx[j] = s_box[lane][j];
      for (int j = 0; j < 3; ++j) ce[j] = s_cen[lane][j];
      for (int j = 0; j < 3; ++j) {
        v[j] = bx[j];
        v[3 + j] = ce[j];
        v[6 + j] = bx[3 + j];
        v[9 + j] = ce[j];
      }
    }
    for (int off = 32; off > 0; off >>= 1)
      for (int i = 0; i < 12; ++i) {
        const float o = __shfl_xor(v[i], off);
        v[i] = i < 6 ? fminf(v[i], o) : fmaxf(v[i], o);
      }
    const int cnt = b - a;
    BNode nd;
    for (int j = 0; j < 3; ++j) { nd.lo[j] = v[j]; nd.hi[j] = v[6 + j]; }
    nd.first = k.lo + a;
    if (cnt <= SR_LEAF_MAX) {
      nd.left = nd.right = -1;
      nd.count = cnt;
      if (lane == 0) bn[node] = nd;
    } else {
      float clo[3], ext[3];
      for (int j = 0; j < 3; ++j) {
        clo[j] = v[3 + j];
        ext[j] = v[9 + j] - v[3 + j];
      }
      BestSplit bs{-1, -1, 0};
      if (!force_median(cnt, depth)) {
        if (lane < 3 * kBins) {
          const int ax = lane / kBins, bb = lane % kBins;
          int c = 0;
          float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
          if (ext[ax] > 0.f) {
            const float scale = kBins / ext[ax];
            for (int j = a; j < b; ++j) {
              if (bin_of(s_cen[j][ax], clo[ax], scale) != bb) continue;
              ++c;
              for (int q = 0; q < 3; ++q) {
                lo[q] = fminf(lo[q], s_box[j][q]);
                hi[q] = fmaxf(hi[q], s_box[j][3 + q]);
              }
            }
          }
          s_cnt[lane] = c;
          for (int q = 0; q < 3; ++q) { s_blo[lane][q] = lo[q]; s_bhi[lane][q] = hi[q]; }
        }
        __syncthreads();
        bs = sah_choose(ext, [&](int ax, int bb, float* lo, float* hi) {
          const int i = ax * kBins + bb;
          for (int q = 0; q < 3; ++q) { lo[q] = s_blo[i][q]; hi[q] = s_bhi[i][q]; }
          return s_cnt[i];
        });
      }
      int mid, pos;
      if (bs.axis >= 0) {
        const bool fl = in && bin_of(ce[bs.axis], clo[bs.axis], kBins / ext[bs.axis]) <= bs.bin;
        const unsigned long long bal = __ballot(fl);
        const int lr = __popcll(bal & ((1ull << lane) - 1ull));
        mid = a + bs.nleft;
        pos = fl ? a + lr : mid + (lane - a - lr);
      } else {
        const int ax = widest_axis(ext);
        int r = 0;
        if (in)
          for (int j = a; j < b; ++j) {
            const float cj = s_cen[j][ax];
            r += (cj < ce[ax] || (cj == ce[ax] && j < lane)) ? 1 : 0;
          }
        mid = a + cnt / 2;
        pos = a + r;
      }
      __syncthreads();  // every lane has read the old order
      if (in) {
        s_id[pos] = f;
        for (int j = 0; j < 6; ++j) s_box[pos][j] = bx[j];
        for (int j = 0; j < 3; ++j) s_cen[pos][j] = ce[j];
      }
      int c0 = 0;
      if (lane == 0) c0 = atomicAdd(ctr, 2);
      c0 = __shfl(c0, 0);
      nd.left = c0;
      nd.right = c0 + 1;
      nd.count = 0;
      if (lane == 0) bn[node] = nd;
      if (sp >= kDepthLimit + 2) {  // cannot happen: the binary depth is bounded by kDepthLimit
        if (lane == 0) atomicAdd(err, 1);
        break;
      }
      s_stack[sp++] = make_int4(c0 + 1, mid, b, depth + 1);  // right half later
      __syncthreads();
      node = c0;
      b = mid;
      ++depth;
      continue;
    }
    if (sp == 0) break;
    const int4 t = s_stack[--sp];
    node = t.x;
    a = t.y;
    b = t.z;
    depth = t.w;
  }
  __syncthreads();
  if (lane < m) idx[k.lo + lane] = s_id[lane];
}

// ------------------------------------------------------------------ 4-wide collapse
// The host's Collapse::emit, breadth-first: a record per 4-wide node holds
// its binary slots (greedy: open the inner entry of largest area until four
// slots; uniform: children or grandchildren) and the record indices of its
// inner children.  Subtree sizes (nodes, faces) bottom-up, then preorder
// indices and leaf offsets top-down give the host's node array and face
// order: a node, then each inner child's subtree in slot order; a node's
// leaf faces before its descendants'.
struct C4 {
  int bin;       // binary node
  int ns;        // slots used
  int slot[4];   // binary nodes in the slots
  int child[4];  // record of an inner slot, -1 for a leaf
  int nleaf;     // faces in leaf slots
  int size;      // 4-wide nodes in the subtree
  int tsize;     // faces in the subtree
  int pre;       // preorder index = node index
  int toff;      // first face position of the subtree
  int pad;
};

__global__ void k_c4_root(C4* __restrict__ c4, int* __restrict__ lv) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  for (int i = 0; i < 2 * (kMaxLevels + 2); ++i) lv[i] = 0;
  lv[kMaxLevels + 2] = 1;  // lcnt[0]
  c4[0].bin = 0;
  c4[0].pre = 0;
  c4[0].toff = 0;
}

// lv[0 .. kMaxLevels + 1]: first record of level L; lv[kMaxLevels + 2 + L]: records of level L
__global__ void k_c4_level(const BNode* __restrict__ bn, C4* __restrict__ c4, int* __restrict__ lv, int L,
                           int greedy) {
  int* loff = lv;
  int* lcnt = lv + kMaxLevels + 2;
  const int base = loff[L], cnt = lcnt[L];
  const int it = blockIdx.x * blockDim.x + threadIdx.x;
  if (it == 0) loff[L + 1] = base + cnt;
  if (it >= cnt) return;
  C4& c = c4[base + it];
  const BNode& nd = bn[c.bin];
  int slots[4];
  int ns = 0;
  if (nd.count > 0) {  // a leaf root: one node with one leaf slot
    slots[ns++] = c.bin;
  } else if (greedy) {
    slots[0] = nd.left;
    slots[1] = nd.right;
    ns = 2;
    while (ns < 4) {
      int best = -1;
      float ba = -1.f;
      for (int q = 0; q < ns; ++q) {
        const BNode& cq = bn[slots[q]];
        if (cq.count > 0) continue;
        const float ar = box_area(cq.lo, cq.hi);
        if (ar > ba) { ba = ar; best = q; }
      }
      if (best < 0) break;
      const BNode& cb = bn[slots[best]];
      for (int q = ns; q > best + 1; --q) slots[q] = slots[q - 1];
      slots[best] = cb.left;
      slots[best + 1] = cb.right;
      ++ns;
    }
  } else {
    for (int ch : {nd.left, nd.right}) {
      const BNode& cn = bn[ch];
      if (cn.count > 0) slots[ns++] = ch;
      else { slots[ns++] = cn.left; slots[ns++] = cn.right; }
    }
  }
  int nleaf = 0;
  c.ns = ns;
  for (int s = 0; s < 4; ++s) {
    c.slot[s] = s < ns ? slots[s] : -1;
    c.child[s] = -1;
    if (s >= ns) continue;
    const BNode& cn = bn[slots[s]];
    if (cn.count > 0) {
      nleaf += cn.count;
    } else {
      const int r = base + cnt + atomicAdd(lcnt + L + 1, 1);
      c4[r].bin = slots[s];
      c.child[s] = r;
    }
  }
  c.nleaf = nleaf;
}

__global__ void k_c4_size(C4* __restrict__ c4, int base, int cnt) {
  const int it = blockIdx.x * blockDim.x + threadIdx.x;
  if (it >= cnt) return;
  C4& c = c4[base + it];
  int size = 1, tsize = c.nleaf;
  for (int s = 0; s < c.ns; ++s)
    if (c.child[s] >= 0) {
      size += c4[c.child[s]].size;
      tsize += c4[c.child[s]].tsize;
    }
  c.size = size;
  c.tsize = tsize;
}

__global__ void k_c4_pre(C4* __restrict__ c4, int base, int cnt) {
  const int it = blockIdx.x * blockDim.x + threadIdx.x;
  if (it >= cnt) return;
  const C4& c = c4[base + it];
  int pre = c.pre + 1, toff = c.toff + c.nleaf;
  for (int s = 0; s < c.ns; ++s)
    if (c.child[s] >= 0) {
      C4& d = c4[c.child[s]];
      d.pre = pre;
      d.toff = toff;
      pre += d.size;
      toff += d.tsize;
    }
}

__global__ void k_c4_emit(const C4* __restrict__ c4, int total, const BNode* __restrict__ bn,
                          const unsigned* __restrict__ idx, float pad, Node4* __restrict__ nodes,
                          unsigned* __restrict__ order) {
  const int it = blockIdx.x * blockDim.x + threadIdx.x;
  if (it >= total) return;
  const C4 c = c4[it];
  Node4 nd;
  int run = 0;
  for (int s = 0; s < 4; ++s) {
    if (s >= c.ns) {
      nd.lo_x[s] = nd.lo_y[s] = nd.lo_z[s] = 3.0e38f;
      nd.hi_x[s] = nd.hi_y[s] = nd.hi_z[s] = -3.0e38f;
      nd.child[s] = SR_EMPTY;
      continue;
    }
    const BNode b = bn[c.slot[s]];
    nd.lo_x[s] = b.lo[0] - pad;
    nd.lo_y[s] = b.lo[1] - pad;
    nd.lo_z[s] = b.lo[2] - pad;
    nd.hi_x[s] = b.hi[0] + pad;
    nd.hi_y[s] = b.hi[1] + pad;
    nd.hi_z[s] = b.hi[2] + pad;
    if (c.child[s] < 0) {
      const int first = c.toff + run;
      for (int q = 0; q < b.count; ++q) order[first + q] = idx[b.first + q];
      run += b.count;
      nd.child[s] = (int)(SR_LEAF_BIT | ((uint32_t)first << 3) | (uint32_t)b.count);
    } else {
      nd.child[s] = c4[c.child[s]].pre;
    }
  }
  for (int q = 0; q < 4; ++q) nd.pad[q] = 0;
  nodes[c.pre] = nd;
}

__global__ void k_tris(const unsigned* __restrict__ order, const float* __restrict__ v, int n,
                       float4* __restrict__ tris) {
#pragma clang fp contract(off)
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n + SR_LEAF_MAX - 1) return;
  if (s >= n) {  // zero records past the end (a leaf's four records are read at once)
    tris[(size_t)s * 3 + 0] = tris[(size_t)s * 3 + 1] = tris[(size_t)s * 3 + 2] = make_float4(0.f, 0.f, 0.f, 0.f);
    return;
  }
  const int f = (int)order[s];
  const float* p = v + (size_t)f * 9;
  tris[(size_t)s * 3 + 0] = make_float4(p[0], p[1], p[2], __int_as_float(f));
  tris[(size_t)s * 3 + 1] = make_float4(p[3] - p[0], p[4] - p[1], p[5] - p[2], 0.f);
  tris[(size_t)s * 3 + 2] = make_float4(p[6] - p[0], p[7] - p[1], p[8] - p[2], 0.f);
}

__global__ void k_emit_flag(const int* __restrict__ mat, const unsigned char* __restrict__ mat_any, int n,
                            int* __restrict__ flag) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f < n) flag[f] = mat_any[mat[f]] ? 1 : 0;
}

__global__ void k_compact(const int* __restrict__ flag, const int* __restrict__ pos, int n, int* __restrict__ ids) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f < n && flag[f]) ids[pos[f]] = f;
}

// se[m * B + b]: surface_emission of material m at band b.  One thread per
// band counts its lights; one thread assigns slots in first-appearance order
// (band by band, face order within a band), as the host build does.
__global__ void k_band_counts(const int* __restrict__ ids, int ne, const int* __restrict__ mat,
                              const unsigned char* __restrict__ se, int B, int* __restrict__ count) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  int c = 0;
  for (int i = 0; i < ne; ++i) c += se[mat[ids[i]] * B + b] ? 1 : 0;
  count[b] = c;
}

__global__ void k_slots(const int* __restrict__ ids, int ne, const int* __restrict__ mat,
                        const unsigned char* __restrict__ se, int B, int* __restrict__ slot_of,
                        int* __restrict__ slots, int* __restrict__ n_slots) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  int ns = 0;
  for (int i = 0; i < ne; ++i) slot_of[i] = -1;
  for (int b = 0; b < B; ++b)
    for (int i = 0; i < ne; ++i)
      if (se[mat[ids[i]] * B + b] && slot_of[i] < 0) {
        slot_of[i] = ns;
        slots[ns++] = ids[i];
      }
  *n_slots = ns;
}

__global__ void k_lights(const int* __restrict__ ids, int ne, const int* __restrict__ mat,
                         const unsigned char* __restrict__ se, int B, const float* __restrict__ area,
                         const int* __restrict__ slot_of, int max_lights, LightRec* __restrict__ lr,
                         float* __restrict__ larea) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  double tot = 0.0;
  for (int i = 0; i < ne; ++i)
    if (se[mat[ids[i]] * B + b]) tot += (double)area[ids[i]];
  larea[b] = (float)tot;
  double c = 0.0;
  int k = 0;
  for (int i = 0; i < ne; ++i) {
    const int f = ids[i];
    if (!se[mat[f] * B + b]) continue;
    c += (double)area[f] / tot;
    LightRec L;
    L.cdf = (float)c;
    L.prim = f;
    L.slot = slot_of[i];
    L.pad = 0;
    lr[(size_t)b * max_lights + k] = L;
    ++k;
  }
  if (k > 0) lr[(size_t)b * max_lights + k - 1].cdf = 1.0f;
}

__global__ void k_lverts(const int* __restrict__ slots, int ns, const float* __restrict__ v,
                         float4* __restrict__ lv) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns) return;
  const float* p = v + (size_t)slots[s] * 9;
  for (int k = 0; k < 3; ++k) lv[(size_t)s * 3 + k] = make_float4(p[3 * k], p[3 * k + 1], p[3 * k + 2], 0.f);
}

__global__ void k_slot_mat(const int* __restrict__ slots, int ns, const int* __restrict__ mat, int* __restrict__ out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < ns) out[s] = mat[slots[s]];
}

struct Tmp {  // build-lifetime device allocation
  void* p = nullptr;
  ~Tmp() { reset(); }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
  }
  template <class T>
  T* get(size_t count) {
    if (hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) p = nullptr;
    return (T*)p;
  }
};

int blocks(long long n, int bs = 256) { return (int)std::max<long long>(1, (n + bs - 1) / bs); }

}  // namespace

#define BCHK(x)                                                              \
  do {                                                                       \
    hipError_t e_ = (x);                                                     \
    if (e_ != hipSuccess) {                                                  \
      std::snprintf(err, errlen, "%s: %s", #x, hipGetErrorString(e_));       \
      return SR_EHIP;                                                        \
    }                                                                        \
  } while (0)
#define BALLOC(ptr)                                                          \
  do {                                                                       \
    if (!(ptr)) {                                                            \
      std::snprintf(err, errlen, "device allocation failed (%s)", #ptr);    \
      return SR_ENOMEM;                                                      \
    }                                                                        \
  } while (0)

void sr_free_scene_out(DeviceSceneOut* o) {
  void** ps[] = {&o->nodes, &o->tris, &o->shade, &o->lights, &o->lverts, &o->light_count, &o->light_area};
  for (void** q : ps) {
    if (*q) (void)hipFree(*q);
    *q = nullptr;
  }
}

static void* take(size_t bytes) {
  void* p = nullptr;
  if (hipMalloc(&p, std::max<size_t>(bytes, 16)) != hipSuccess) return nullptr;
  return p;
}

int sr_build_scene_device(const float* tris_v, const float* tris_n, const int* tri_mat, int n, const int* mat_i32,
                          int n_mats, int B, hipStream_t st, DeviceSceneOut* out, char* err, size_t errlen) {
  *out = DeviceSceneOut{};
  err[0] = 0;
  // surface_emission flags per (material, band), and "emits in some band"
  std::vector<unsigned char> se_h((size_t)n_mats * B), any_h((size_t)n_mats, 0);
  for (int m = 0; m < n_mats; ++m)
    for (int b = 0; b < B; ++b) {
      se_h[(size_t)m * B + b] = mat_i32[((size_t)m * 2 + 1) * B + b] ? 1 : 0;
      any_h[m] |= se_h[(size_t)m * B + b];
    }
  Tmp t_se, t_any;
  unsigned char* se = t_se.get<unsigned char>(se_h.size());
  unsigned char* mat_any = t_any.get<unsigned char>(any_h.size());
  BALLOC(se);
  BALLOC(mat_any);
  BCHK(hipMemcpyAsync(se, se_h.data(), se_h.size(), hipMemcpyHostToDevice, st));
  BCHK(hipMemcpyAsync(mat_any, any_h.data(), any_h.size(), hipMemcpyHostToDevice, st));

  // faces
  out->shade = take((size_t)std::max(n, 1) * 4 * sizeof(float4));
  BALLOC(out->shade);
  out->shade_bytes = (size_t)std::max(n, 1) * 4 * sizeof(float4);
  if (n == 0) BCHK(hipMemsetAsync(out->shade, 0, out->shade_bytes, st));  // the placeholder record, as the host's
  Tmp t_area, t_box, t_cen, t_bounds;
  float* area = t_area.get<float>(n);
  float* fbox = t_box.get<float>((size_t)n * 6);
  float* cen = t_cen.get<float>((size_t)n * 3);
  unsigned* bounds = t_bounds.get<unsigned>(8);
  BALLOC(area); BALLOC(fbox); BALLOC(cen); BALLOC(bounds);
  {
    unsigned init[8] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u, 0u, 0u, 0u};
    BCHK(hipMemcpyAsync(bounds, init, sizeof(init), hipMemcpyHostToDevice, st));
  }
  if (n > 0)
    hipLaunchKernelGGL(k_faces, dim3(std::min(blocks(n), kFaceBlocks)), dim3(256), 0, st, tris_v, tris_n, tri_mat, n, n_mats,
                       (float4*)out->shade, area, fbox, cen, bounds, bounds + 7);
  BCHK(hipGetLastError());
  unsigned hb[8];
  BCHK(hipMemcpyAsync(hb, bounds, sizeof(hb), hipMemcpyDeviceToHost, st));
  BCHK(hipStreamSynchronize(st));
  if (hb[7]) {
    std::snprintf(err, errlen, "%u faces have a material id outside [0, %d)", hb[7], n_mats);
    return SR_EINVAL;
  }
  float maxabs = 1e-3f;
  if (n > 0 && (hb[6] & 0x80000000u)) {  // a positive float in the ordered encoding
    const unsigned bits = hb[6] & 0x7FFFFFFFu;
    std::memcpy(&maxabs, &bits, sizeof(bits));
  }
  const float pad = 4e-6f * maxabs;  // as the host builder

  // binary SAH tree, the host builder's splits
  const int cap_big = n / (kSmall + 1) + 2;              // tasks of > kSmall faces in one level
  const int cap_chunks = n / kChunk + cap_big + 1;        // chunks of one level
  Tmp t_idx, t_idx2, t_mk, t_mk2, t_bn, t_ta, t_tb, t_sm, t_bnd, t_bins, t_spl, t_cc, t_coff, t_cl, t_cpre, t_mb, t_me,
      t_ctr, t_err, t_scan, t_seg;
  unsigned* idx = t_idx.get<unsigned>(n);
  unsigned* idx2 = t_idx2.get<unsigned>(n);
  unsigned long long* mkey = t_mk.get<unsigned long long>(n);
  unsigned long long* mkey2 = t_mk2.get<unsigned long long>(n);
  BNode* bn = t_bn.get<BNode>((size_t)2 * n + 2);
  Task* cur = t_ta.get<Task>(cap_big);
  Task* nxt = t_tb.get<Task>(cap_big);
  Task* small = t_sm.get<Task>(std::max(n, 1));
  unsigned* tbnd = t_bnd.get<unsigned>((size_t)cap_big * 12);
  unsigned* bins = t_bins.get<unsigned>((size_t)cap_big * 3 * kBins * kBinWords);
  Split* spl = t_spl.get<Split>(cap_big);
  int* cc = t_cc.get<int>(cap_big + 1);
  int* coff = t_coff.get<int>(cap_big + 1);
  int* cl = t_cl.get<int>(cap_chunks);
  int* cpre = t_cpre.get<int>(cap_chunks);
  int* mbeg = t_mb.get<int>(cap_big);
  int* mend = t_me.get<int>(cap_big);
  const int n_ctr = 2 + 2 * (kDepthLimit + 2);
  int* ctr = t_ctr.get<int>(n_ctr);
  int* kerr = t_err.get<int>(1);
  BALLOC(idx); BALLOC(idx2); BALLOC(mkey); BALLOC(mkey2); BALLOC(bn); BALLOC(cur); BALLOC(nxt); BALLOC(small);
  BALLOC(tbnd); BALLOC(bins); BALLOC(spl); BALLOC(cc); BALLOC(coff); BALLOC(cl); BALLOC(cpre); BALLOC(mbeg);
  BALLOC(mend); BALLOC(ctr); BALLOC(kerr);
  size_t scan_bytes = 0, scan_b2 = 0;
  BCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, cl, cpre, cap_chunks, st));
  BCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_b2, cc, coff, cap_big + 1, st));
  void* scan_tmp = t_scan.get<char>(std::max(scan_bytes, scan_b2));
  BALLOC(scan_tmp);
  scan_bytes = std::max(scan_bytes, scan_b2);
  size_t seg_bytes = 0;
  void* seg_tmp = nullptr;

  BCHK(hipMemsetAsync(ctr, 0, n_ctr * sizeof(int), st));
  BCHK(hipMemsetAsync(kerr, 0, sizeof(int), st));
  int depth = 0;
  int n_nodes = 0;
  if (n > 0) {
    hipLaunchKernelGGL(k_iota, dim3(blocks(n)), dim3(256), 0, st, idx, n);
    hipLaunchKernelGGL(k_root, dim3(1), dim3(64), 0, st, cur, small, ctr, n);
    BCHK(hipGetLastError());
    int T = n > kSmall ? 1 : 0;
    for (int L = 0; T > 0; ++L) {
      if (L > kDepthLimit) {
        std::snprintf(err, errlen, "binary build exceeded %d levels", kDepthLimit);
        return SR_EINVAL;
      }
      const int C = n / kChunk + T + 1;  // >= the level's chunks
      hipLaunchKernelGGL(k_level_init, dim3(blocks((long long)T * 3 * kBins * kBinWords + 1)), dim3(256), 0, st,
                         cur, T, tbnd, bins, cc);
      BCHK(hipcub::DeviceScan::ExclusiveSum(scan_tmp, scan_bytes, cc, coff, T + 1, st));
      hipLaunchKernelGGL(k_bounds, dim3(C), dim3(256), 0, st, cur, T, coff, idx, fbox, cen, tbnd);
      hipLaunchKernelGGL(k_bins, dim3(C), dim3(256), 0, st, cur, T, coff, idx, fbox, cen, tbnd, bins);
      hipLaunchKernelGGL(k_split, dim3(blocks(T, 64)), dim3(64), 0, st, cur, T, tbnd, bins, bn, spl, nxt, small, ctr,
                         L, mbeg, mend);
      BCHK(hipGetLastError());
      int h2[2] = {0, 0};  // next level's tasks, median tasks of this level
      BCHK(hipMemcpyAsync(h2, ctr + 2 + 2 * L, sizeof(h2), hipMemcpyDeviceToHost, st));
      BCHK(hipStreamSynchronize(st));
      if (h2[0] > cap_big || h2[1] > T) {
        std::snprintf(err, errlen, "binary build task overflow (%d tasks, capacity %d)", h2[0], cap_big);
        return SR_EINVAL;
      }
      if (h2[1] > 0) {  // object-median splits: stable sort of each range by centroid
        hipLaunchKernelGGL(k_mkeys, dim3(C), dim3(256), 0, st, cur, T, coff, spl, idx, cen, mkey);
        size_t need = 0;
        BCHK(hipcub::DeviceSegmentedRadixSort::SortPairs(nullptr, need, mkey, mkey2, idx, idx2, n, h2[1], mbeg, mend,
                                                         0, 64, st));
        if (need > seg_bytes) {
          t_seg.reset();
          seg_tmp = t_seg.get<char>(need);
          BALLOC(seg_tmp);
          seg_bytes = need;
        }
        BCHK(hipcub::DeviceSegmentedRadixSort::SortPairs(seg_tmp, seg_bytes, mkey, mkey2, idx, idx2, n, h2[1], mbeg,
                                                         mend, 0, 64, st));
      }
      hipLaunchKernelGGL(k_pcount, dim3(C), dim3(256), 0, st, cur, T, coff, spl, idx, cen, cl);
      BCHK(hipcub::DeviceScan::ExclusiveSum(scan_tmp, scan_bytes, cl, cpre, C, st));
      hipLaunchKernelGGL(k_pscatter, dim3(C), dim3(256), 0, st, cur, T, coff, spl, idx, cen, cpre, idx2, kerr);
      hipLaunchKernelGGL(k_copyback, dim3(C), dim3(256), 0, st, cur, T, coff, idx2, idx);
      BCHK(hipGetLastError());
      std::swap(cur, nxt);
      T = h2[0];
    }
    int hs[2] = {0, 0};  // binary nodes so far, small tasks
    BCHK(hipMemcpyAsync(hs, ctr, sizeof(hs), hipMemcpyDeviceToHost, st));
    BCHK(hipStreamSynchronize(st));
    if (hs[1] > 0)
      hipLaunchKernelGGL(k_small, dim3(hs[1]), dim3(64), 0, st, small, hs[1], idx, fbox, cen, bn, ctr, kerr);
    BCHK(hipGetLastError());
    int herr = 0;
    BCHK(hipMemcpyAsync(&herr, kerr, sizeof(int), hipMemcpyDeviceToHost, st));
    BCHK(hipStreamSynchronize(st));
    if (herr) {
      std::snprintf(err, errlen, "inconsistent split in the binary build (%d faces)", herr);
      return SR_EINVAL;
    }
  }

  // collapse into Node4 in the host builder's preorder
  Tmp t_c4, t_lv, t_order;
  C4* c4 = t_c4.get<C4>(std::max(n, 1) + 1);
  int* lv = t_lv.get<int>(2 * (kMaxLevels + 2));
  unsigned* order = t_order.get<unsigned>(std::max(n, 1));
  BALLOC(c4); BALLOC(lv); BALLOC(order);
  if (n > 0) {
    std::vector<int> hl(2 * (kMaxLevels + 2), 0);
    for (int greedy = 1; greedy >= 0; --greedy) {
      hipLaunchKernelGGL(k_c4_root, dim3(1), dim3(64), 0, st, c4, lv);
      for (int L = 0; L < kMaxLevels; ++L) {  // records of level L + 1 <= min(4^(L+1), n)
        const long long cap = std::min<long long>((long long)n, 1LL << std::min(2 * L, 40));
        hipLaunchKernelGGL(k_c4_level, dim3(blocks(cap)), dim3(256), 0, st, bn, c4, lv, L, greedy);
      }
      BCHK(hipGetLastError());
      BCHK(hipMemcpyAsync(hl.data(), lv, hl.size() * sizeof(int), hipMemcpyDeviceToHost, st));
      BCHK(hipStreamSynchronize(st));
      if (hl[kMaxLevels + 2 + kMaxLevels] == 0) break;  // no node at depth kMaxLevels
      if (!greedy) {
        std::snprintf(err, errlen, "device BVH deeper than %d 4-wide levels (the traversal stack bound)",
                      kMaxLevels);
        return SR_EOVERFLOW;
      }
    }
    const int* loff = hl.data();
    const int* lcnt = hl.data() + kMaxLevels + 2;
    for (int L = 0; L < kMaxLevels; ++L)
      if (lcnt[L] > 0) depth = L + 1;
    n_nodes = loff[depth - 1] + lcnt[depth - 1];
    for (int L = depth - 1; L >= 0; --L)
      hipLaunchKernelGGL(k_c4_size, dim3(blocks(lcnt[L])), dim3(256), 0, st, c4, loff[L], lcnt[L]);
    for (int L = 0; L + 1 < depth; ++L)
      hipLaunchKernelGGL(k_c4_pre, dim3(blocks(lcnt[L])), dim3(256), 0, st, c4, loff[L], lcnt[L]);
    BCHK(hipGetLastError());
    out->nodes = take((size_t)n_nodes * sizeof(Node4));
    BALLOC(out->nodes);
    hipLaunchKernelGGL(k_c4_emit, dim3(blocks(n_nodes)), dim3(256), 0, st, c4, n_nodes, bn, idx, pad,
                       (Node4*)out->nodes, order);
    BCHK(hipGetLastError());
  } else {  // empty scene: one node with four empty slots
    Node4 t;
    std::memset(&t, 0, sizeof(t));
    for (int s = 0; s < 4; ++s) {
      t.lo_x[s] = t.lo_y[s] = t.lo_z[s] = 3.0e38f;
      t.hi_x[s] = t.hi_y[s] = t.hi_z[s] = -3.0e38f;
      t.child[s] = SR_EMPTY;
    }
    n_nodes = 1;
    out->nodes = take(sizeof(Node4));
    BALLOC(out->nodes);
    BCHK(hipMemcpyAsync(out->nodes, &t, sizeof(t), hipMemcpyHostToDevice, st));
    BCHK(hipStreamSynchronize(st));
  }
  out->nodes_bytes = (size_t)n_nodes * sizeof(Node4);
  out->n_nodes = n_nodes;
  out->depth = depth;

  // triangle records in leaf order, plus SR_LEAF_MAX - 1 zero records
  out->tris_bytes = ((size_t)n + SR_LEAF_MAX - 1) * 3 * sizeof(float4);
  out->tris = take(out->tris_bytes);
  BALLOC(out->tris);
  hipLaunchKernelGGL(k_tris, dim3(blocks(n + SR_LEAF_MAX - 1)), dim3(256), 0, st, order, tris_v, n,
                     (float4*)out->tris);
  BCHK(hipGetLastError());

  // light lists
  Tmp t_flag, t_pos, t_ids, t_lscan, t_bc;
  int* flag = t_flag.get<int>(std::max(n, 1));
  int* pos = t_pos.get<int>(std::max(n, 1));
  int* bcount = t_bc.get<int>(std::max(B, 1));
  BALLOC(flag); BALLOC(pos); BALLOC(bcount);
  int ne = 0;
  if (n > 0) {
    hipLaunchKernelGGL(k_emit_flag, dim3(blocks(n)), dim3(256), 0, st, tri_mat, mat_any, n, flag);
    size_t lscan_bytes = 0;
    BCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, lscan_bytes, flag, pos, n, st));
    void* lscan_tmp = t_lscan.get<char>(lscan_bytes);
    BALLOC(lscan_tmp);
    BCHK(hipcub::DeviceScan::ExclusiveSum(lscan_tmp, lscan_bytes, flag, pos, n, st));
    int lastpos = 0, lastflag = 0;
    BCHK(hipMemcpyAsync(&lastpos, pos + n - 1, sizeof(int), hipMemcpyDeviceToHost, st));
    BCHK(hipMemcpyAsync(&lastflag, flag + n - 1, sizeof(int), hipMemcpyDeviceToHost, st));
    BCHK(hipStreamSynchronize(st));
    ne = lastpos + lastflag;
  }
  int* ids = t_ids.get<int>(std::max(ne, 1));
  BALLOC(ids);
  if (ne > 0) hipLaunchKernelGGL(k_compact, dim3(blocks(n)), dim3(256), 0, st, flag, pos, n, ids);
  hipLaunchKernelGGL(k_band_counts, dim3(blocks(B, 64)), dim3(64), 0, st, ids, ne, tri_mat, se, B, bcount);
  BCHK(hipGetLastError());
  out->h_light_count.assign((size_t)B, 0);
  BCHK(hipMemcpyAsync(out->h_light_count.data(), bcount, (size_t)B * sizeof(int), hipMemcpyDeviceToHost, st));
  BCHK(hipStreamSynchronize(st));
  int max_lights = 0;
  for (int b = 0; b < B; ++b) max_lights = std::max(max_lights, out->h_light_count[b]);
  out->max_lights = max_lights;
  const int ml = std::max(max_lights, 1);
  out->lights_bytes = (size_t)B * ml * sizeof(LightRec);
  out->lights = take(out->lights_bytes);
  out->light_count_bytes = (size_t)std::max(B, 1) * sizeof(int);
  out->light_count = take(out->light_count_bytes);
  out->light_area_bytes = (size_t)std::max(B, 1) * sizeof(float);
  out->light_area = take(out->light_area_bytes);
  BALLOC(out->lights); BALLOC(out->light_count); BALLOC(out->light_area);
  BCHK(hipMemsetAsync(out->lights, 0, out->lights_bytes, st));
  BCHK(hipMemcpyAsync(out->light_count, bcount, (size_t)B * sizeof(int), hipMemcpyDeviceToDevice, st));
  Tmp t_slot_of, t_slots, t_ns;
  int* slot_of = t_slot_of.get<int>(std::max(ne, 1));
  int* slots = t_slots.get<int>(std::max(ne, 1));
  int* nslots = t_ns.get<int>(1);
  BALLOC(slot_of); BALLOC(slots); BALLOC(nslots);
  hipLaunchKernelGGL(k_slots, dim3(1), dim3(64), 0, st, ids, ne, tri_mat, se, B, slot_of, slots, nslots);
  hipLaunchKernelGGL(k_lights, dim3(blocks(B, 64)), dim3(64), 0, st, ids, ne, tri_mat, se, B, area, slot_of, ml,
                     (LightRec*)out->lights, (float*)out->light_area);
  BCHK(hipGetLastError());
  int ns = 0;
  out->h_light_area.assign((size_t)B, 0.0f);
  BCHK(hipMemcpyAsync(&ns, nslots, sizeof(int), hipMemcpyDeviceToHost, st));
  BCHK(hipMemcpyAsync(out->h_light_area.data(), out->light_area, (size_t)B * sizeof(float), hipMemcpyDeviceToHost,
                      st));
  BCHK(hipStreamSynchronize(st));
  out->n_slots = ns;
  out->lverts_bytes = (size_t)std::max(ns, 1) * 3 * sizeof(float4);
  out->lverts = take(out->lverts_bytes);
  BALLOC(out->lverts);
  BCHK(hipMemsetAsync(out->lverts, 0, out->lverts_bytes, st));
  if (ns > 0) hipLaunchKernelGGL(k_lverts, dim3(blocks(ns)), dim3(256), 0, st, slots, ns, tris_v, (float4*)out->lverts);
  BCHK(hipGetLastError());
  // slot faces, their materials and vertices for the light tree (built on the host)
  out->h_slots.assign((size_t)ns, 0);
  out->h_slot_mat.assign((size_t)ns, 0);
  out->h_lverts.assign((size_t)ns * 12, 0.0f);
  if (ns > 0) {
    Tmp t_sm;
    int* smat = t_sm.get<int>(ns);
    BALLOC(smat);
    hipLaunchKernelGGL(k_slot_mat, dim3(blocks(ns)), dim3(256), 0, st, slots, ns, tri_mat, smat);
    BCHK(hipGetLastError());
    BCHK(hipMemcpyAsync(out->h_slots.data(), slots, (size_t)ns * sizeof(int), hipMemcpyDeviceToHost, st));
    BCHK(hipMemcpyAsync(out->h_slot_mat.data(), smat, (size_t)ns * sizeof(int), hipMemcpyDeviceToHost, st));
    BCHK(hipMemcpyAsync(out->h_lverts.data(), out->lverts, (size_t)ns * 12 * sizeof(float), hipMemcpyDeviceToHost,
                        st));
    BCHK(hipStreamSynchronize(st));
  }
  BCHK(hipStreamSynchronize(st));
  return 0;
}
