// sr_build.hip -- scene compile on the device: the records sr_upload_scene
// builds on the host (setup_scene, main.py:509-631), built from triangle
// arrays already in device memory.
//
//   faces   normal and area per face in f32 with the host's operation order
//           (main.py:526, 535-536), shading records in original face order;
//   BVH     linear BVH (Karras, "Maximizing parallelism in the construction
//           of BVHs, octrees and k-d trees", HPG 2012) over the 63-bit Morton
//           codes of the centroids (21 bits per axis; equal codes are ordered
//           by face index -- a stable sort -- and split by sorted position,
//           the paper's augmented keys), bottom-up boxes, then collapsed top-down, one level
//           per launch, into the 4-wide Node4 layout: a node's four slots
//           come from repeatedly opening the frontier entry with the largest
//           box area, subtrees of <= SR_LEAF_MAX triangles become leaves;
//   lights  per-band light lists in face order with the f64 CDF of area
//           fractions (main.py:590-627, last entry forced to 1), slots in
//           first-appearance order -- the same records as the host build.
//
// Leaves are contiguous ranges of the sorted keys, so the sorted order is the
// leaf order of the triangle records.  Closest hits do not depend on the tree
// (boxes are conservative, ties go to the smaller face id), so a scene built
// here traces bit-identically to the host-built one.
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

// bounds[0..2] centroid min, [3..5] centroid max (ordered uints), [6] max |coord|
__global__ void k_faces(const float* __restrict__ v, const float* __restrict__ nrm, const int* __restrict__ mat,
                        int n, int n_mats, float4* __restrict__ shade, float* __restrict__ area,
                        float* __restrict__ box, float* __restrict__ cen, unsigned* __restrict__ bounds,
                        unsigned* __restrict__ bad) {
#pragma clang fp contract(off)
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  float cmin[3] = {3.0e38f, 3.0e38f, 3.0e38f}, cmax[3] = {-3.0e38f, -3.0e38f, -3.0e38f}, mabs = 1e-3f;
  if (f < n) {
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
      cmin[ax] = c;
      cmax[ax] = c;
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
  if (__lane_id() == 0) {
    for (int ax = 0; ax < 3; ++ax) {
      atomicMin(bounds + ax, ord_f(cmin[ax]));
      atomicMax(bounds + 3 + ax, ord_f(cmax[ax]));
    }
    atomicMax(bounds + 6, ord_f(mabs));
  }
}

__device__ __forceinline__ unsigned long long spread21(unsigned x) {  // bit b -> bit 3b
  unsigned long long v = x & 0x1FFFFFu;
  v = (v | (v << 32)) & 0x1F00000000FFFFull;
  v = (v | (v << 16)) & 0x1F0000FF0000FFull;
  v = (v | (v << 8)) & 0x100F00F00F00F00Full;
  v = (v | (v << 4)) & 0x10C30C30C30C30C3ull;
  v = (v | (v << 2)) & 0x1249249249249249ull;
  return v;
}

__global__ void k_keys(const float* __restrict__ cen, int n, const unsigned* __restrict__ bounds,
                       unsigned long long* __restrict__ keys, unsigned* __restrict__ idx) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n) return;
  // one cubic grid over the largest extent: a flat scene spends few code
  // bits on its thin axis instead of splitting along it as often as along
  // the others
  float ext = 0.0f;
  for (int ax = 0; ax < 3; ++ax) ext = fmaxf(ext, unord_f(bounds[3 + ax]) - unord_f(bounds[ax]));
  unsigned q[3];
  for (int ax = 0; ax < 3; ++ax) {
    const float lo = unord_f(bounds[ax]);
    float t = ext > 0.0f ? (cen[(size_t)f * 3 + ax] - lo) / ext : 0.0f;
    t = fminf(fmaxf(t * 2097152.0f, 0.0f), 2097151.0f);
    q[ax] = (unsigned)t;
  }
  keys[f] = (spread21(q[0]) << 2) | (spread21(q[1]) << 1) | spread21(q[2]);
  idx[f] = (unsigned)f;
}

// Common-prefix length of sorted keys i and j; equal keys continue with
// the bits of the sorted positions (64 + clz), so every pair differs.
__device__ __forceinline__ int delta(const unsigned long long* k, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  const unsigned long long x = k[i] ^ k[j];
  return x ? __clzll((long long)x) : 64 + __clz(i ^ j);
}

// Binary node refs: >= 0 internal node, < 0 leaf ~s (sorted position s).
__global__ void k_karras(const unsigned long long* __restrict__ k, int n, int* __restrict__ left,
                         int* __restrict__ right, int* __restrict__ first, int* __restrict__ last,
                         int* __restrict__ parent_in, int* __restrict__ parent_leaf) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  const int d = (delta(k, n, i, i + 1) - delta(k, n, i, i - 1)) >= 0 ? 1 : -1;
  const int dmin = delta(k, n, i, i - d);
  int lmax = 2;
  while (delta(k, n, i, i + lmax * d) > dmin) lmax *= 2;
  int l = 0;
  for (int t = lmax / 2; t >= 1; t /= 2)
    if (delta(k, n, i, i + (l + t) * d) > dmin) l += t;
  const int j = i + l * d;
  const int dnode = delta(k, n, i, j);
  int s = 0;
  int t = l;
  do {
    t = (t + 1) >> 1;
    if (delta(k, n, i, i + (s + t) * d) > dnode) s += t;
  } while (t > 1);
  const int g = i + s * d + min(d, 0);
  const int lo = min(i, j), hi = max(i, j);
  const int L = lo == g ? ~g : g;
  const int R = hi == g + 1 ? ~(g + 1) : g + 1;
  left[i] = L;
  right[i] = R;
  first[i] = lo;
  last[i] = hi;
  if (L >= 0) parent_in[L] = i; else parent_leaf[~L] = i;
  if (R >= 0) parent_in[R] = i; else parent_leaf[~R] = i;
}

// Bottom-up boxes of the internal nodes: the second child to arrive at a
// node forms its box and climbs on.
__global__ void k_boxes(const unsigned* __restrict__ order, int n, const float* __restrict__ fbox,
                        const int* __restrict__ left, const int* __restrict__ right,
                        const int* __restrict__ parent_in, const int* __restrict__ parent_leaf,
                        unsigned* __restrict__ arrivals, float* ibox) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n || n < 2) return;
  int node = parent_leaf[s];
  while (node >= 0) {
    __threadfence();
    if (atomicAdd(arrivals + node, 1u) == 0u) return;  // first arrival: the sibling finishes
    __threadfence();
    float lo[3], hi[3];
    const int ch[2] = {left[node], right[node]};
    for (int ax = 0; ax < 3; ++ax) { lo[ax] = 3.0e38f; hi[ax] = -3.0e38f; }
    for (int c = 0; c < 2; ++c) {
      const float* b = ch[c] >= 0 ? ibox + (size_t)ch[c] * 6 : fbox + (size_t)order[~ch[c]] * 6;
      for (int ax = 0; ax < 3; ++ax) {
        // volatile read: the sibling's box was written by another thread
        lo[ax] = fminf(lo[ax], ((volatile const float*)b)[ax]);
        hi[ax] = fmaxf(hi[ax], ((volatile const float*)b)[3 + ax]);
      }
    }
    for (int ax = 0; ax < 3; ++ax) {
      ibox[(size_t)node * 6 + ax] = lo[ax];
      ibox[(size_t)node * 6 + 3 + ax] = hi[ax];
    }
    node = parent_in[node];
  }
}

struct CollapseArgs {
  const unsigned* order;  // sorted position -> face
  const float* fbox;   // per face (original order)
  const float* ibox;   // per internal node
  const int* left;
  const int* right;
  const int* first;
  const int* last;
  Node4* nodes;
  int* node_count;
  const int2* items_in;  // (binary ref, Node4 index)
  const int* n_in;
  int2* items_out;
  int* n_out;
  float pad;
};

__device__ __forceinline__ int ref_size(const CollapseArgs& a, int r) {
  return r >= 0 ? a.last[r] - a.first[r] + 1 : 1;
}

__global__ void k_collapse(CollapseArgs a) {
  const int it = blockIdx.x * blockDim.x + threadIdx.x;
  if (it >= *a.n_in) return;
  const int2 item = a.items_in[it];
  int fr[4];
  int nf = 0;
  if (item.x >= 0 && a.last[item.x] - a.first[item.x] + 1 > SR_LEAF_MAX) {
    fr[0] = a.left[item.x];
    fr[1] = a.right[item.x];
    nf = 2;
  } else {
    fr[0] = item.x;  // the whole tree is one leaf
    nf = 1;
  }
  while (nf < 4) {  // open the frontier subtree (not a leaf) with the largest box area
    int best = -1;
    float barea = -1.0f;
    for (int s = 0; s < nf; ++s) {
      if (fr[s] < 0 || ref_size(a, fr[s]) <= SR_LEAF_MAX) continue;
      const float* b = a.ibox + (size_t)fr[s] * 6;
      const float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
      const float ar = dx * dy + dy * dz + dz * dx;
      if (ar > barea) { best = s; barea = ar; }
    }
    if (best < 0) break;
    const int r = fr[best];
    for (int s = nf; s > best + 1; --s) fr[s] = fr[s - 1];
    fr[best] = a.left[r];
    fr[best + 1] = a.right[r];
    ++nf;
  }
  Node4 nd;
  for (int s = 0; s < 4; ++s) {
    if (s >= nf) {
      nd.lo_x[s] = nd.lo_y[s] = nd.lo_z[s] = 3.0e38f;
      nd.hi_x[s] = nd.hi_y[s] = nd.hi_z[s] = -3.0e38f;
      nd.child[s] = SR_EMPTY;
      continue;
    }
    const int r = fr[s];
    const float* b = r >= 0 ? a.ibox + (size_t)r * 6 : a.fbox + (size_t)a.order[~r] * 6;
    nd.lo_x[s] = b[0] - a.pad;
    nd.lo_y[s] = b[1] - a.pad;
    nd.lo_z[s] = b[2] - a.pad;
    nd.hi_x[s] = b[3] + a.pad;
    nd.hi_y[s] = b[4] + a.pad;
    nd.hi_z[s] = b[5] + a.pad;
    const int sz = ref_size(a, r);
    if (sz <= SR_LEAF_MAX) {
      const int f0 = r >= 0 ? a.first[r] : ~r;
      nd.child[s] = (int)(SR_LEAF_BIT | ((uint32_t)f0 << 3) | (uint32_t)sz);
    } else {
      const int c = atomicAdd(a.node_count, 1);
      nd.child[s] = c;
      const int o = atomicAdd(a.n_out, 1);
      a.items_out[o] = make_int2(r, c);
    }
  }
  for (int q = 0; q < 4; ++q) nd.pad[q] = 0;
  a.nodes[item.y] = nd;
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

struct Tmp {  // build-lifetime device allocation
  void* p = nullptr;
  ~Tmp() {
    if (p) (void)hipFree(p);
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
    hipLaunchKernelGGL(k_faces, dim3(blocks(n)), dim3(256), 0, st, tris_v, tris_n, tri_mat, n, n_mats,
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

  // sorted keys and the binary tree
  Tmp t_keys, t_keys2, t_idx, t_idx2, t_sort;
  unsigned long long* keys = t_keys.get<unsigned long long>(n);
  unsigned long long* skeys = t_keys2.get<unsigned long long>(n);
  unsigned* idx = t_idx.get<unsigned>(n);
  unsigned* order = t_idx2.get<unsigned>(n);
  BALLOC(keys); BALLOC(skeys); BALLOC(idx); BALLOC(order);
  if (n > 0) hipLaunchKernelGGL(k_keys, dim3(blocks(n)), dim3(256), 0, st, cen, n, bounds, keys, idx);
  size_t sort_bytes = 0;
  BCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, keys, skeys, idx, order, n, 0, 63, st));
  void* sort_tmp = t_sort.get<char>(sort_bytes);
  BALLOC(sort_tmp);
  // radix sort is stable: equal codes stay in face order
  if (n > 0) BCHK(hipcub::DeviceRadixSort::SortPairs(sort_tmp, sort_bytes, keys, skeys, idx, order, n, 0, 63, st));

  const int ni = std::max(n - 1, 1);
  Tmp t_l, t_r, t_f, t_la, t_pi, t_pl, t_arr, t_ib;
  int* left = t_l.get<int>(ni);
  int* right = t_r.get<int>(ni);
  int* first = t_f.get<int>(ni);
  int* last = t_la.get<int>(ni);
  int* parent_in = t_pi.get<int>(ni);
  int* parent_leaf = t_pl.get<int>(std::max(n, 1));
  unsigned* arrivals = t_arr.get<unsigned>(ni);
  float* ibox = t_ib.get<float>((size_t)ni * 6);
  BALLOC(left); BALLOC(right); BALLOC(first); BALLOC(last); BALLOC(parent_in); BALLOC(parent_leaf);
  BALLOC(arrivals); BALLOC(ibox);
  if (n >= 2) {
    BCHK(hipMemsetAsync(parent_in, 0xFF, (size_t)ni * sizeof(int), st));  // root has parent -1
    BCHK(hipMemsetAsync(arrivals, 0, (size_t)ni * sizeof(unsigned), st));
    hipLaunchKernelGGL(k_karras, dim3(blocks(n - 1)), dim3(256), 0, st, skeys, n, left, right, first, last,
                       parent_in, parent_leaf);
    hipLaunchKernelGGL(k_boxes, dim3(blocks(n)), dim3(256), 0, st, order, n, fbox, left, right, parent_in,
                       parent_leaf, arrivals, ibox);
    BCHK(hipGetLastError());
  }

  // collapse into Node4, one level per launch
  // every 4-wide node has >= 2 slots and the leaves partition the faces
  const size_t max_nodes = (size_t)std::max(n, 4);
  Tmp t_nodes, t_cnt, t_ia, t_ib2;
  Node4* nodes = t_nodes.get<Node4>(max_nodes);
  int* counts = t_cnt.get<int>(2 + kMaxLevels);  // [0] node count, [1 + L] items of level L
  int2* items_a = t_ia.get<int2>(max_nodes);
  int2* items_b = t_ib2.get<int2>(max_nodes);
  BALLOC(nodes); BALLOC(counts); BALLOC(items_a); BALLOC(items_b);
  int depth = 0;
  int n_nodes = 0;
  if (n > 0) {
    std::vector<int> ci(2 + kMaxLevels, 0);
    ci[0] = 1;  // the root
    ci[1] = 1;  // one item at level 0
    BCHK(hipMemcpyAsync(counts, ci.data(), ci.size() * sizeof(int), hipMemcpyHostToDevice, st));
    const int2 root = make_int2(n >= 2 ? 0 : ~0, 0);
    BCHK(hipMemcpyAsync(items_a, &root, sizeof(int2), hipMemcpyHostToDevice, st));
    int2* in = items_a;
    int2* outq = items_b;
    for (int L = 0; L < kMaxLevels; ++L) {  // levels 0 .. kMaxLevels-1: depth <= kMaxLevels
      CollapseArgs a;
      a.order = order; a.fbox = fbox; a.ibox = ibox;
      a.left = left; a.right = right; a.first = first; a.last = last;
      a.nodes = nodes; a.node_count = counts;
      a.items_in = in; a.n_in = counts + 1 + L;
      a.items_out = outq; a.n_out = counts + 2 + L;
      a.pad = pad;
      // items at level L <= min(4^L, nodes)
      const long long cap = std::min<long long>((long long)max_nodes, 1LL << std::min(2 * L, 40));
      hipLaunchKernelGGL(k_collapse, dim3(blocks(cap)), dim3(256), 0, st, a);
      std::swap(in, outq);
    }
    BCHK(hipGetLastError());
    BCHK(hipMemcpyAsync(ci.data(), counts, ci.size() * sizeof(int), hipMemcpyDeviceToHost, st));
    BCHK(hipStreamSynchronize(st));
    if (ci[1 + kMaxLevels] > 0) {
      std::snprintf(err, errlen, "device BVH deeper than %d 4-wide levels (the traversal stack bound)",
                    kMaxLevels);
      return SR_EOVERFLOW;
    }
    n_nodes = ci[0];
    for (int L = 0; L < kMaxLevels; ++L)
      if (ci[1 + L] > 0) depth = L + 1;
  } else {  // empty scene: one node with four empty slots
    Node4 t;
    std::memset(&t, 0, sizeof(t));
    for (int s = 0; s < 4; ++s) {
      t.lo_x[s] = t.lo_y[s] = t.lo_z[s] = 3.0e38f;
      t.hi_x[s] = t.hi_y[s] = t.hi_z[s] = -3.0e38f;
      t.child[s] = SR_EMPTY;
    }
    BCHK(hipMemcpyAsync(nodes, &t, sizeof(t), hipMemcpyHostToDevice, st));
    n_nodes = 1;
  }
  out->nodes = take((size_t)n_nodes * sizeof(Node4));
  BALLOC(out->nodes);
  out->nodes_bytes = (size_t)n_nodes * sizeof(Node4);
  BCHK(hipMemcpyAsync(out->nodes, nodes, out->nodes_bytes, hipMemcpyDeviceToDevice, st));
  out->n_nodes = n_nodes;
  out->depth = depth;

  // triangle records in leaf (sorted) order, plus SR_LEAF_MAX - 1 zero records
  out->tris_bytes = ((size_t)n + SR_LEAF_MAX - 1) * 3 * sizeof(float4);
  out->tris = take(out->tris_bytes);
  BALLOC(out->tris);
  hipLaunchKernelGGL(k_tris, dim3(blocks(n + SR_LEAF_MAX - 1)), dim3(256), 0, st, order, tris_v, n,
                     (float4*)out->tris);
  BCHK(hipGetLastError());

  // light lists
  Tmp t_flag, t_pos, t_ids, t_scan, t_bc;
  int* flag = t_flag.get<int>(std::max(n, 1));
  int* pos = t_pos.get<int>(std::max(n, 1));
  int* bcount = t_bc.get<int>(std::max(B, 1));
  BALLOC(flag); BALLOC(pos); BALLOC(bcount);
  int ne = 0;
  if (n > 0) {
    hipLaunchKernelGGL(k_emit_flag, dim3(blocks(n)), dim3(256), 0, st, tri_mat, mat_any, n, flag);
    size_t scan_bytes = 0;
    BCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, flag, pos, n, st));
    void* scan_tmp = t_scan.get<char>(scan_bytes);
    BALLOC(scan_tmp);
    BCHK(hipcub::DeviceScan::ExclusiveSum(scan_tmp, scan_bytes, flag, pos, n, st));
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
  BCHK(hipStreamSynchronize(st));
  return 0;
}
