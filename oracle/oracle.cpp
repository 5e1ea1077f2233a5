// oracle.cpp -- CPU restatement of the reference spectral path tracer.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/, the
// smoke() entry and bench.py's cpu_baseline leg load it, never the product.
// It is written from Xyzonox/taichi-spectral-render main.py read as text;
// every function cites the reference lines it restates.  It shares nothing
// with the HIP library except the RNG definition (Philox-4x32-10, restated
// here independently and pinned by Random123's known-answer vectors) and the
// flattened scene arrays it is given.
//
// Parity status per function (how each is pinned; see tests/):
//   philox / uniform            Random123 KAT vectors (test_oracle_units.py)
//   intersect_triangle / trace  brute force + closed-form quad hits
//   fresnel_spectral            normal-incidence closed form, real Fresnel,
//                               numpy complex arithmetic
//   ggx_D / ggx_G / ggx sample  normalisation integral, chi^2 histogram
//   hg eval / sample            normalisation integral, chi^2 histogram
//   cosine hemisphere / sphere  chi^2 histogram, unit-length invariant
//   scatter_volume free flight  exponential histogram, slab transmittance
//   shadow_transmittance        hand-computed slab cases
//   sample_direct_light         parallel-rectangle irradiance closed form
//   cast_rays                   light-visible pixels = closed form,
//                               zero-light scene = 0, escaped paths = 0
//   EEM categorical band pick   row-normalised pmf histogram
//   light-tree NEE (optional)   the build's own extension (main.py:541 is a
//                               TODO): the tree arrays are given, the
//                               descent and pdf walk are restated from their
//                               definition in IEEE f32; pinned on a tree
//                               built in the test (probabilities sum to at
//                               most 1, and to 1 where every light faces the
//                               point; stratified sampler frequencies equal
//                               the pdf; lights facing the point keep p > 0;
//                               tests/test_oracle_lighttree.py) and by the
//                               GPU tree mode's expectation tests
//
// Deliberate, documented deviations shared with the HIP path:
//   * light CDF last entry forced to 1.0 (reference main.py:619-621 leaves
//     it; the fallback to BVH-order tri 0 at main.py:1309 is then dead);
//   * light lists are in original face order, not BVH order (same pmf);
//   * non-finite path contributions are dropped and counted instead of
//     poisoning the accumulator (reference main.py:1589-1590 adds NaNs);
//   * trace ties (equal t) resolve to the lowest original face id.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

static const float PI_F = 3.14159265358979323846f;
static const float EPSILON = 1e-4f;   // main.py:19
static const float ESCAPE = 1e2f;     // main.py:20
static const int SCATTER_BOUNCES = 5; // main.py:23

// ---------------------------------------------------------------- vec3
struct V3 {
  float x, y, z;
};
static inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
static inline V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 operator-(V3 a) { return v3(-a.x, -a.y, -a.z); }
static inline V3 operator*(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
static inline V3 operator*(float s, V3 a) { return v3(a.x * s, a.y * s, a.z * s); }
static inline V3 operator/(V3 a, float s) { return v3(a.x / s, a.y / s, a.z / s); }
static inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline V3 cross(V3 a, V3 b) {
  return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline float length(V3 a) { return std::sqrt(dot(a, a)); }
static inline V3 normalize(V3 a) { return a / length(a); }
// GLSL reflect / refract as taichi.math defines them.
static inline V3 reflect(V3 x, V3 n) { return x - 2.0f * dot(n, x) * n; }
static inline V3 refract(V3 x, V3 n, float eta) {
  float dxn = dot(x, n);
  float k = 1.0f - eta * eta * (1.0f - dxn * dxn);
  V3 r = v3(0.f, 0.f, 0.f);
  if (k >= 0.0f) r = eta * x - (eta * dxn + std::sqrt(k)) * n;
  return r;
}
static inline float norm_sqr(V3 a) { return dot(a, a); }
static inline float tmax(float a, float b) { return std::fmax(a, b); }
static inline float tmin(float a, float b) { return std::fmin(a, b); }
static inline float clampf(float x, float lo, float hi) { return tmin(tmax(x, lo), hi); }

// ---------------------------------------------------------------- RNG
// Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).
static inline void philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Uniforms standing in for ti.random().  A path draws them in fixed blocks:
// block (segment k, slot s) is Philox(ctr = {c0, c1, c2, k << 3 | s},
// key = seed), four words mapped to [0, 1) with 24 bits each, and each
// sampling step of a segment owns one slot (the step's draws are words
// 0, 1, ... of its block in call order).  The reference draws one stream per
// thread (main.py:9, ti.random); any assignment of independent uniforms to
// the draws gives the same estimator, and this one lets the GPU kernel form
// a step's block once for all lanes that reach the step.
// Without a slot the generator is a plain stream: call n returns word n & 3
// of block n >> 2.
enum { SLOT_EXT = 0, SLOT_NEE = 1, SLOT_SCAT = 2, SLOT_OPQ = 3, SLOT_AUX = 4 };
struct Rng {
  uint32_t key[2];
  uint32_t c0, c1, c2;
  uint32_t n = 0;    // plain stream: words drawn; slotted: words drawn from the block
  uint32_t seg = 0;  // path segment of the slotted blocks
  bool slotted = false;
  uint32_t buf[4];
  // Position the draws at word `first` of block (seg, sl).
  void slot(int sl, int first = 0) {
    uint32_t ctr[4] = {c0, c1, c2, (seg << 3) | (uint32_t)sl};
    philox4x32_10(ctr, key, buf);
    slotted = true;
    n = (uint32_t)first;
  }
  float next() {
    uint32_t x;
    if (slotted) {
      if (n >= 4u) { std::fprintf(stderr, "oracle: more than 4 draws in one slot\n"); std::abort(); }
      x = buf[n++];
    } else {
      if ((n & 3u) == 0u) {
        uint32_t ctr[4] = {c0, c1, c2, n >> 2};
        philox4x32_10(ctr, key, buf);
      }
      x = buf[n & 3u];
      ++n;
    }
    return (float)(x >> 8) * (1.0f / 16777216.0f);
  }
};

// ---------------------------------------------------------------- scene
struct Tri {
  V3 v0, v1, v2, n0, n1, n2, normal;
  int mat;
  float area;
};
struct LNode {  // one light-tree node, 64 bytes as the device stores it
  float c[3], r2;  // bounding sphere: centre, squared radius
  float axis[3], cos_o, sin_o;  // normal cone
  float r;         // radius
  int left;   // inner: first child (the second is left + 1); leaf: ~slot
  int right;  // inner: left + 1; leaf: -1
  int parent, face;
  int pad[2];
};
static_assert(sizeof(LNode) == 64, "LNode mirrors the 64-byte device node");
struct Mat {  // MaterialSample, main.py:277-289
  float n, k, r, ss, em, g;
  int tv, se, eem;
};
struct BNode {  // BVHNode, main.py:267-275
  V3 mn, mx;
  int left, right, ntris, start;
};

struct Scene {
  int B = 0;
  float wl_min = 0, wl_max = 0, wl_step = 0;
  std::vector<Tri> tris;       // original face order
  std::vector<int> order;      // reference-BVH leaf order -> original id
  std::vector<BNode> nodes;
  int M = 0, E = 0;
  std::vector<float> mat_f32;  // [M][6][B]
  std::vector<int> mat_i32;    // [M][2][B]
  std::vector<int> mat_eem;    // [M]
  std::vector<float> eem;      // [E][B][B]; id 0 unused
  std::vector<int> nlights;    // [B]
  std::vector<float> light_area;  // [B]
  std::vector<std::vector<float>> light_cdf;  // [B][n]
  std::vector<std::vector<int>> light_idx;    // [B][n]
  V3 cam_o, cam_lookat, cam_up, cam_u, cam_v, cam_w;
  bool brute = false;  // trace by brute force instead of the reference BVH
  // optional light tree (many-light NEE): nodes in sibling pairs (root 0,
  // node 1 unused, the children of an inner node at an even index and the
  // next), power [B][n], leaf of each face (-1: not a light slot)
  bool tree = false;
  std::vector<LNode> lt_nodes;
  std::vector<float> lt_power;
  std::vector<int> lt_leaf;
};

static inline Mat get_mat(const Scene& s, int m, int b) {  // main.py:980-996
  const int B = s.B;
  Mat r;
  const float* f = &s.mat_f32[((size_t)m * 6) * B];
  r.n = f[0 * B + b];
  r.k = f[1 * B + b];
  r.r = f[2 * B + b];
  r.ss = f[3 * B + b];
  r.em = f[4 * B + b];
  r.g = f[5 * B + b];
  const int* i = &s.mat_i32[((size_t)m * 2) * B];
  r.tv = i[0 * B + b];
  r.se = i[1 * B + b];
  r.eem = s.mat_eem[m];
  return r;
}
static inline float eem_at(const Scene& s, int id, int i, int j) {
  return s.eem[((size_t)id * s.B + i) * s.B + j];
}
// (WAVELENGTH_MIN * 1e-9 + idx * WAVELENGTH_STEP * 1e-9), main.py:1057/1203/1462
static inline float lam_m(const Scene& s, int b) {
  return (float)(s.wl_min * 1e-9) + (float)b * s.wl_step * 1e-9f;
}

// ---------------------------------------------------------------- BVH (reference shape)
// Restates build_bvh main.py:437-507: binary SAH with an exhaustive sweep
// over every split of the centroid-sorted list on each axis, leaves of at
// most 10 triangles, a leaf when the best cost >= n, and a 1e-6 pad on
// degenerate axes (get_aabb, main.py:453-461).  Sorting is stable (numpy's
// default argsort is not), so only tie order on equal centroids can differ.
struct Box {
  float mn[3], mx[3];
};
static inline Box box_empty() {
  Box b;
  for (int a = 0; a < 3; ++a) { b.mn[a] = INFINITY; b.mx[a] = -INFINITY; }
  return b;
}
static inline void box_grow(Box& b, const Box& o) {
  for (int a = 0; a < 3; ++a) { b.mn[a] = std::min(b.mn[a], o.mn[a]); b.mx[a] = std::max(b.mx[a], o.mx[a]); }
}
static inline Box box_pad(Box b) {
  for (int a = 0; a < 3; ++a)
    if (b.mn[a] == b.mx[a]) { b.mn[a] -= 1e-6f; b.mx[a] += 1e-6f; }
  return b;
}
static inline float box_sa(const Box& b) {
  float d0 = b.mx[0] - b.mn[0], d1 = b.mx[1] - b.mn[1], d2 = b.mx[2] - b.mn[2];
  return 2.0f * (d0 * d1 + d1 * d2 + d2 * d0);
}

struct Builder {
  const std::vector<Box>& tb;
  const std::vector<float>& cen;  // [N][3]
  Scene& s;
  std::vector<int> pool[3];  // per-axis sorted index lists for the current range
  std::vector<uint8_t> side;
  std::vector<Box> pre;
  Builder(const std::vector<Box>& b, const std::vector<float>& c, Scene& sc) : tb(b), cen(c), s(sc) {}

  Box range_box(const std::vector<int>& idx, size_t lo, size_t hi) {
    Box b = box_empty();
    for (size_t i = lo; i < hi; ++i) box_grow(b, tb[idx[i]]);
    return box_pad(b);
  }
  // A leaf lists its triangles in the order of the caller's `indices`
  // array (main.py:474-477), which is pool[base] over [lo, hi).
  int make_leaf(size_t lo, size_t hi, const Box& bb, int base) {
    int ni = (int)s.nodes.size();
    BNode n;
    n.mn = v3(bb.mn[0], bb.mn[1], bb.mn[2]);
    n.mx = v3(bb.mx[0], bb.mx[1], bb.mx[2]);
    n.left = n.right = -1;
    n.ntris = (int)(hi - lo);
    n.start = (int)s.order.size();
    for (size_t i = lo; i < hi; ++i) s.order.push_back(pool[base][i]);
    s.nodes.push_back(n);
    return ni;
  }

  // indices live in pool[a][lo..hi) for a = 0..2, each sorted on axis a;
  // `base` names which pool holds the caller's `indices` order.
  int subdivide(size_t lo, size_t hi, int base) {
    const size_t n = hi - lo;
    Box bb = range_box(pool[0], lo, hi);
    if (n <= 10) return make_leaf(lo, hi, bb, base);
    int best_axis = -1;
    size_t best_split = 0;
    float best_cost = INFINITY;
    float parent_sa = box_sa(bb);
    if (pre.size() < n + 1) pre.resize(n + 1);
    for (int a = 0; a < 3; ++a) {
      const std::vector<int>& L = pool[a];
      Box acc = box_empty();
      for (size_t i = 0; i < n; ++i) { box_grow(acc, tb[L[lo + i]]); pre[i + 1] = acc; }
      Box suf = box_empty();
      std::vector<float> rcost(n + 1);
      for (size_t i = n; i-- > 1;) {
        box_grow(suf, tb[L[lo + i]]);
        rcost[i] = box_sa(box_pad(suf)) * (float)(n - i);
      }
      for (size_t i = 1; i < n; ++i) {
        float cost = (box_sa(box_pad(pre[i])) * (float)i + rcost[i]) / parent_sa;
        if (cost < best_cost) { best_cost = cost; best_axis = a; best_split = i; }
      }
    }
    if (!(best_cost < (float)n)) return make_leaf(lo, hi, bb, base);
    // stable partition of every axis list by side of the split
    const std::vector<int>& S = pool[best_axis];
    for (size_t i = 0; i < n; ++i) side[S[lo + i]] = (i < best_split) ? 0 : 1;
    std::vector<int> tmp(n);
    for (int a = 0; a < 3; ++a) {
      if (a == best_axis) continue;
      size_t w = 0;
      for (size_t i = 0; i < n; ++i) if (side[pool[a][lo + i]] == 0) tmp[w++] = pool[a][lo + i];
      for (size_t i = 0; i < n; ++i) if (side[pool[a][lo + i]] == 1) tmp[w++] = pool[a][lo + i];
      std::copy(tmp.begin(), tmp.end(), pool[a].begin() + lo);
    }
    int ni = (int)s.nodes.size();
    s.nodes.push_back(BNode());
    int l = subdivide(lo, lo + best_split, best_axis);
    int r = subdivide(lo + best_split, hi, best_axis);
    BNode& nd = s.nodes[ni];
    nd.mn = v3(bb.mn[0], bb.mn[1], bb.mn[2]);
    nd.mx = v3(bb.mx[0], bb.mx[1], bb.mx[2]);
    nd.left = l;
    nd.right = r;
    nd.ntris = 0;
    nd.start = 0;
    return ni;
  }
};

static void build_reference_bvh(Scene& s) {
  const size_t N = s.tris.size();
  s.nodes.clear();
  s.order.clear();
  if (N == 0) return;
  std::vector<Box> tb(N);
  std::vector<float> cen(N * 3);
  for (size_t i = 0; i < N; ++i) {
    const Tri& t = s.tris[i];
    const float* p0 = &t.v0.x; const float* p1 = &t.v1.x; const float* p2 = &t.v2.x;
    for (int a = 0; a < 3; ++a) {
      tb[i].mn[a] = std::min(p0[a], std::min(p1[a], p2[a]));
      tb[i].mx[a] = std::max(p0[a], std::max(p1[a], p2[a]));
      cen[i * 3 + a] = (p0[a] + p1[a] + p2[a]) / 3.0f;
    }
  }
  Builder bl(tb, cen, s);
  for (int a = 0; a < 3; ++a) {
    bl.pool[a].resize(N);
    std::iota(bl.pool[a].begin(), bl.pool[a].end(), 0);
    std::stable_sort(bl.pool[a].begin(), bl.pool[a].end(),
                     [&](int x, int y) { return cen[x * 3 + a] < cen[y * 3 + a]; });
  }
  bl.side.resize(N);
  // The root's `indices` is arange(N) (main.py:445); its leaf order only
  // matters when N <= 10.
  if (N <= 10) {
    Box bb = box_empty();
    for (size_t i = 0; i < N; ++i) box_grow(bb, tb[i]);
    bb = box_pad(bb);
    BNode n;
    n.mn = v3(bb.mn[0], bb.mn[1], bb.mn[2]);
    n.mx = v3(bb.mx[0], bb.mx[1], bb.mx[2]);
    n.left = n.right = -1;
    n.ntris = (int)N;
    n.start = 0;
    for (size_t i = 0; i < N; ++i) s.order.push_back((int)i);
    s.nodes.push_back(n);
    return;
  }
  bl.subdivide(0, N, 0);
}

// ---------------------------------------------------------------- intersection
// intersect_aabb, main.py:797-805 (minnum/maxnum semantics for tm.min/max)
static inline bool intersect_aabb(V3 p, V3 inv, const BNode& n) {
  V3 a = (n.mn - p), b = (n.mx - p);
  V3 tlo = v3(a.x * inv.x, a.y * inv.y, a.z * inv.z);
  V3 thi = v3(b.x * inv.x, b.y * inv.y, b.z * inv.z);
  V3 t0 = v3(tmin(tlo.x, thi.x), tmin(tlo.y, thi.y), tmin(tlo.z, thi.z));
  V3 t1 = v3(tmax(tlo.x, thi.x), tmax(tlo.y, thi.y), tmax(tlo.z, thi.z));
  float te = tmax(t0.x, tmax(t0.y, t0.z));
  float tx = tmin(t1.x, tmin(t1.y, t1.z));
  return te < tx && tx > 0.0f;
}

// intersect_triangle, main.py:807-829 (Moller-Trumbore; returns t,u,v or -1)
static inline V3 intersect_triangle(V3 pos, V3 dir, const Tri& tri) {
  V3 e1 = tri.v1 - tri.v0;
  V3 e2 = tri.v2 - tri.v0;
  V3 pvec = cross(dir, e2);
  float det = dot(e1, pvec);
  float t = -1.0f, u = -1.0f, v = -1.0f;
  if (std::fabs(det) > 1e-8f) {
    float inv_det = 1.0f / det;
    V3 tvec = pos - tri.v0;
    float ut = dot(tvec, pvec) * inv_det;
    if (ut >= 0.0f && ut <= 1.0f) {
      V3 qvec = cross(tvec, e1);
      float vt = dot(dir, qvec) * inv_det;
      if (vt >= 0.0f && ut + vt <= 1.0f) {
        float tt = dot(e2, qvec) * inv_det;
        if (tt > EPSILON) { t = tt; u = ut; v = vt; }
      }
    }
  }
  return v3(t, u, v);
}

struct Hit {  // HitRecord, main.py:318-328
  float t = 1e9f;
  V3 normal{0, 0, 0}, shading_normal{0, 0, 0};
  int material_id = 0;
  float u = 0, v = 0;
  int is_hit = 0, is_front_face = 0, tri_idx = -1;
};

// trace, main.py:831-864.  tri_idx is reported as the original face id.
static Hit trace(const Scene& s, V3 position, V3 direction) {
  Hit hit;
  if (s.brute) {
    for (size_t i = 0; i < s.tris.size(); ++i) {
      V3 hd = intersect_triangle(position, direction, s.tris[i]);
      float t = hd.x;
      if (t > EPSILON && t < hit.t) {
        hit.is_hit = 1; hit.t = t; hit.material_id = s.tris[i].mat;
        hit.tri_idx = (int)i; hit.u = hd.y; hit.v = hd.z;
      }
    }
  } else if (!s.nodes.empty()) {
    V3 inv = v3(1.0f / direction.x, 1.0f / direction.y, 1.0f / direction.z);
    int stack[256];
    int sp = 1;
    stack[0] = 0;
    while (sp > 0) {
      sp -= 1;
      const BNode& node = s.nodes[stack[sp]];
      if (!intersect_aabb(position, inv, node)) continue;
      if (node.ntris > 0) {
        for (int i = 0; i < node.ntris; ++i) {
          int oid = s.order[node.start + i];
          V3 hd = intersect_triangle(position, direction, s.tris[oid]);
          float t = hd.x;
          if (t > EPSILON && t < hit.t) {
            hit.is_hit = 1; hit.t = t; hit.material_id = s.tris[oid].mat;
            hit.tri_idx = oid; hit.u = hd.y; hit.v = hd.z;
          }
        }
      } else {
        stack[sp++] = node.left;
        stack[sp++] = node.right;
      }
    }
  }
  if (hit.is_hit) {
    const Tri& t = s.tris[hit.tri_idx];
    hit.normal = t.normal;
    hit.shading_normal = normalize((1.0f - hit.u - hit.v) * t.n0 + hit.u * t.n1 + hit.v * t.n2);
    hit.is_front_face = dot(direction, hit.normal) < 0.0f ? 1 : 0;
  }
  return hit;
}

// ---------------------------------------------------------------- sampling (main.py:682-792)
static inline V3 random_in_unit_sphere(Rng& rng) {
  float z = rng.next() * 2.0f - 1.0f;
  float r_xy = std::sqrt(tmax(0.0f, 1.0f - z * z));
  float phi = rng.next() * 2.0f * PI_F;
  return v3(r_xy * std::cos(phi), r_xy * std::sin(phi), z);
}
static inline void onb(V3 n, V3& t, V3& b) {
  V3 up = std::fabs(n.y) < 0.999f ? v3(0, 1, 0) : v3(1, 0, 0);
  t = normalize(cross(up, n));
  b = cross(n, t);
}
static inline V3 cosine_weighted_hemisphere_direction(V3 normal, Rng& rng) {
  float r1 = rng.next();
  float r2 = rng.next();
  float phi = 2.0f * PI_F * r1;
  float cos_theta_sq = 1.0f - r2;
  float cos_theta = std::sqrt(cos_theta_sq);
  float sin_theta = std::sqrt(r2);
  V3 t, b;
  onb(normal, t, b);
  return t * std::cos(phi) * sin_theta + b * std::sin(phi) * sin_theta + normal * cos_theta;
}
static inline float evaluate_henyey_greenstein(float cos_theta, float g) {
  float g2 = g * g;
  float denom = 1.0f + g2 - 2.0f * g * cos_theta;
  return (1.0f - g2) / (4.0f * PI_F * denom * std::sqrt(denom) + 1e-6f);
}
static inline V3 sample_henyey_greenstein(V3 w, float g, Rng& rng) {
  float cos_theta;
  if (std::fabs(g) < 1e-4f) {
    cos_theta = 2.0f * rng.next() - 1.0f;
  } else {
    float rand_sq = (1.0f - g * g) / (1.0f - g + 2.0f * g * rng.next());
    cos_theta = (1.0f + g * g - rand_sq * rand_sq) / (2.0f * g);
  }
  float sin_theta = std::sqrt(tmax(0.0f, 1.0f - cos_theta * cos_theta));
  float phi = 2.0f * PI_F * rng.next();
  V3 u, v;
  onb(w, u, v);
  return u * std::cos(phi) * sin_theta + v * std::sin(phi) * sin_theta + w * cos_theta;
}

// ---------------------------------------------------------------- microfacet / Fresnel (main.py:866-978)
struct C2 { float re, im; };
static inline C2 cmul(C2 a, C2 b) { return C2{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
static inline C2 cdiv(C2 a, C2 b) {
  C2 bc{b.re, -b.im};
  float m = b.re * b.re + b.im * b.im;
  C2 p = cmul(a, bc);
  return C2{p.re / m, p.im / m};
}
static inline float cmod2(C2 a) { return a.re * a.re + a.im * a.im; }
static inline float signf(float x) { return x > 0.f ? 1.f : (x < 0.f ? -1.f : 0.f); }
static inline C2 csqrt(C2 z) {
  float mod = std::sqrt(z.re * z.re + z.im * z.im);
  float re = std::sqrt((mod + z.re) / 2.0f);
  float im = signf(z.im) * std::sqrt((mod - z.re) / 2.0f);
  if (z.im == 0.0f) im = 0.0f;
  return C2{re, im};
}
static inline float ggx_D(float NdotH, float roughness) {  // main.py:900-907
  float alpha = roughness * roughness;
  float a2 = alpha * alpha;
  float NdotH2 = NdotH * NdotH;
  float denom = (NdotH2 * (a2 - 1.0f) + 1.0f);
  return a2 / (PI_F * denom * denom);
}
static inline float ggx_G(V3 n, V3 v, V3 l, float roughness) {  // main.py:909-923
  float NdotV = tmax(dot(n, v), 0.0f);
  float alpha = roughness * roughness;
  alpha = tmax(alpha, 1e-4f);
  float alpha_sq = alpha * alpha;
  float k_v = alpha_sq / 2.0f;
  float g1v = NdotV / (NdotV * (1.0f - k_v) + k_v + 1e-7f);
  float NdotL = tmax(dot(n, l), 0.0f);
  float k_l = alpha_sq / 2.0f;
  float g1l = NdotL / (NdotL * (1.0f - k_l) + k_l + 1e-7f);
  return g1v * g1l;
}
static inline float fresnel_spectral(float cos_i, float ni, float ki, float nt, float kt) {  // main.py:925-966
  float c = clampf(cos_i, -1.0f, 1.0f);
  float ang = std::acos(c);
  C2 ein{ni * ni - ki * ki, 2 * ni * ki};
  C2 etr{nt * nt - kt * kt, 2 * nt * kt};
  float cos_t = std::cos(ang);
  float s = std::sin(ang);
  float sin_sq = s * s;
  C2 cosc{cos_t, 0.0f};
  C2 sinc{sin_sq, 0.0f};
  C2 left = cmul(csqrt(ein), cosc);
  C2 em = cmul(ein, sinc);
  C2 right = csqrt(C2{etr.re - em.re, etr.im - em.im});
  C2 rs = cdiv(C2{left.re - right.re, left.im - right.im}, C2{left.re + right.re, left.im + right.im});
  C2 l2 = cmul(etr, left);
  C2 r2 = cmul(ein, right);
  C2 rp = cdiv(C2{l2.re - r2.re, l2.im - r2.im}, C2{l2.re + r2.re, l2.im + r2.im});
  return 0.5f * (cmod2(rs) + cmod2(rp));
}
static inline V3 importance_sample_ggx(float x0, float x1, V3 n, float alpha) {  // main.py:968-978
  float a2 = alpha * alpha;
  float phi = 2.0f * PI_F * x0;
  float cos_theta = std::sqrt((1.0f - x1) / (1.0f + (a2 - 1.0f) * x1));
  float sin_theta = std::sqrt(1.0f - cos_theta * cos_theta);
  V3 h = v3(std::cos(phi) * sin_theta, std::sin(phi) * sin_theta, cos_theta);
  V3 t, b;
  onb(n, t, b);
  return normalize(t * h.x + b * h.y + n * h.z);
}

// ---------------------------------------------------------------- path structures
struct Ray {  // main.py:291-304
  V3 position, direction;
  float throughput = 1.0f;
  int active = 0, transport = 0, path_length = 0, has_terminated = 0, vol_mat_id = 0;
  float surface_pdf = 1.0f;
  int penetrated_tri_idx = 0, is_reflected = 0;
};
struct VolumeEvent {  // main.py:330-344
  int hit_action = 0;
  float travel_distance = 0.f, estimator = 0.f;
  V3 normal{0, 0, 0}, shading_normal{0, 0, 0};
  int is_front_face = 0, hit_mat_id = 0, tri_idx = 0;
  V3 direction{0, 0, 0};
  int new_wavelength_idx = 0, penetrated_tri_idx = 0;
};
struct SurfaceEvent {  // main.py:346-353
  int did_reflect = 0;
  V3 direction{0, 0, 0};
  float estimator = 0.f, pdf = 0.f;
  int new_wavelength = 0;
};
struct NEE {  // main.py:355-359
  float contribution_factor = 0.f, emission = 0.f, weight = 0.f;
};

struct Counters {
  uint64_t segments = 0, traversals = 0, nonfinite = 0;
  // Accounting for the HIP kernel's shadow-walk culling (not part of the
  // reference semantics, which are executed in full here): shadow
  // traversals whose result cannot reach the image -- the walk of a path
  // with throughput 0, and every step that starts in a dense medium
  // (sigma_t * EPSILON >= 14.5, so exp(-sigma_t t) < 1e-6 for any hit) with
  // the light not yet passed -- and the non-finite NEE terms among those
  // walks' contributions (dropped either way).  The kernel does not trace
  // them; the tests compare its counters with these.
  uint64_t culled = 0, culled_nf = 0;
  // Sources of the non-finite values (deviation 3: dropped, the path ends):
  // scatter_surface estimators that are not finite, split into the rough
  // BTDF with n_i == n_t (leaving a rough dielectric: refract returns the
  // incident ray and the estimator is 0/0, main.py:1166-1185) and any other;
  // of the former, those on a path's last segment (never read again); and
  // non-finite NEE and emitter contributions.
  uint64_t nf_btdf_eta1 = 0, nf_btdf_eta1_last = 0, nf_other_est = 0, nf_direct = 0, nf_indirect = 0;
};
static inline bool shadow_dense(float sigma_t) { return sigma_t * EPSILON >= 14.5f; }

// OR_PLANT (0 = none) builds the oracle with one deliberate mistake, so
// that a test can show that its checks fail on it (tests/test_oracle_bsdf.py
// builds the variants with `make plants`).  Never set in the checker itself.
#ifndef OR_PLANT
#define OR_PLANT 0
#endif

// Oren-Nayar factor shared by scatter_surface and evaluate_bsdf
static inline float oren_nayar(float roughness, float NdotL, float NdotV, V3 l, V3 v, V3 n) {
#if OR_PLANT == 6
  float sigma = roughness;  // planted: sigma without pi / 2
#else
  float sigma = roughness * PI_F / 2;
#endif
  float sigma2 = sigma * sigma;
  float A = 1.0f - sigma2 / (2.0f * (sigma2 + 0.33f));
  float B = 0.45f * sigma2 / (sigma2 + 0.09f);
  float theta_i = std::acos(NdotL);
  float theta_o = std::acos(NdotV);
  float alpha = tmax(theta_i, theta_o);
  float beta = tmin(theta_i, theta_o);
  float cos_phi_diff = 0.0f;
  if (NdotL > 1e-4f && NdotV > 1e-4f) {
    V3 pi = normalize(l - n * NdotL);
    V3 po = normalize(v - n * NdotV);
    cos_phi_diff = tmax(0.0f, dot(pi, po));
  }
#if OR_PLANT == 1
  return A + B * cos_phi_diff * std::tan(alpha) * std::sin(beta);  // planted: sin / tan swapped
#else
  return A + B * cos_phi_diff * std::sin(alpha) * std::tan(beta);
#endif
}

// scatter_surface, main.py:1000-1188
static SurfaceEvent scatter_surface(const Scene& s, const Ray& ray, const VolumeEvent& ev, Rng& rng,
                                    V3* h_out = nullptr) {
  SurfaceEvent res;
  const int a = ray.active;
  Mat sm = get_mat(s, ev.hit_mat_id, a);
  Mat vm = get_mat(s, ray.vol_mat_id, a);
  V3 V = -ray.direction;
  V3 N = ev.shading_normal;
  float ni = vm.n, ki = vm.k, nt = sm.n, kt = sm.k;
  rng.slot(SLOT_SCAT);
  float x0 = rng.next();
  float x1 = rng.next();
  V3 H = importance_sample_ggx(x0, x1, N, sm.r);
  if (h_out) *h_out = H;  // (probe: the sampled half vector)
  float ia = tmax(1e-4f, dot(V, H));
  float F = fresnel_spectral(ia, ni, ki, nt, kt);
  float rnd = clampf(rng.next(), 0.0f, 0.99f);
  if (rnd < F) {
    V3 R = reflect(ray.direction, H);
    if (sm.r < 0.008f) {
      res.direction = R; res.did_reflect = 1; res.estimator = F; res.pdf = 1e9f;
    } else {
      float NdotL = tmax(1e-4f, dot(N, R));
      float NdotH = tmax(1e-4f, dot(N, H));
      float NdotV = tmax(1e-4f, dot(N, V));
      float D = ggx_D(NdotH, sm.r);
      float G = ggx_G(N, V, R, sm.r);
      float bsdf = (D * G * F) / (4.0f * NdotL * NdotV);
#if OR_PLANT == 2
      float pdf = (D * NdotH) / ia;  // planted: the 4 of the Jacobian dropped
#else
      float pdf = (D * NdotH) / (4.0f * ia);
#endif
      res.direction = R; res.did_reflect = 1;
      float cos_theta = tmax(0.0f, dot(N, R));
      res.estimator = (bsdf * cos_theta) / pdf;
      res.pdf = pdf;
    }
  } else {
    float absorb = (4.0f * PI_F * kt) / lam_m(s, a);
    float dist_to_event = -std::log(-rng.next() + 1.0f + 1e-7f) / (sm.ss + absorb + 1e-6f);
    if (!sm.tv) {
      res.did_reflect = 1;
      rng.slot(SLOT_OPQ);
      V3 S = cosine_weighted_hemisphere_direction(N, rng);
      float NdotL = tmax(1e-4f, dot(N, S));
      float NdotV = tmax(1e-4f, dot(N, V));
      float albedo = sm.ss / (sm.ss + absorb + 1e-6f);
      float bsdf = (1.0f - F) * (albedo / PI_F) * oren_nayar(sm.r, NdotL, NdotV, S, V, N);
      float pdf = 0.318309886184f * NdotL;
      float qy = 0.0f;
      int eid = sm.eem;
      res.new_wavelength = a;
      if (eid > 0) {
        for (int j = 0; j < s.B; ++j) { float q = eem_at(s, eid, a, j); if (q > qy) qy = q; }
        float r3 = rng.next();
        if (qy > 0.0f && r3 < qy) {
          float normalizer = 0.0f;
          for (int j = 0; j < s.B; ++j) normalizer += eem_at(s, eid, a, j);
          float r4 = rng.next();
          float cum = 0.0f;
          int nw = a;
          for (int j = 0; j < s.B; ++j) {
            cum += eem_at(s, eid, a, j) / normalizer;
            if (r4 < cum) { nw = j; break; }
          }
          res.new_wavelength = nw;
          Mat fm = get_mat(s, ev.hit_mat_id, nw);
          float ab2 = (4.0f * PI_F * fm.k) / lam_m(s, nw);
          bsdf *= std::exp(-ab2 * dist_to_event);
        }
      }
      res.estimator = (bsdf * NdotL) / pdf;
      res.pdf = pdf;
      res.direction = S;
    } else {
      float ior_ratio = ni / nt;
      V3 Rf = refract(ray.direction, H, ior_ratio);
      if (norm_sqr(Rf) < 1e-4f) {
        V3 R = reflect(ray.direction, H);
        if (sm.r < 0.008f) {
          res.direction = R; res.did_reflect = 1; res.estimator = F; res.pdf = 1e9f;
        } else {
          float NdotL = tmax(1e-4f, dot(N, R));
          float NdotH = tmax(1e-4f, dot(N, H));
          float NdotV = tmax(1e-4f, dot(N, V));
          float D = ggx_D(NdotH, sm.r);
          float G = ggx_G(N, V, R, sm.r);
          float bsdf = (D * G * F) / (4.0f * NdotL * NdotV);
          float pdf = (D * NdotH) / (4.0f * ia);
          res.direction = R; res.did_reflect = 1;
          float cos_theta = std::fabs(dot(N, R));
          res.estimator = (bsdf * cos_theta) / pdf;
          res.pdf = pdf;
        }
      } else {
        res.did_reflect = 0;
        res.direction = Rf;
        if (sm.r < 0.0018f) {
          res.estimator = 1.0f - F;
          res.pdf = 1e9f;
        } else {
          float LdotH = dot(Rf, H);
          float VdotH = dot(V, H);
          float NdotH = tmax(1e-4f, dot(N, H));
          float NdotL = tmax(1e-4f, std::fabs(dot(N, Rf)));
          float NdotV = tmax(1e-4f, std::fabs(dot(N, V)));
          float G = ggx_G(N, V, Rf, sm.r);
          float D = ggx_D(NdotH, sm.r);
          float denom = ni * VdotH + nt * LdotH;
          float dsq = denom * denom;
          float t1 = std::fabs(LdotH * VdotH) / (NdotL * NdotV);
          float t2 = (nt * nt * (1.0f - F) * D * G) / dsq;
          float bsdf = t1 * t2;
#if OR_PLANT == 3
          float pdf = (D * NdotH * (1.0f - F) * nt * nt) / dsq;  // planted: |L.H| dropped
#else
          float pdf = (D * NdotH * (1.0f - F) * nt * nt * std::fabs(LdotH)) / dsq;
#endif
          res.estimator = (bsdf * NdotL) / pdf;
          res.pdf = pdf;
        }
      }
    }
  }
  return res;
}

// scatter_volume, main.py:1190-1292
static VolumeEvent scatter_volume(const Scene& s, const Ray& ray, Rng& rng, Counters& cnt) {
  VolumeEvent res;
  const int a = ray.active;
  Mat vm = get_mat(s, ray.vol_mat_id, a);
  res.new_wavelength_idx = a;
  res.direction = ray.direction;
  res.penetrated_tri_idx = ray.penetrated_tri_idx;
  float sa = (4.0f * PI_F * vm.k) / lam_m(s, a);
  rng.slot(SLOT_EXT);
  float d_vol = -std::log(-rng.next() + 1.0f - 1e-7f) / (vm.ss + sa + 1e-6f);
  Hit hit = trace(s, ray.position, ray.direction);
  cnt.traversals++;
  float d_surf = hit.t;
  float travel = -1.0f;
  if (d_surf < ESCAPE) {
    travel = d_surf;
    res.penetrated_tri_idx = hit.tri_idx;
    if (d_vol + EPSILON < travel) travel = d_vol;
  }
  res.travel_distance = travel;
  res.is_front_face = hit.is_front_face;
  float ff = (float)hit.is_front_face;
  res.normal = hit.normal * ff - hit.normal * (1.0f - ff);
  res.shading_normal = hit.shading_normal * ff - hit.shading_normal * (1.0f - ff);
  float T = std::exp(-(sa + vm.ss) * travel);
  float emittance = ((-T + 1.0f) * vm.em) / (sa + 1e-6f);
  float atten = T + emittance;
  if (travel == d_surf) {
    res.estimator = atten / (1.0f + 1e-6f);
    res.hit_action = 1;
    res.hit_mat_id = hit.material_id;
    res.tri_idx = hit.tri_idx;
  } else if (travel >= 0.0f) {
    res.hit_mat_id = -1;
    res.tri_idx = -1;
    float rand_event = rng.next();
    float scatter_type = sa / (vm.ss + sa + 1e-6f);
    if (rand_event <= scatter_type) {
      float qy = 0.0f;
      int eid = vm.eem;
      if (eid > 0) {
        for (int j = 0; j < s.B; ++j) { float q = eem_at(s, eid, a, j); if (q > qy) qy = q; }
        float rf = rng.next();
        if (qy > 0.0f && rf < qy) {
          res.hit_action = 3;
          float normalizer = 0.0f;
          for (int j = 0; j < s.B; ++j) normalizer += eem_at(s, eid, a, j);
          float r2 = rng.next();
          float cum = 0.0f;
          int nw = a;
          for (int j = 0; j < s.B; ++j) {
            cum += eem_at(s, eid, a, j) / normalizer;
            if (r2 < cum) { nw = j; break; }
          }
          res.new_wavelength_idx = nw;
          rng.slot(SLOT_AUX, 2);
          res.direction = random_in_unit_sphere(rng);
          float em2 = ((-(1.0f - T) + 1.0f) * get_mat(s, ray.vol_mat_id, nw).em) / (sa + 1e-6f);
          res.estimator = (1.0f - T) + em2;
        } else {
          res.estimator = atten;
        }
      } else {
        res.estimator = atten / (1.0f + 1e-6f);
      }
    } else {
      res.hit_action = 2;
      res.direction = sample_henyey_greenstein(ray.direction, vm.g, rng);
      res.estimator = atten;
    }
  }
  return res;
}

// evaluate_bsdf, main.py:1360-1444 (the transmission branch is unreachable:
// NdotV*NdotL > 0 always holds after the 1e-4 clamps, main.py:1370-1373)
static void evaluate_bsdf(const Scene& s, int vol_id, int surf_id, V3 V, V3 L, V3 N, int b, float& bsdf, float& pdf) {
  Mat vm = get_mat(s, vol_id, b);
  Mat sm = get_mat(s, surf_id, b);
  float n_i = vm.n, k_i = vm.k, n_t = sm.n, k_t = sm.k;
  float NdotL = tmax(1e-4f, dot(N, L));
  float NdotV = tmax(1e-4f, dot(N, V));
  bsdf = 0.0f;
  pdf = 0.0f;
  if (NdotV * NdotL > 0.0f) {
    V3 H = normalize(V + L);
    if (dot(H, N) < 0.0f) H = -H;
    float VdotH = tmax(1e-4f, dot(V, H));
    float F = fresnel_spectral(VdotH, n_i, k_i, n_t, k_t);
    if (sm.r > 0.008f) {
      float NdotH = tmax(1e-4f, dot(N, H));
      float D = ggx_D(NdotH, sm.r);
      float G = ggx_G(N, V, L, sm.r);
#if OR_PLANT == 5
      bsdf += (D * G * F) / (4.0f * NdotH * NdotV);  // planted: N.H for N.L
#else
      bsdf += (D * G * F) / (4.0f * NdotL * NdotV);
#endif
#if OR_PLANT == 4
      pdf += (D * NdotH * F) / VdotH;  // planted: the 4 dropped
#else
      pdf += (D * NdotH * F) / (4.0f * VdotH);
#endif
    }
    float absorb = (4.0f * PI_F * k_t * 1e9f) / (s.wl_min + (float)b * s.wl_step);
    float albedo = sm.ss / (sm.ss + absorb + 1e-6f);
    float on = oren_nayar(sm.r, NdotL, NdotV, L, V, N);
    float opq = (1.0f - (float)sm.tv);
    bsdf += (1.0f - F) * opq * (albedo / PI_F) * on;
    pdf += NdotL * 0.318309886184f * (1.0f - F) * opq;
  }
}

// shadow_transmittance, main.py:1446-1499
static float shadow_transmittance(const Scene& s, V3 origin, V3 dir, float max_dist, int tri_idx, int start_vol, int b,
                                  Counters& cnt, bool cull = false) {
  float total = 1.0f;
  float traveled = 0.0f;
  V3 current_pos = origin;
  int vol = start_vol;
  V3 spos = current_pos, sdir = dir;
  for (int i = 0; i < SCATTER_BOUNCES; ++i) {
    Mat vm = get_mat(s, vol, b);
    float absorb = (4.0f * PI_F * vm.k) / lam_m(s, b);
    if (max_dist - traveled >= 0.0f && shadow_dense(absorb + vm.ss)) cull = true;
    Hit hit = trace(s, spos, sdir);
    cnt.traversals++;
    if (cull) cnt.culled++;
    float dts = hit.t;
    total *= std::exp(-(absorb + vm.ss) * hit.t);
    if (hit.t >= (max_dist - traveled) || hit.tri_idx == tri_idx) {
      total *= std::exp(-(absorb + vm.ss) * (max_dist - traveled));
      break;
    }
    Mat sm = get_mat(s, hit.material_id, b);
    float n_i = vm.n;
    float n_t = sm.n;
    if (!hit.is_front_face) n_t = get_mat(s, 0, b).n;
    float ff = (float)hit.is_front_face;
    V3 inc = hit.shading_normal * ff - hit.shading_normal * (1.0f - ff);
    float ratio = n_i / n_t;
    V3 refr = refract(sdir, inc, ratio);
    float rt = (dot(sdir, refr) - 1.0f) * 37.9743f;
    float af = 1.0f / (1.0f + rt * rt * rt * rt * (3.0f - 2.0f * rt));
    total *= af;
    if (total < 1e-6f || !sm.tv) { total = 0.0f; break; }
    traveled += dts + EPSILON;
    spos = current_pos + sdir * (dts + EPSILON);
    vol = hit.material_id * hit.is_front_face;
  }
  return total;
}

// Light-tree pick.  A node's importance for a point x is its band power P
// (emitting area) times a bound on the emitters' cosine towards x over the
// squared distance to the node centre (floored at 1e-3 r^2 for points near
// the centre of a large node):  with theta the angle between the node's
// cone axis and x - centre, theta_o the cone half angle and theta_u the
// half angle of the bounding sphere seen from x, the bound is
// cos(max(0, theta - theta_o - theta_u)), 1 inside the sphere, and 0 when
// that angle reaches pi/2 (every emitter faces away).  Children are picked
// in proportion to their importance, with one uniform rescaled per level.
static float lt_importance(const LNode& nd, float P, V3 x) {
  if (!(P > 0.0f)) return 0.0f;
  V3 dv = x - v3(nd.c[0], nd.c[1], nd.c[2]);
  float r2 = nd.r2;
  float d2 = dot(dv, dv);
  float cb = 1.0f;
  if (d2 > r2) {
    float d = std::sqrt(d2);
    float th = std::acos(clampf(dot(v3(nd.axis[0], nd.axis[1], nd.axis[2]), dv) / d, -1.0f, 1.0f));
    float th_o = std::atan2(nd.sin_o, nd.cos_o);
    float th_u = std::asin(tmin(1.0f, nd.r / d));
    float a = tmax(0.0f, th - th_o - th_u);
    if (a >= 0.5f * PI_F) return 0.0f;
    cb = std::cos(a);
  }
  return P * cb / tmax(d2, 1e-3f * r2 + 1e-12f);
}

static int lt_sample(const Scene& s, int band, V3 x, float u, float& p) {
  const int n = (int)s.lt_nodes.size();
  p = 1.0f;
  if (s.lt_nodes[0].right < 0) return 0;
  int f = s.lt_nodes[0].left;
  for (int level = 0; level < 64; ++level) {  // bounded whatever the arrays hold
    float il = lt_importance(s.lt_nodes[f], s.lt_power[(size_t)band * n + f], x);
    float ir = lt_importance(s.lt_nodes[f + 1], s.lt_power[(size_t)band * n + f + 1], x);
    float sum = il + ir;
    if (!(sum > 0.0f)) return -1;
    float pl = il / sum;
    int c;
    if (u < pl || !(ir > 0.0f)) {
      u = tmin(u / pl, 0x1.fffffep-1f);
      p *= pl;
      c = f;
    } else {
      float pr = ir / sum;
      u = tmin((u - pl) / pr, 0x1.fffffep-1f);
      p *= pr;
      c = f + 1;
    }
    if (s.lt_nodes[c].right < 0) return c;
    f = s.lt_nodes[c].left;
  }
  return -1;
}

static float lt_pdf(const Scene& s, int band, V3 x, int leaf) {
  const int n = (int)s.lt_nodes.size();
  float p = 1.0f;
  int level = 0;
  for (int c = leaf; c != 0; c = s.lt_nodes[c].parent) {
    if (++level > 64) return 0.0f;
    int f = c & ~1;
    float il = lt_importance(s.lt_nodes[f], s.lt_power[(size_t)band * n + f], x);
    float ir = lt_importance(s.lt_nodes[f + 1], s.lt_power[(size_t)band * n + f + 1], x);
    float sum = il + ir;
    if (!(sum > 0.0f)) return 0.0f;
    p *= (c == f ? il : ir) / sum;
  }
  return p;
}

// sample_direct_light, main.py:1296-1358 (with the light tree, the picked
// light's area over its pick probability takes the place of the band's
// total emitting area in the estimator and the MIS pdf)
static NEE sample_direct_light(const Scene& s, const Ray& ray, V3 normVec, int surf_id, Rng& rng, Counters& cnt) {
  NEE res;
  V3 p = ray.position;
  const int wl = ray.active;
  int nl = s.nlights[wl];
  if (nl > 0) {
    rng.slot(SLOT_NEE);
    float r = rng.next();
    int lt_idx = 0;
    float area = s.light_area[wl];
    if (s.tree) {
      float pk;
      int leaf = lt_sample(s, wl, p, r, pk);
      if (leaf < 0) return res;  // every light of the band faces away
      lt_idx = s.lt_nodes[leaf].face;
      area = s.tris[lt_idx].area / pk;
    } else {
      for (int i = 0; i < nl; ++i) {
        if (r < s.light_cdf[wl][i]) { lt_idx = s.light_idx[wl][i]; break; }
      }
    }
    const Tri& lt = s.tris[lt_idx];
    float bx = rng.next(), by = rng.next();
    if (bx + by > 1.0f) { bx = 1.0f - bx; by = 1.0f - by; }
    V3 pl = lt.v0 + bx * (lt.v1 - lt.v0) + by * (lt.v2 - lt.v0);
    V3 dl = normalize(pl - p);
    float dist = length(pl - p);
    float tr = shadow_transmittance(s, p, dl, dist, lt_idx, ray.vol_mat_id, wl, cnt, ray.throughput == 0.0f);
    if (tr > 1e-6f) {
      V3 ln = lt.normal;
      float cl = tmax(0.0f, dot(ln, -dl));
      float d2 = dist * dist;
      float gterm = cl / (d2 + 1e-4f);
      float bsdf = 0.0f, cos_theta = 1.0f, pdf = 1.0f;
      if (surf_id != -1) {
        cos_theta = tmax(0.0f, dot(normVec, dl));
        evaluate_bsdf(s, ray.vol_mat_id, surf_id, -ray.direction, dl, normVec, wl, bsdf, pdf);
      } else {
        cos_theta = tmax(0.0f, dot(ray.direction, dl));
        bsdf = evaluate_henyey_greenstein(cos_theta, get_mat(s, ray.vol_mat_id, wl).g);
        pdf = bsdf;
      }
      res.contribution_factor = bsdf * cos_theta * tr * gterm * area;
      res.emission = get_mat(s, lt.mat, ray.transport).em;
      float pdf_light = d2 / (tmax(1e-6f, cl) * area);
      res.weight = (pdf_light * pdf_light) / (pdf_light * pdf_light + pdf * pdf + 1e-6f);
    }
  }
  return res;
}

// One single-band path of cast_rays (main.py:1502-1590) for pixel (i, j).
// Contributions are added to acc[band] (the pixel's B-vector).
static void cast_ray_path(const Scene& s, int W, int H, int i, int j, int band, Rng& rng, int depth, float* acc,
                          Counters& cnt) {
  rng.seg = 0;
  rng.slot(SLOT_AUX);
  float u = ((float)i + rng.next()) / (float)W;
  float v = ((float)j + rng.next()) / (float)H;
  const float aspect = (float)W / (float)H;
  V3 rd = normalize(s.cam_u * (u * 2.0f - 1.0f) * aspect * 1.0f + s.cam_v * (v * 2.0f - 1.0f) * 1.0f - s.cam_w * 2.0f);
  Ray ray;
  ray.position = s.cam_o;
  ray.direction = rd;
  ray.active = band;
  ray.transport = band;
  ray.vol_mat_id = 0;
  ray.throughput = 1.0f;
  ray.path_length = 0;
  ray.has_terminated = 0;
  ray.surface_pdf = 1.0f;
  int active_field = band;  // active_wavelength_idx[i, j]
  V3 nee_pos = ray.position;  // where the last NEE was taken (light-tree MIS pdf)
  for (int k = 0; k < depth; ++k) {
    if (ray.has_terminated) break;
    rng.seg = (uint32_t)k;
    // A non-finite throughput (e.g. the 0/0 rough BTDF estimator when a path
    // leaves a rough dielectric, main.py:1176-1185 with n_i == n_t) would
    // make every later contribution NaN; the path ends here instead.
    if (!std::isfinite(ray.throughput)) { cnt.nonfinite++; break; }
    float direct = 0.0f, indirect = 0.0f;
    bool nee_tp0 = false;  // this segment's NEE ran with throughput 0 (culled by the kernel)
    VolumeEvent ev = scatter_volume(s, ray, rng, cnt);
    cnt.segments++;
    if (ev.travel_distance < 0.0f) break;
    ray.position = ray.position + ray.direction * ev.travel_distance;
    ray.throughput *= ev.estimator;
    ray.penetrated_tri_idx = ev.penetrated_tri_idx;
    if (ev.hit_action == 1) {
      Mat sm = get_mat(s, ev.hit_mat_id, ray.transport);
      if (sm.se) {
        const Tri& lt = s.tris[ev.tri_idx];
        float d2 = ev.travel_distance * ev.travel_distance;
        float cl = tmax(0.0f, dot(lt.normal, -ray.direction));
        float lden = cl * s.light_area[ray.active] + 1e-6f;
        if (s.tree && cl > 0.0f && ray.path_length > 0) {
          // the tree's pick probability of this face from the previous NEE
          // vertex; a face the band's NEE cannot pick keeps the total area
          int leaf = s.lt_leaf[ev.tri_idx];
          if (leaf >= 0 && s.lt_power[(size_t)ray.active * s.lt_nodes.size() + leaf] > 0.0f) {
            float pk = lt_pdf(s, ray.active, nee_pos, leaf);
            if (pk > 0.0f) lden = cl * (lt.area / pk) + 1e-6f;
          }
        }
        float pdf_light = (d2 * tmin(1.0f, (float)ray.path_length)) / lden;
        float w = (ray.surface_pdf * ray.surface_pdf) /
                  (pdf_light * pdf_light + ray.surface_pdf * ray.surface_pdf + 1e-6f);
        indirect = ray.throughput * sm.em * w;
        ray.has_terminated = 1;
      } else {
        NEE ns = sample_direct_light(s, ray, ev.shading_normal, ev.hit_mat_id, rng, cnt);
        nee_pos = ray.position;
        nee_tp0 = ray.throughput == 0.0f;
        direct = ray.throughput * ns.emission * ns.contribution_factor * ns.weight;
        ray.path_length += 1;
        SurfaceEvent se = scatter_surface(s, ray, ev, rng);
        if (!std::isfinite(se.estimator)) {
          const Mat sm = get_mat(s, ev.hit_mat_id, ray.active), vm = get_mat(s, ray.vol_mat_id, ray.active);
          if (!se.did_reflect && sm.tv && sm.r >= 0.0018f && vm.n == sm.n) {
            cnt.nf_btdf_eta1++;
            if (k == depth - 1) cnt.nf_btdf_eta1_last++;
          } else {
            cnt.nf_other_est++;
          }
        }
        ray.surface_pdf = se.pdf;
        ray.direction = se.direction;
        ray.throughput *= se.estimator;
        float dr = (float)se.did_reflect;
        ray.position = ray.position + ev.normal * EPSILON * (dr - (1.0f - dr));
        ray.vol_mat_id = ev.hit_mat_id * ev.is_front_face * (1 - se.did_reflect) + ray.vol_mat_id * se.did_reflect;
        ray.is_reflected = se.did_reflect;
      }
    } else {
      NEE ns = sample_direct_light(s, ray, ev.shading_normal, ev.hit_mat_id, rng, cnt);
      nee_pos = ray.position;
      nee_tp0 = ray.throughput == 0.0f;
      direct = ray.throughput * ns.emission * ns.contribution_factor * ns.weight;
      if (ev.hit_action == 2) {
        ray.direction = ev.direction;
        ray.path_length += 1;
      } else if (ev.hit_action == 3) {
        ray.direction = ev.direction;
        ray.active = ev.new_wavelength_idx;
        active_field = ray.active;
        ray.path_length += 1;
      }
    }
    if (std::isfinite(indirect)) acc[active_field] += indirect; else { cnt.nonfinite++; cnt.nf_indirect++; }
    if (std::isfinite(direct)) {
      acc[active_field] += direct;
    } else {
      cnt.nonfinite++;
      cnt.nf_direct++;
      if (nee_tp0) cnt.culled_nf++;
    }
  }
}

}  // namespace orc

using namespace orc;

extern "C" {

void or_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out) { philox4x32_10(ctr, key, out); }

// Fill out[n] with the uniform stream of one path key.
void or_uniforms(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, int n, float* out) {
  Rng r;
  r.key[0] = (uint32_t)seed; r.key[1] = (uint32_t)(seed >> 32);
  r.c0 = c0; r.c1 = c1; r.c2 = c2; r.n = 0;
  for (int i = 0; i < n; ++i) out[i] = r.next();
}

// Scene construction: restates setup_scene main.py:509-631 on flattened
// arrays (tris_v/tris_n: [N][3][3] f32 vertex positions and per-vertex
// shading normals; mat_f32 [M][6][B]; mat_i32 [M][2][B]; eem [E][B][B]
// with E counting the unused id 0; cam = origin, lookat, up).
void* or_scene_create(int B, float wl_min, float wl_max, const float* tris_v, const float* tris_n, const int* tri_mat,
                      int N, const float* mat_f32, const int* mat_i32, const int* mat_eem, int M, const float* eem,
                      int E, const float* cam, int brute) {
  Scene* s = new Scene();
  s->B = B;
  s->wl_min = wl_min;
  s->wl_max = wl_max;
  s->wl_step = (float)(((double)wl_max - (double)wl_min) / (double)(B - 1));
  s->brute = brute != 0;
  s->tris.resize(N);
  for (int f = 0; f < N; ++f) {
    Tri& t = s->tris[f];
    const float* p = tris_v + (size_t)f * 9;
    const float* q = tris_n + (size_t)f * 9;
    t.v0 = v3(p[0], p[1], p[2]);
    t.v1 = v3(p[3], p[4], p[5]);
    t.v2 = v3(p[6], p[7], p[8]);
    t.n0 = v3(q[0], q[1], q[2]);
    t.n1 = v3(q[3], q[4], q[5]);
    t.n2 = v3(q[6], q[7], q[8]);
    V3 nrm = cross(t.v1 - t.v0, t.v2 - t.v0);  // main.py:526
    float area = length(nrm) / 2.0f;            // main.py:535
    if (area > 1e-6f) nrm = nrm / (area * 2.0f);  // main.py:536
    t.normal = nrm;
    t.area = area;
    t.mat = tri_mat[f];
  }
  s->M = M;
  s->E = E;
  s->mat_f32.assign(mat_f32, mat_f32 + (size_t)M * 6 * B);
  s->mat_i32.assign(mat_i32, mat_i32 + (size_t)M * 2 * B);
  s->mat_eem.assign(mat_eem, mat_eem + M);
  s->eem.assign((size_t)std::max(E, 1) * B * B, 0.0f);
  if (E > 0) std::copy(eem, eem + (size_t)E * B * B, s->eem.begin());
  // per-band light lists, main.py:590-627 (original face order; CDF in f64,
  // stored f32, last entry forced to 1)
  s->nlights.assign(B, 0);
  s->light_area.assign(B, 0.0f);
  s->light_cdf.assign(B, {});
  s->light_idx.assign(B, {});
  for (int b = 0; b < B; ++b) {
    std::vector<int> ids;
    double tot = 0.0;
    for (int f = 0; f < N; ++f) {
      int m = s->tris[f].mat;
      if (mat_i32[((size_t)m * 2 + 1) * B + b]) { ids.push_back(f); tot += (double)s->tris[f].area; }
    }
    s->nlights[b] = (int)ids.size();
    if (!ids.empty()) {
      s->light_area[b] = (float)tot;
      double c = 0.0;
      for (size_t i = 0; i < ids.size(); ++i) {
        c += (double)s->tris[ids[i]].area / tot;
        s->light_cdf[b].push_back((float)c);
        s->light_idx[b].push_back(ids[i]);
      }
      s->light_cdf[b].back() = 1.0f;
    }
  }
  // camera, update_camera main.py:633-640
  s->cam_o = v3(cam[0], cam[1], cam[2]);
  s->cam_lookat = v3(cam[3], cam[4], cam[5]);
  s->cam_up = v3(cam[6], cam[7], cam[8]);
  s->cam_w = normalize(s->cam_o - s->cam_lookat);
  s->cam_u = normalize(cross(s->cam_up, s->cam_w));
  s->cam_v = cross(s->cam_w, s->cam_u);
  if (!s->brute) build_reference_bvh(*s);
  return s;
}

void or_scene_destroy(void* p) { delete (Scene*)p; }

// Light tree for NEE (enable = 0 returns to the area-CDF pick): n nodes of
// 64 bytes in sibling pairs, power [B][n] f32, keys {face, leaf}[n_keys].
int or_scene_light_tree(void* p, int enable, int n, const void* nodes, const float* power, int n_keys,
                        const int* keys) {
  Scene* s = (Scene*)p;
  s->tree = false;
  if (!enable) return 0;
  if (n < 2 || n % 2 != 0) return -1;
  s->lt_nodes.resize(n);
  std::memcpy(s->lt_nodes.data(), nodes, (size_t)n * sizeof(LNode));
  s->lt_power.assign(power, power + (size_t)s->B * n);
  s->lt_leaf.assign(s->tris.size(), -1);
  for (int k = 0; k < n_keys; ++k) {
    if (keys[2 * k] < 0 || keys[2 * k] >= (int)s->tris.size() || keys[2 * k + 1] < 0 || keys[2 * k + 1] >= n) return -1;
    s->lt_leaf[keys[2 * k]] = keys[2 * k + 1];
  }
  s->tree = true;
  return 0;
}

int or_scene_info(void* p, int* n_nodes, int* max_lights) {
  Scene* s = (Scene*)p;
  *n_nodes = (int)s->nodes.size();
  int m = 0;
  for (int b = 0; b < s->B; ++b) m = std::max(m, s->nlights[b]);
  *max_lights = m;
  return 0;
}

// Closest hit for n rays; prim = original face id or -1, t = 1e9 on miss.
void or_trace(void* p, int n, const float* orig, const float* dir, float* out_t, int* out_prim, float* out_uv) {
  const Scene& s = *(Scene*)p;
#pragma omp parallel for schedule(dynamic, 256)
  for (int r = 0; r < n; ++r) {
    Hit h = trace(s, v3(orig[r * 3], orig[r * 3 + 1], orig[r * 3 + 2]), v3(dir[r * 3], dir[r * 3 + 1], dir[r * 3 + 2]));
    out_t[r] = h.t;
    out_prim[r] = h.tri_idx;
    out_uv[r * 2] = h.u;
    out_uv[r * 2 + 1] = h.v;
  }
}

// Renders blocks [first_block, first_block + n_blocks) for the pixel
// columns [i0, i1) and updates the running mean accum[W][H][B] in place
// (average_and_reset_block main.py:1592-1607 applied once per block).
// RNG key per path: c0 = i*H + j, c1 = block, c2 = band | (stream << 8).
// stats = {segments, traversals, nonfinite, culled shadow traversals, culled
// non-finite NEE terms, then the sources of non-finite values: rough-BTDF
// n_i == n_t estimators, those on a last segment, other estimators,
// non-finite NEE terms, non-finite emitter terms} (see Counters).
int or_render(void* p, int W, int H, uint64_t seed, uint32_t stream, int first_block, int n_blocks, int depth,
              int i0, int i1, float* accum, double* stats) {
  const Scene& s = *(Scene*)p;
  const int B = s.B;
  uint64_t seg = 0, trav = 0, nf = 0, cul = 0, cnf = 0, e1 = 0, e1l = 0, eo = 0, nd = 0, ni = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : seg, trav, nf, cul, cnf, e1, e1l, eo, nd, ni)
  for (int i = i0; i < i1; ++i) {
    std::vector<float> blk(B);
    Counters cnt;
    for (int j = 0; j < H; ++j) {
      float* a = accum + ((size_t)i * H + j) * B;
      for (int bl = first_block; bl < first_block + n_blocks; ++bl) {
        std::fill(blk.begin(), blk.end(), 0.0f);
        for (int band = 0; band < B; ++band) {
          Rng rng;
          rng.key[0] = (uint32_t)seed; rng.key[1] = (uint32_t)(seed >> 32);
          rng.c0 = (uint32_t)(i * H + j); rng.c1 = (uint32_t)bl; rng.c2 = (uint32_t)band | (stream << 8);
          rng.n = 0;
          cast_ray_path(s, W, H, i, j, band, rng, depth, blk.data(), cnt);
        }
        const float nb = (float)(bl + 1);
        for (int k = 0; k < B; ++k) a[k] = (a[k] * (nb - 1.0f) + blk[k]) / nb;
      }
    }
    seg += cnt.segments; trav += cnt.traversals; nf += cnt.nonfinite;
    cul += cnt.culled; cnf += cnt.culled_nf;
    e1 += cnt.nf_btdf_eta1; e1l += cnt.nf_btdf_eta1_last; eo += cnt.nf_other_est;
    nd += cnt.nf_direct; ni += cnt.nf_indirect;
  }
  if (stats) {
    stats[0] = (double)seg; stats[1] = (double)trav; stats[2] = (double)nf;
    stats[3] = (double)cul; stats[4] = (double)cnf;
    stats[5] = (double)e1; stats[6] = (double)e1l; stats[7] = (double)eo;
    stats[8] = (double)nd; stats[9] = (double)ni;
  }
  return 0;
}

// Light-tree pick on given arrays (nodes as 64-byte records in sibling
// pairs, power [B][n]): for each of the k points x[k][3] and uniforms u[k],
// the leaf lt_sample picks (-1: none) and its probability; or_lt_pdf gives
// lt_pdf of leaves[k] at x[k].
static Scene lt_scene(int n, const void* nodes, const float* power, int B) {
  Scene s;
  s.B = B;
  s.lt_nodes.resize(n);
  std::memcpy(s.lt_nodes.data(), nodes, (size_t)n * sizeof(LNode));
  s.lt_power.assign(power, power + (size_t)B * n);
  return s;
}
void or_lt_sample(int n, const void* nodes, const float* power, int B, int band, int k, const float* x,
                  const float* u, int* leaf, float* prob) {
  const Scene s = lt_scene(n, nodes, power, B);
  for (int i = 0; i < k; ++i) leaf[i] = lt_sample(s, band, v3(x[3 * i], x[3 * i + 1], x[3 * i + 2]), u[i], prob[i]);
}
void or_lt_pdf(int n, const void* nodes, const float* power, int B, int band, int k, const float* x, const int* leaves,
               float* prob) {
  const Scene s = lt_scene(n, nodes, power, B);
  for (int i = 0; i < k; ++i) prob[i] = lt_pdf(s, band, v3(x[3 * i], x[3 * i + 1], x[3 * i + 2]), leaves[i]);
}

// Unit-level entry points for closed-form tests.
float or_fresnel(float c, float ni, float ki, float nt, float kt) { return fresnel_spectral(c, ni, ki, nt, kt); }
float or_ggx_D(float nh, float r) { return ggx_D(nh, r); }
float or_ggx_G(const float* n, const float* v, const float* l, float r) {
  return ggx_G(v3(n[0], n[1], n[2]), v3(v[0], v[1], v[2]), v3(l[0], l[1], l[2]), r);
}
float or_hg_eval(float c, float g) { return evaluate_henyey_greenstein(c, g); }

// ---- BSDF probes on a one-band, two-material scene: material 0 is the
// medium the ray travels in (n_i, k_i), material 1 the surface it hits.
// mat = {n_i, k_i, n_t, k_t, roughness, scattering_coefficient,
// is_true_volume}; wl_nm is the band's wavelength.
static Scene probe_scene(const float* mat, float wl_nm) {
  Scene s;
  s.B = 1;
  s.wl_min = wl_nm;
  s.wl_max = wl_nm + 1.0f;
  s.wl_step = 1.0f;
  s.M = 2;
  s.mat_f32.assign(12, 0.0f);
  s.mat_i32.assign(4, 0);
  s.mat_eem.assign(2, 0);
  s.mat_f32[0] = mat[0];   // n_i
  s.mat_f32[1] = mat[1];   // k_i
  s.mat_i32[0] = 1;        // the medium is a volume
  s.mat_f32[6] = mat[2];   // n_t
  s.mat_f32[7] = mat[3];   // k_t
  s.mat_f32[8] = mat[4];   // roughness
  s.mat_f32[9] = mat[5];   // scattering_coefficient
  s.mat_i32[2] = (int)mat[6];
  s.E = 1;
  s.eem.assign(1, 0.0f);
  s.nlights.assign(1, 0);
  s.light_area.assign(1, 0.0f);
  return s;
}

float or_oren_nayar(float r, float NdotL, float NdotV, const float* l, const float* v, const float* n) {
  return oren_nayar(r, NdotL, NdotV, v3(l[0], l[1], l[2]), v3(v[0], v[1], v[2]), v3(n[0], n[1], n[2]));
}

// n samples of scatter_surface for a ray with direction dir hitting a
// surface with shading normal nrm; sample i uses Philox counter {i, 0, 0}
// (slotted blocks of segment 0) with key seed.  out[i] = {direction (3),
// estimator, pdf, did_reflect, half vector (3)}.
void or_scatter_surface(const float* mat, float wl_nm, const float* dir, const float* nrm, uint64_t seed, int n,
                        float* out) {
  Scene s = probe_scene(mat, wl_nm);
  for (int i = 0; i < n; ++i) {
    Ray ray;
    ray.position = v3(0, 0, 0);
    ray.direction = v3(dir[0], dir[1], dir[2]);
    ray.active = ray.transport = 0;
    ray.vol_mat_id = 0;
    VolumeEvent ev;
    ev.hit_mat_id = 1;
    ev.is_front_face = 1;
    ev.shading_normal = ev.normal = v3(nrm[0], nrm[1], nrm[2]);
    Rng rng;
    rng.key[0] = (uint32_t)seed; rng.key[1] = (uint32_t)(seed >> 32);
    rng.c0 = (uint32_t)i; rng.c1 = 0; rng.c2 = 0;
    rng.seg = 0;
    V3 H;
    SurfaceEvent se = scatter_surface(s, ray, ev, rng, &H);
    float* o = out + (size_t)i * 9;
    o[0] = se.direction.x; o[1] = se.direction.y; o[2] = se.direction.z;
    o[3] = se.estimator; o[4] = se.pdf; o[5] = (float)se.did_reflect;
    o[6] = H.x; o[7] = H.y; o[8] = H.z;
  }
}

// evaluate_bsdf for n (V, L) pairs about the normal nrm: out[i] = {bsdf, pdf}.
void or_eval_bsdf(const float* mat, float wl_nm, int n, const float* V, const float* L, const float* nrm, float* out) {
  Scene s = probe_scene(mat, wl_nm);
  const V3 N = v3(nrm[0], nrm[1], nrm[2]);
  for (int i = 0; i < n; ++i) {
    float b, p;
    evaluate_bsdf(s, 0, 1, v3(V[3 * i], V[3 * i + 1], V[3 * i + 2]), v3(L[3 * i], L[3 * i + 1], L[3 * i + 2]), N, 0,
                  b, p);
    out[2 * i] = b;
    out[2 * i + 1] = p;
  }
}
// Samplers driven by a path key; out[n*3] directions.
void or_sample(int which, uint64_t seed, int n, const float* axis, float param, float* out) {
  V3 ax = v3(axis[0], axis[1], axis[2]);
  for (int i = 0; i < n; ++i) {
    Rng r;
    r.key[0] = (uint32_t)seed; r.key[1] = (uint32_t)(seed >> 32);
    r.c0 = (uint32_t)i; r.c1 = 0; r.c2 = 0; r.n = 0;
    V3 d;
    if (which == 0) d = cosine_weighted_hemisphere_direction(ax, r);
    else if (which == 1) d = sample_henyey_greenstein(ax, param, r);
    else if (which == 2) { float a = r.next(); float b = r.next(); d = importance_sample_ggx(a, b, ax, param); }
    else d = random_in_unit_sphere(r);
    out[i * 3] = d.x; out[i * 3 + 1] = d.y; out[i * 3 + 2] = d.z;
  }
}
// shadow_transmittance / sample_direct_light / scatter_volume probes
float or_shadow(void* p, const float* o, const float* d, float max_dist, int tri, int vol, int band) {
  Counters c;
  return shadow_transmittance(*(Scene*)p, v3(o[0], o[1], o[2]), v3(d[0], d[1], d[2]), max_dist, tri, vol, band, c);
}
// NEE at a point: out = {factor, emission, weight} averaged over n samples
void or_nee(void* p, const float* pos, const float* dir, const float* nrm, int surf, int vol, int band, uint64_t seed,
            int n, double* out) {
  const Scene& s = *(Scene*)p;
  double acc[3] = {0, 0, 0};
  for (int i = 0; i < n; ++i) {
    Rng r;
    r.key[0] = (uint32_t)seed; r.key[1] = (uint32_t)(seed >> 32);
    r.c0 = (uint32_t)i; r.c1 = 1; r.c2 = 0; r.n = 0;
    Ray ray;
    ray.position = v3(pos[0], pos[1], pos[2]);
    ray.direction = v3(dir[0], dir[1], dir[2]);
    ray.active = ray.transport = band;
    ray.vol_mat_id = vol;
    Counters c;
    NEE e = sample_direct_light(s, ray, v3(nrm[0], nrm[1], nrm[2]), surf, r, c);
    acc[0] += (double)e.contribution_factor * e.weight;
    acc[1] += e.emission;
    acc[2] += (double)e.contribution_factor;
  }
  out[0] = acc[0] / n; out[1] = acc[1] / n; out[2] = acc[2] / n;
}
// scatter_volume free-flight probe: returns travel distances of n events
void or_volume_events(void* p, const float* pos, const float* dir, int vol, int band, uint64_t seed, int n,
                      float* travel, int* action, float* est, int* new_band) {
  const Scene& s = *(Scene*)p;
  for (int i = 0; i < n; ++i) {
    Rng r;
    r.key[0] = (uint32_t)seed; r.key[1] = (uint32_t)(seed >> 32);
    r.c0 = (uint32_t)i; r.c1 = 2; r.c2 = 0; r.n = 0;
    Ray ray;
    ray.position = v3(pos[0], pos[1], pos[2]);
    ray.direction = v3(dir[0], dir[1], dir[2]);
    ray.active = ray.transport = band;
    ray.vol_mat_id = vol;
    Counters c;
    VolumeEvent e = scatter_volume(s, ray, r, c);
    travel[i] = e.travel_distance; action[i] = e.hit_action; est[i] = e.estimator; new_band[i] = e.new_wavelength_idx;
  }
}

}  // extern "C"
