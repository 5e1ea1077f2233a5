// bvh4_build.cpp -- host builder of the 4-wide BVH (128-byte node lines).
//
// Replaces the reference's build_bvh (main.py:437-507: binary SAH with an
// O(n^2)-per-node exhaustive sweep and leaves of <= 10 triangles), which is
// infeasible at 1M triangles.  Closest-hit results do not depend on the
// tree's shape, only on the triangle set, so this builder is free to choose
// its own: binned SAH (16 bins per axis) into a binary tree with leaves of
// <= 4 triangles, then a collapse to 4-wide nodes that fills each node by
// opening its largest-area inner entry until it has four slots.  Depth is
// bounded (SR_MAX_BVH4_DEPTH): the binary builder switches to object-median
// splits when a subtree runs short of depth budget, and a collapse that
// would come out deeper is redone taking children or grandchildren (half the
// binary depth), so the per-lane traversal stack can never overflow.  Child
// boxes are padded outward so the f32 slab test is conservative for every ray
// that hits a contained triangle.
//
// Every step is deterministic -- stable partitions, a stable sort for the
// median split, first-minimum ties in the SAH scan and the collapse -- so
// the device builder (sr_build.hip) reproduces this tree, node array and
// triangle order byte for byte.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "sr_layout.h"

namespace {

struct AABB {
  float lo[3], hi[3];
  void empty() {
    for (int a = 0; a < 3; ++a) { lo[a] = INFINITY; hi[a] = -INFINITY; }
  }
  void grow(const AABB& o) {
    for (int a = 0; a < 3; ++a) { lo[a] = std::min(lo[a], o.lo[a]); hi[a] = std::max(hi[a], o.hi[a]); }
  }
  void grow(const float* p) {
    for (int a = 0; a < 3; ++a) { lo[a] = std::min(lo[a], p[a]); hi[a] = std::max(hi[a], p[a]); }
  }
  float area() const {
    float d0 = hi[0] - lo[0], d1 = hi[1] - lo[1], d2 = hi[2] - lo[2];
    if (!(d0 >= 0.f)) return 0.f;
    return 2.0f * (d0 * d1 + d1 * d2 + d2 * d0);
  }
};

struct B2 {
  AABB box;
  int left = -1, right = -1;
  int first = 0, count = 0;  // leaf when count > 0
};

struct Build2 {
  const std::vector<AABB>& tb;
  const std::vector<float>& cen;  // [N][3]
  std::vector<int>& idx;
  std::vector<B2> nodes;
  int depth_limit;

  Build2(const std::vector<AABB>& b, const std::vector<float>& c, std::vector<int>& i, int dl)
      : tb(b), cen(c), idx(i), depth_limit(dl) {}

  static int ceil_log2(size_t n) {
    int l = 0;
    size_t v = 1;
    while (v < n) { v <<= 1; ++l; }
    return l;
  }

  int build(size_t lo, size_t hi, int depth) {
    int ni = (int)nodes.size();
    nodes.emplace_back();
    AABB box, cb;
    box.empty();
    cb.empty();
    for (size_t i = lo; i < hi; ++i) {
      box.grow(tb[idx[i]]);
      cb.grow(&cen[(size_t)idx[i] * 3]);
    }
    const size_t n = hi - lo;
    if (n <= SR_LEAF_MAX) {
      nodes[ni].box = box;
      nodes[ni].first = (int)lo;
      nodes[ni].count = (int)n;
      return ni;
    }
    // leaves needed below: ceil(n / LEAF_MAX) -> levels needed by a median
    // split: ceil_log2 of that; force median when the budget is tight.
    const int need = ceil_log2((n + SR_LEAF_MAX - 1) / SR_LEAF_MAX);
    const bool force_median = depth + need + 1 >= depth_limit;
    size_t mid = lo;
    bool done = false;
    if (!force_median) {
      const int NB = 16;
      float best = INFINITY;
      int best_axis = -1, best_bin = -1;
      for (int a = 0; a < 3; ++a) {
        float ext = cb.hi[a] - cb.lo[a];
        if (!(ext > 0.f)) continue;
        AABB bb[NB];
        int bc[NB];
        for (int b = 0; b < NB; ++b) { bb[b].empty(); bc[b] = 0; }
        const float scale = NB / ext;
        for (size_t i = lo; i < hi; ++i) {
          int t = idx[i];
          int b = (int)((cen[(size_t)t * 3 + a] - cb.lo[a]) * scale);
          b = std::min(std::max(b, 0), NB - 1);
          bb[b].grow(tb[t]);
          bc[b]++;
        }
        float rA[NB];
        int rC[NB];
        AABB acc;
        acc.empty();
        int cnt = 0;
        for (int b = NB - 1; b > 0; --b) {
          acc.grow(bb[b]);
          cnt += bc[b];
          rA[b] = acc.area();
          rC[b] = cnt;
        }
        acc.empty();
        cnt = 0;
        for (int b = 0; b < NB - 1; ++b) {
          acc.grow(bb[b]);
          cnt += bc[b];
          if (cnt == 0 || rC[b + 1] == 0) continue;
          float cost = acc.area() * (float)cnt + rA[b + 1] * (float)rC[b + 1];
          if (cost < best) { best = cost; best_axis = a; best_bin = b; }
        }
      }
      if (best_axis >= 0) {
        const int a = best_axis;
        const float ext = cb.hi[a] - cb.lo[a];
        const float scale = NB / ext;
        auto it = std::stable_partition(idx.begin() + lo, idx.begin() + hi, [&](int t) {
          int b = (int)((cen[(size_t)t * 3 + a] - cb.lo[a]) * scale);
          b = std::min(std::max(b, 0), NB - 1);
          return b <= best_bin;
        });
        mid = (size_t)(it - idx.begin());
        done = mid > lo && mid < hi;
      }
    }
    if (!done) {
      int a = 0;
      float e = -1.f;
      for (int k = 0; k < 3; ++k) {
        float ek = cb.hi[k] - cb.lo[k];
        if (ek > e) { e = ek; a = k; }
      }
      mid = lo + n / 2;
      std::stable_sort(idx.begin() + lo, idx.begin() + hi,
                       [&](int x, int y) { return cen[(size_t)x * 3 + a] < cen[(size_t)y * 3 + a]; });
    }
    int l = build(lo, mid, depth + 1);
    int r = build(mid, hi, depth + 1);
    nodes[ni].box = box;
    nodes[ni].left = l;
    nodes[ni].right = r;
    nodes[ni].count = 0;
    return ni;
  }
};

struct Collapse {
  const std::vector<B2>& b2;
  const std::vector<int>& idx;
  float pad;
  std::vector<Node4>& out;
  std::vector<int>& tri_order;
  int max_depth = 0;
  bool greedy = true;

  Collapse(const std::vector<B2>& b, const std::vector<int>& i, float p, std::vector<Node4>& o, std::vector<int>& t)
      : b2(b), idx(i), pad(p), out(o), tri_order(t) {}

  // Emit the 4-wide node for binary node `n` (inner) in preorder.
  int emit(int n, int depth) {
    max_depth = std::max(max_depth, depth);
    int slots[4];
    int ns = 0;
    const B2& nd = b2[n];
    if (greedy) {
      // open the inner frontier entry with the largest box area until the
      // node has four slots (fewer, fuller nodes than taking grandchildren)
      slots[0] = nd.left;
      slots[1] = nd.right;
      ns = 2;
      while (ns < 4) {
        int best = -1;
        float ba = -1.f;
        for (int q = 0; q < ns; ++q) {
          const B2& c = b2[slots[q]];
          if (c.count > 0) continue;
          const float ar = c.box.area();
          if (ar > ba) { ba = ar; best = q; }
        }
        if (best < 0) break;
        const B2& c = b2[slots[best]];
        for (int q = ns; q > best + 1; --q) slots[q] = slots[q - 1];
        slots[best] = c.left;
        slots[best + 1] = c.right;
        ++ns;
      }
    } else {  // children, or grandchildren of inner children: depth <= binary depth / 2
      for (int c : {nd.left, nd.right}) {
        const B2& cn = b2[c];
        if (cn.count > 0) slots[ns++] = c;
        else { slots[ns++] = cn.left; slots[ns++] = cn.right; }
      }
    }
    int me = (int)out.size();
    out.emplace_back();
    Node4 tmp;
    std::memset(&tmp, 0, sizeof(tmp));
    for (int s = 0; s < 4; ++s) {
      if (s >= ns) {
        tmp.lo_x[s] = tmp.lo_y[s] = tmp.lo_z[s] = 3.0e38f;
        tmp.hi_x[s] = tmp.hi_y[s] = tmp.hi_z[s] = -3.0e38f;
        tmp.child[s] = SR_EMPTY;
        continue;
      }
      const B2& cn = b2[slots[s]];
      tmp.lo_x[s] = cn.box.lo[0] - pad;
      tmp.lo_y[s] = cn.box.lo[1] - pad;
      tmp.lo_z[s] = cn.box.lo[2] - pad;
      tmp.hi_x[s] = cn.box.hi[0] + pad;
      tmp.hi_y[s] = cn.box.hi[1] + pad;
      tmp.hi_z[s] = cn.box.hi[2] + pad;
      if (cn.count > 0) {
        uint32_t first = (uint32_t)tri_order.size();
        for (int k = 0; k < cn.count; ++k) tri_order.push_back(idx[cn.first + k]);
        tmp.child[s] = sr_leaf_ref(first, (uint32_t)cn.count);
        max_depth = std::max(max_depth, depth + 1);
      } else {
        tmp.child[s] = -1;  // patched below
      }
    }
    for (int s = 0; s < ns; ++s) {
      if (tmp.child[s] == -1) tmp.child[s] = emit(slots[s], depth + 1);
    }
    out[me] = tmp;
    return me;
  }
};

}  // namespace

// Builds the 4-wide BVH over n triangles (v: [n][3][3] f32).  Writes the
// node array and the leaf-order -> original-face permutation.  Returns the
// 4-wide depth (root = 0, leaf slots one below their node), or -1.
int sr_build_bvh4(const float* v, int n, std::vector<Node4>& nodes, std::vector<int>& tri_order) {
  nodes.clear();
  tri_order.clear();
  if (n <= 0) return 0;
  std::vector<AABB> tb((size_t)n);
  std::vector<float> cen((size_t)n * 3);
  float maxabs = 1e-3f;
  for (int i = 0; i < n; ++i) {
    const float* p = v + (size_t)i * 9;
    tb[i].empty();
    for (int k = 0; k < 3; ++k) tb[i].grow(p + 3 * k);
    for (int a = 0; a < 3; ++a) {
      cen[(size_t)i * 3 + a] = (p[a] + p[3 + a] + p[6 + a]) * (1.0f / 3.0f);
      maxabs = std::max(maxabs, std::max(std::fabs(tb[i].lo[a]), std::fabs(tb[i].hi[a])));
    }
  }
  // Slab-test rounding is ~1e-7 relative to t; 4e-6 of the scene scale
  // keeps every containing box hit for rays that start inside the scene.
  const float pad = 4e-6f * maxabs;
  std::vector<int> idx((size_t)n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  Build2 b(tb, cen, idx, 2 * SR_MAX_BVH4_DEPTH);
  b.nodes.reserve((size_t)n / 2 + 8);
  int root = b.build(0, (size_t)n, 0);
  Collapse c(b.nodes, idx, pad, nodes, tri_order);
  if (b.nodes[root].count > 0) {
    // single leaf: wrap it in one 4-wide node
    B2 wrap;
    wrap.box = b.nodes[root].box;
    wrap.left = root;
    wrap.right = -1;
    Node4 t;
    std::memset(&t, 0, sizeof(t));
    for (int s = 0; s < 4; ++s) {
      t.lo_x[s] = t.lo_y[s] = t.lo_z[s] = 3.0e38f;
      t.hi_x[s] = t.hi_y[s] = t.hi_z[s] = -3.0e38f;
      t.child[s] = SR_EMPTY;
    }
    const AABB& bx = b.nodes[root].box;
    t.lo_x[0] = bx.lo[0] - pad; t.lo_y[0] = bx.lo[1] - pad; t.lo_z[0] = bx.lo[2] - pad;
    t.hi_x[0] = bx.hi[0] + pad; t.hi_y[0] = bx.hi[1] + pad; t.hi_z[0] = bx.hi[2] + pad;
    for (int k = 0; k < b.nodes[root].count; ++k) tri_order.push_back(idx[b.nodes[root].first + k]);
    t.child[0] = sr_leaf_ref(0, (uint32_t)b.nodes[root].count);
    nodes.push_back(t);
    return 1;
  }
  c.emit(root, 0);
  if (c.max_depth > SR_MAX_BVH4_DEPTH) {
    // the greedy collapse can deepen a lopsided tree: take grandchildren,
    // which keeps the 4-wide depth at half the bounded binary depth
    nodes.clear();
    tri_order.clear();
    c.greedy = false;
    c.max_depth = 0;
    c.emit(root, 0);
  }
  if ((int)tri_order.size() != n) return -1;
  return c.max_depth;
}
