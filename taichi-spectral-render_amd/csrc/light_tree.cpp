// light_tree.cpp -- host builder of the light tree used by the many-light
// NEE mode (SR_RENDER_LIGHT_TREE).  The reference picks a light by its area
// fraction of the band's emitting area (main.py:1306-1313, a CDF scan) and
// leaves a light BVH as a TODO (main.py:541).  Here the light slots (the
// emissive faces, main.py:590-627) are organised as a binary tree: median
// splits of the slot centroids along the longest axis, each node holding the
// bounding sphere (the one around the bounding box) and the normal cone of
// the emitters below (Conty Estevez & Kulla, "Importance Sampling of Many
// Lights with Adaptive Tree Splitting", 2018: cone union) and the emitting
// area per band.  The kernel descends it with one uniform, choosing
// children in proportion to a conservative importance (area x orientation
// bound / distance^2), so every emitter that can face the shading point
// keeps a non-zero probability.
//
// Deterministic (stable sort, fixed arithmetic), so host and device scene
// builds, which both call it, produce the same arrays.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "sr_layout.h"

namespace {

constexpr double kMargin = 2e-3;  // radians

struct Cone {
  double ax[3];
  double th;  // half angle, pi = every direction
};

double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

Cone cone_union(Cone a, Cone b) {
  const double kPi = 3.14159265358979323846;
  if (a.th >= kPi) return a;
  if (b.th >= kPi) return b;
  if (b.th > a.th) std::swap(a, b);
  const double td = std::acos(std::max(-1.0, std::min(1.0, dot3(a.ax, b.ax))));
  if (std::min(td + b.th, kPi) <= a.th) return a;
  const double tn = 0.5 * (a.th + td + b.th);
  if (tn >= kPi) return Cone{{a.ax[0], a.ax[1], a.ax[2]}, kPi};
  const double tr = tn - a.th;
  double w[3] = {a.ax[1] * b.ax[2] - a.ax[2] * b.ax[1], a.ax[2] * b.ax[0] - a.ax[0] * b.ax[2],
                 a.ax[0] * b.ax[1] - a.ax[1] * b.ax[0]};
  const double wl = std::sqrt(dot3(w, w));
  if (wl < 1e-12) return Cone{{a.ax[0], a.ax[1], a.ax[2]}, kPi};  // opposite axes
  for (double& x : w) x /= wl;
  // rotate a.ax about w (perpendicular to it) by tr towards b.ax
  const double wa[3] = {w[1] * a.ax[2] - w[2] * a.ax[1], w[2] * a.ax[0] - w[0] * a.ax[2],
                        w[0] * a.ax[1] - w[1] * a.ax[0]};
  Cone r;
  const double c = std::cos(tr), s = std::sin(tr);
  for (int k = 0; k < 3; ++k) r.ax[k] = a.ax[k] * c + wa[k] * s;
  const double l = std::sqrt(dot3(r.ax, r.ax));
  for (double& x : r.ax) x /= l;
  r.th = tn;
  return r;
}

struct Builder {
  int B;
  const std::vector<double>& cen;    // [n][3]
  const std::vector<float>& box;     // [n][6]
  const std::vector<Cone>& cone;     // [n]
  const std::vector<double>& power;  // [n][B]
  const std::vector<int>& face;      // [n]
  std::vector<int> idx;
  std::vector<LightNode> nodes;
  std::vector<double> npow;  // [nodes][B] while building; stored [B][nodes]
  std::vector<Cone> ncone;
  std::vector<float> nbox;  // [nodes][6]: lo, hi

  int build(int lo_i, int hi_i, int parent) {
    const int me = (int)nodes.size();
    nodes.emplace_back();
    npow.resize(npow.size() + B, 0.0);
    ncone.emplace_back();
    nbox.resize(nbox.size() + 6);
    float lo[3], hi[3];
    LightNode nd;
    std::memset(&nd, 0, sizeof(nd));
    nd.parent = parent;
    if (hi_i - lo_i == 1) {
      const int s = idx[lo_i];
      for (int k = 0; k < 3; ++k) { lo[k] = box[s * 6 + k]; hi[k] = box[s * 6 + 3 + k]; }
      nd.left = ~s;
      nd.right = -1;
      nd.face = face[s];
      for (int b = 0; b < B; ++b) npow[(size_t)me * B + b] = power[(size_t)s * B + b];
      ncone[me] = cone[s];
    } else {
      double cl[3] = {1e300, 1e300, 1e300}, ch[3] = {-1e300, -1e300, -1e300};
      for (int i = lo_i; i < hi_i; ++i)
        for (int k = 0; k < 3; ++k) {
          cl[k] = std::min(cl[k], cen[(size_t)idx[i] * 3 + k]);
          ch[k] = std::max(ch[k], cen[(size_t)idx[i] * 3 + k]);
        }
      int ax = 0;
      for (int k = 1; k < 3; ++k)
        if (ch[k] - cl[k] > ch[ax] - cl[ax]) ax = k;
      std::stable_sort(idx.begin() + lo_i, idx.begin() + hi_i,
                       [&](int a, int b) { return cen[(size_t)a * 3 + ax] < cen[(size_t)b * 3 + ax]; });
      const int mid = lo_i + (hi_i - lo_i) / 2;
      const int l = build(lo_i, mid, me);
      const int r = build(mid, hi_i, me);
      for (int k = 0; k < 3; ++k) {
        lo[k] = std::min(nbox[(size_t)l * 6 + k], nbox[(size_t)r * 6 + k]);
        hi[k] = std::max(nbox[(size_t)l * 6 + 3 + k], nbox[(size_t)r * 6 + 3 + k]);
      }
      nd.left = l;
      nd.right = r;
      nd.face = -1;
      for (int b = 0; b < B; ++b) npow[(size_t)me * B + b] = npow[(size_t)l * B + b] + npow[(size_t)r * B + b];
      ncone[me] = cone_union(ncone[l], ncone[r]);
    }
    for (int k = 0; k < 3; ++k) {
      nbox[(size_t)me * 6 + k] = lo[k];
      nbox[(size_t)me * 6 + 3 + k] = hi[k];
    }
    // the sphere around the box, in the f32 operations (uncontracted) a
    // kernel would use to form it
    float ext[3];
    for (int k = 0; k < 3; ++k) {
      nd.c[k] = (lo[k] + hi[k]) * 0.5f;
      ext[k] = hi[k] - lo[k];
    }
    nd.r2 = 0.25f * (ext[0] * ext[0] + ext[1] * ext[1] + ext[2] * ext[2]);
    nd.r = std::sqrt(nd.r2);
    const Cone& c = ncone[me];
    for (int k = 0; k < 3; ++k) nd.axis[k] = (float)c.ax[k];
    // the half angle widened by kMargin (f32 rounding of the stored axis,
    // of the cosines and of the kernel's angle arithmetic), as cos and sin
    const double th = std::min(c.th + kMargin, 3.14159265358979323846);
    nd.cos_o = (float)std::cos(th);
    nd.sin_o = (float)std::sin(th);
    nodes[me] = nd;
    return me;
  }
};

}  // namespace

// slots: emissive faces in light-slot order; lverts [n][3][4] f32 (xyz_);
// slot_mat: material of each slot; se [M][B]: surface_emission flags.
// Writes nodes (sibling pairs, root 0, node 1 unused), power [B][nodes]
// (emitting area per band, f32 of the f64 sums) and keys (face -> leaf,
// sorted by face).
int sr_build_light_tree(int n, const float* lverts, const int* slots, const int* slot_mat, const unsigned char* se,
                        int B, std::vector<LightNode>& nodes, std::vector<float>& power, std::vector<LightKey>& keys) {
  nodes.clear();
  power.clear();
  keys.clear();
  if (n <= 0) return 0;
  std::vector<double> cen((size_t)n * 3), pw((size_t)n * B, 0.0);
  std::vector<float> box((size_t)n * 6);
  std::vector<Cone> cone((size_t)n);
  std::vector<int> face(slots, slots + n);
  for (int s = 0; s < n; ++s) {
    const float* p = lverts + (size_t)s * 12;
    float e1[3], e2[3], nn[3];
    for (int k = 0; k < 3; ++k) {
      e1[k] = p[4 + k] - p[k];
      e2[k] = p[8 + k] - p[k];
    }
    nn[0] = e1[1] * e2[2] - e1[2] * e2[1];
    nn[1] = e1[2] * e2[0] - e1[0] * e2[2];
    nn[2] = e1[0] * e2[1] - e1[1] * e2[0];
    // face area as the scene compile forms it (main.py:535)
    const float area = std::sqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]) / 2.0f;
    for (int k = 0; k < 3; ++k) {
      box[(size_t)s * 6 + k] = std::min(std::min(p[k], p[4 + k]), p[8 + k]);
      box[(size_t)s * 6 + 3 + k] = std::max(std::max(p[k], p[4 + k]), p[8 + k]);
      cen[(size_t)s * 3 + k] = ((double)p[k] + p[4 + k] + p[8 + k]) / 3.0;
    }
    const double l = std::sqrt((double)nn[0] * nn[0] + (double)nn[1] * nn[1] + (double)nn[2] * nn[2]);
    Cone c;
    if (l > 0.0) {
      for (int k = 0; k < 3; ++k) c.ax[k] = nn[k] / l;
      c.th = 0.0;
    } else {  // degenerate face: no orientation bound
      c.ax[0] = 0.0; c.ax[1] = 1.0; c.ax[2] = 0.0;
      c.th = 3.14159265358979323846;
    }
    cone[s] = c;
    for (int b = 0; b < B; ++b)
      if (se[(size_t)slot_mat[s] * B + b]) pw[(size_t)s * B + b] = area;
  }
  Builder bl{B, cen, box, cone, pw, face, {}, {}, {}, {}, {}};
  bl.idx.resize(n);
  for (int s = 0; s < n; ++s) bl.idx[s] = s;
  bl.nodes.reserve((size_t)2 * n);
  bl.build(0, n, -1);
  // Sibling-pair layout: root at 0, an unused node at 1, and the two
  // children of every inner node at an even index and the next (breadth
  // first), so one descent step reads both children as one 128-byte line
  // and their powers in one band row as one 8-byte pair.
  const int nb = (int)bl.nodes.size();
  const int nn = nb + 1;
  std::vector<int> at((size_t)nb, -1), order;
  order.reserve((size_t)nn);
  at[0] = 0;
  order.push_back(0);
  order.push_back(-1);  // the unused node
  for (size_t q = 0; q < order.size(); ++q) {
    const int o = order[q];
    if (o < 0 || bl.nodes[o].right < 0) continue;
    at[bl.nodes[o].left] = (int)order.size();
    order.push_back(bl.nodes[o].left);
    at[bl.nodes[o].right] = (int)order.size();
    order.push_back(bl.nodes[o].right);
  }
  nodes.assign((size_t)nn, LightNode{});
  power.assign((size_t)nn * B, 0.0f);
  for (int i = 0; i < nn; ++i) {
    const int o = order[i];
    LightNode& d = nodes[i];
    if (o < 0) {
      std::memset(&d, 0, sizeof(d));
      d.left = ~0;
      d.right = -1;
      d.parent = -1;
      d.face = -1;
      continue;
    }
    d = bl.nodes[o];
    if (d.right >= 0) {
      d.left = at[d.left];
      d.right = at[d.right];
    }
    d.parent = d.parent >= 0 ? at[d.parent] : -1;
    for (int b = 0; b < B; ++b) power[(size_t)b * nn + i] = (float)bl.npow[(size_t)o * B + b];
  }
  for (int i = 0; i < nn; ++i)
    if (order[i] >= 0 && nodes[i].right < 0) keys.push_back(LightKey{nodes[i].face, i});
  std::sort(keys.begin(), keys.end(), [](const LightKey& a, const LightKey& b) { return a.face < b.face; });
  return nn;
}
