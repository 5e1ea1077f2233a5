// host_check.cpp -- CPU check of the host scene builders (bvh4_build.cpp,
// light_tree.cpp) on adversarial inputs, built with AddressSanitizer and
// UndefinedBehaviorSanitizer by `make asan`.  Every input is checked for
// the builders' structural guarantees: the BVH's leaf order is a
// permutation of the faces, every child reference is in range, leaves hold
// 1..SR_LEAF_MAX faces, every face lies inside each box on its root path and
// the depth fits the traversal stack; the light tree has 2n nodes in
// sibling pairs, each emitter face owns one leaf, and each band's root power
// is the sum of its leaves.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "sr_layout.h"

int sr_build_bvh4(const float* v, int n, std::vector<Node4>& nodes, std::vector<int>& tri_order);
int sr_build_light_tree(int n, const float* lverts, const int* slots, const int* slot_mat, const unsigned char* se,
                        int B, std::vector<LightNode>& nodes, std::vector<float>& power, std::vector<LightKey>& keys);

static int g_fail = 0;
#define CHECK(c, ...)                                   \
  do {                                                  \
    if (!(c)) {                                         \
      std::fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
      std::fprintf(stderr, __VA_ARGS__);                \
      std::fprintf(stderr, "\n");                       \
      ++g_fail;                                         \
      return;                                           \
    }                                                   \
  } while (0)

static bool inside(const Node4& nd, int s, const float* tri) {
  for (int k = 0; k < 3; ++k) {
    const float* p = tri + 3 * k;
    if (p[0] < nd.lo_x[s] || p[0] > nd.hi_x[s] || p[1] < nd.lo_y[s] || p[1] > nd.hi_y[s] || p[2] < nd.lo_z[s] ||
        p[2] > nd.hi_z[s])
      return false;
  }
  return true;
}

static void check_bvh(const char* name, const std::vector<float>& v) {
  const int n = (int)(v.size() / 9);
  std::vector<Node4> nodes;
  std::vector<int> order;
  const int depth = sr_build_bvh4(v.data(), n, nodes, order);
  CHECK(depth >= 0, "%s: build failed", name);
  if (n == 0) {
    CHECK(nodes.empty() && order.empty(), "%s: empty scene", name);
    return;
  }
  CHECK(3 * depth + 1 <= SR_STACK_SIZE, "%s: depth %d exceeds the traversal stack", name, depth);
  CHECK((int)order.size() == n, "%s: %zu faces in leaf order, expected %d", name, order.size(), n);
  std::vector<int> seen(n, 0);
  for (int f : order) {
    CHECK(f >= 0 && f < n, "%s: face %d out of range", name, f);
    seen[f]++;
  }
  for (int f = 0; f < n; ++f) CHECK(seen[f] == 1, "%s: face %d appears %d times", name, f, seen[f]);
  // walk the tree: refs in range, leaves cover [0, n) exactly once, every
  // face inside each (node, slot) box on its root path
  std::vector<int> covered(n, 0);
  struct Rec {
    const std::vector<Node4>& nodes;
    const std::vector<float>& v;
    const std::vector<int>& order;
    std::vector<int>& covered;
    int n;
    int max_level;
    bool ok = true;
    char msg[256] = {0};
    void walk(int node, int level, std::vector<std::pair<int, int>>& up) {
      if (!ok) return;
      if (node < 0 || node >= (int)nodes.size()) {
        ok = false;
        std::snprintf(msg, sizeof(msg), "node %d out of range", node);
        return;
      }
      max_level = std::max(max_level, level);
      const Node4& nd = nodes[node];
      for (int s = 0; s < 4; ++s) {
        const int c = nd.child[s];
        if (c == SR_EMPTY) continue;
        up.push_back({node, s});
        if (c >= 0) {
          walk(c, level + 1, up);
        } else {
          const uint32_t r = (uint32_t)c;
          const int first = (int)((r & ~SR_LEAF_BIT) >> 3), cnt = (int)(r & 7u);
          if (cnt < 1 || cnt > SR_LEAF_MAX || first < 0 || first + cnt > n) {
            ok = false;
            std::snprintf(msg, sizeof(msg), "leaf (%d, %d) out of range", first, cnt);
            return;
          }
          for (int i = first; i < first + cnt; ++i) {
            covered[i]++;
            const float* tri = v.data() + (size_t)order[i] * 9;
            for (auto& ns : up)
              if (!inside(nodes[ns.first], ns.second, tri)) {
                ok = false;
                std::snprintf(msg, sizeof(msg), "face %d outside the box of node %d slot %d", order[i], ns.first,
                              ns.second);
                return;
              }
          }
        }
        up.pop_back();
      }
    }
  } rec{nodes, v, order, covered, n, 0};
  std::vector<std::pair<int, int>> up;
  rec.walk(0, 1, up);
  CHECK(rec.ok, "%s: %s", name, rec.msg);
  for (int i = 0; i < n; ++i) CHECK(covered[i] == 1, "%s: leaf slot %d covered %d times", name, i, covered[i]);
  CHECK(rec.max_level == depth, "%s: walked depth %d, reported %d", name, rec.max_level, depth);
}

static void check_light_tree(const char* name, int n, int B, std::mt19937& rng) {
  std::uniform_real_distribution<float> U(-1.0f, 1.0f);
  std::vector<float> lv((size_t)n * 12);
  std::vector<int> slots(n), smat(n);
  const int M = 3;
  std::vector<unsigned char> se((size_t)M * B);
  for (int m = 0; m < M; ++m)
    for (int b = 0; b < B; ++b) se[(size_t)m * B + b] = (unsigned char)((m + b) % 3 != 0);
  for (int s = 0; s < n; ++s) {
    for (int k = 0; k < 3; ++k) {
      lv[(size_t)s * 12 + 4 * k + 0] = U(rng);
      lv[(size_t)s * 12 + 4 * k + 1] = U(rng);
      lv[(size_t)s * 12 + 4 * k + 2] = U(rng);
      lv[(size_t)s * 12 + 4 * k + 3] = 0.0f;
    }
    slots[s] = 3 * s + 1;  // face ids, increasing
    smat[s] = s % M;
  }
  std::vector<LightNode> nodes;
  std::vector<float> power;
  std::vector<LightKey> keys;
  const int rc = sr_build_light_tree(n, lv.data(), slots.data(), smat.data(), se.data(), B, nodes, power, keys);
  CHECK(rc == (int)nodes.size(), "%s: build returned %d for %zu nodes", name, rc, nodes.size());
  if (n == 0) {
    CHECK(nodes.empty(), "%s: empty", name);
    return;
  }
  const int nn = (int)nodes.size();
  CHECK(power.size() == (size_t)B * nn, "%s: power table %zu, expected %d", name, power.size(), B * nn);
  CHECK((int)keys.size() == n, "%s: %zu keys for %d emitters", name, keys.size(), n);
  for (int i = 0; i + 1 < (int)keys.size(); ++i) CHECK(keys[i].face < keys[i + 1].face, "%s: keys unsorted", name);
  std::vector<int> leaf_of_slot(n, -1);
  for (int i = 0; i < nn; ++i) {
    const LightNode& d = nodes[i];
    if (i == 1) continue;  // unused slot of the root's pair
    if (d.right < 0) {     // leaf
      const int s = ~d.left;
      CHECK(s >= 0 && s < n, "%s: leaf %d names slot %d", name, i, s);
      CHECK(leaf_of_slot[s] < 0, "%s: slot %d in two leaves", name, s);
      leaf_of_slot[s] = i;
      CHECK(d.face == slots[s], "%s: leaf %d face %d, slot face %d", name, i, d.face, slots[s]);
    } else {
      CHECK(d.left >= 2 && d.left % 2 == 0 && d.right == d.left + 1 && d.right < nn, "%s: node %d children %d %d",
            name, i, d.left, d.right);
      CHECK(nodes[d.left].parent == i && nodes[d.right].parent == i, "%s: node %d parent links", name, i);
      for (int b = 0; b < B; ++b) {
        const float p = power[(size_t)b * nn + i], q = power[(size_t)b * nn + d.left] + power[(size_t)b * nn + d.right];
        CHECK(std::fabs(p - q) <= 1e-5f * std::fabs(p) + 1e-12f, "%s: node %d band %d power %g vs children %g", name,
              i, b, p, q);
      }
    }
  }
  for (int s = 0; s < n; ++s) CHECK(leaf_of_slot[s] >= 0, "%s: slot %d has no leaf", name, s);
  for (const LightKey& k : keys) CHECK(k.leaf >= 0 && k.leaf < nn && nodes[k.leaf].face == k.face, "%s: key", name);
}

static std::vector<float> random_tris(int n, std::mt19937& rng, float size) {
  std::uniform_real_distribution<float> U(-1.0f, 1.0f), S(-size, size);
  std::vector<float> v((size_t)n * 9);
  for (int i = 0; i < n; ++i) {
    float c[3] = {U(rng), U(rng), U(rng)};
    for (int k = 0; k < 3; ++k)
      for (int a = 0; a < 3; ++a) v[(size_t)i * 9 + 3 * k + a] = c[a] + S(rng);
  }
  return v;
}

int main() {
  std::mt19937 rng(12345);
  char name[64];
  check_bvh("empty", {});
  for (int n : {1, 2, 3, 4, 5, 7, 16, 17, 63, 64, 65, 200, 1000, 20000}) {
    std::snprintf(name, sizeof(name), "random %d", n);
    check_bvh(name, random_tris(n, rng, 0.05f));
  }
  {  // every centroid identical: the build can only split by median
    std::vector<float> v = random_tris(3000, rng, 0.0f);
    for (size_t i = 0; i < v.size(); ++i) v[i] = (float)((i % 9) / 3);
    check_bvh("coincident centroids", v);
  }
  {  // degenerate faces (zero area) and a huge one
    std::vector<float> v = random_tris(500, rng, 0.0f);
    for (int k = 0; k < 9; ++k) v[k] = (k % 3 == 0) ? 1e6f * (k / 3 - 1) : 0.0f;
    check_bvh("degenerate", v);
  }
  {  // faces at x = -2^-k: the centroid distribution that exhausts the binary depth budget
    std::vector<float> v;
    for (int k = 0; k < 120; ++k)
      for (int r = 0; r < 40; ++r) {
        const float x = -std::ldexp(1.0f, -k), y = 0.01f * r;
        const float t[9] = {x, y, 0.f, x, y + 0.005f, 0.f, x, y, 0.005f};
        v.insert(v.end(), t, t + 9);
      }
    check_bvh("geometric x", v);
  }
  {  // a grid heightfield (shared edges, many equal bounds)
    const int nx = 64, nz = 32;
    std::vector<float> v;
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < nz; ++j) {
        const float x0 = i / (float)nx, x1 = (i + 1) / (float)nx, z0 = j / (float)nz, z1 = (j + 1) / (float)nz;
        const float a[9] = {x0, 0.f, z0, x1, 0.f, z0, x1, 0.f, z1}, b[9] = {x0, 0.f, z0, x1, 0.f, z1, x0, 0.f, z1};
        v.insert(v.end(), a, a + 9);
        v.insert(v.end(), b, b + 9);
      }
    check_bvh("flat grid", v);
  }
  for (int n : {0, 1, 2, 3, 5, 64, 257, 1282}) {
    std::snprintf(name, sizeof(name), "light tree %d", n);
    check_light_tree(name, n, 7, rng);
  }
  if (g_fail) {
    std::fprintf(stderr, "host_check: %d failure(s)\n", g_fail);
    return 1;
  }
  std::printf("host_check: all builder checks passed\n");
  return 0;
}
