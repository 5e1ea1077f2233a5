// sr_layout.h -- device-memory layout of a compiled scene (host and device).
//
// HBM layout (all arrays 16-byte aligned, one allocation each):
//   nodes   Node4[n_nodes]      128 B lines: SoA child boxes + child refs
//   tris    TriRec[n_tris + 3]  48 B: v0|prim, e1, e2 in BVH leaf order, then
//                               SR_LEAF_MAX - 1 zero records (a leaf's four
//                               records are read at once)
//   shade   ShadeRec[n_tris]    64 B: face normal|area, n0|mat, n1, n2,
//                               indexed by ORIGINAL face id (fetched once
//                               per closest hit, never during traversal)
//   mats    MatRec[M][B]        32 B per (material, band)
//   lights  LightRec[B][Lmax]   per-band area CDF over emissive faces
//   lverts  float4[n_slots][3]  vertices of emissive faces (light slots)
#pragma once
#include <stdint.h>

#define SR_STACK_SIZE 40     // per-lane LDS traversal stack entries
#define SR_MAX_BVH4_DEPTH 13 // builder guarantee: 3 * depth + 1 <= stack
#define SR_LEAF_MAX 4        // triangles per leaf (the traversal reads 4 records per leaf)
#define SR_MAX_BANDS 64

// Child reference in Node4::child: >= 0 is an inner node index; with the
// sign bit set it is a leaf: bits 3..30 = first TriRec, bits 0..2 = count
// (1..SR_LEAF_MAX, never 0, so no leaf equals SR_EMPTY).  An empty slot has
// child = SR_EMPTY and an inverted (never-hit) box.
#define SR_EMPTY ((int)0x80000000)
#define SR_LEAF_BIT 0x80000000u

struct alignas(16) Node4 {
  float lo_x[4], hi_x[4], lo_y[4], hi_y[4], lo_z[4], hi_z[4];
  int child[4];
  int pad[4];
};
static_assert(sizeof(Node4) == 128, "Node4 must be one 128-byte line");

struct alignas(16) MatRec {  // one (material, band) of get_mat, main.py:980-996
  float n, k, r, ss, em, g;
  int tv, se;
};
static_assert(sizeof(MatRec) == 32, "MatRec is 32 bytes");

struct alignas(16) LightRec {
  float cdf;
  int prim;  // original face id
  int slot;  // index into lverts
  int pad;
};

// Light tree node (many-light NEE, SR_RENDER_LIGHT_TREE): bounding sphere
// of the emitters below (the sphere around their bounding box: centre and
// squared half diagonal, both in f32 as the kernel would form them, and the
// radius) and their normal cone -- axis and the cosine and sine of its half
// angle widened by a safety margin (cos_o = -1: every direction).  Inner:
// the children at left and left + 1 = right; leaf: left = ~slot,
// right = -1, face = its original face id.  The first 48 bytes are what the
// descent reads.  Per-band emitting area in a separate [B][nodes] array.
struct alignas(16) LightNode {
  float c[3];
  float r2;
  float axis[3];
  float cos_o;
  float sin_o;
  float r;
  int left;
  int right;
  int parent;  // -1 at the root
  int face;    // leaf: original face id; inner: -1
  int pad[2];
};
static_assert(sizeof(LightNode) == 64, "LightNode is 64 bytes");

struct LightKey {  // emissive face -> its leaf, sorted by face
  int face, leaf;
};

static inline int sr_leaf_ref(uint32_t first, uint32_t count) {
  return (int)(SR_LEAF_BIT | (first << 3) | count);
}
