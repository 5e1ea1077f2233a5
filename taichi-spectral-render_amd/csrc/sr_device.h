// sr_device.h -- device functions of the spectral path tracer (gfx950).
//
// The shading functions restate Xyzonox/taichi-spectral-render main.py
// (each cites its lines).  tm.min/tm.max map to fminf/fmaxf (NaN-ignoring,
// as LLVM minnum/maxnum).  The ray/triangle test is compiled without FMA
// contraction and with IEEE division so closest-hit t is bit-identical to
// an un-contracted f32 evaluation on the host.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sr_layout.h"

#define SR_PI 3.14159265358979323846f
#define SR_EPSILON 1e-4f  // main.py:19
#define SR_ESCAPE 1e2f    // main.py:20
#define SR_SCATTER_BOUNCES 5  // main.py:23

// ------------------------------------------------------------------ vec3
struct f3 {
  float x, y, z;
};
__device__ __forceinline__ f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ f3 operator-(f3 a) { return mk(-a.x, -a.y, -a.z); }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ f3 operator*(float s, f3 a) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ f3 operator/(f3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
  return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float length(f3 a) { return sqrtf(dot(a, a)); }
// Shading arithmetic whose results feed no exact identity: the hardware
// reciprocal, square root and reciprocal square root (about 1 ulp each;
// rcp(0) = inf and rcp(inf) = 0, so 0/0 and x/0 keep their IEEE outcomes).
// IEEE division and sqrt stay where the reference's exact cases live: the
// ray/triangle test, refract (eta = 1 must return the ray), the relative
// index n_i / n_t and the emitter estimator atten / (1 + 1e-6).
__device__ __forceinline__ float frcp(float b) { return __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ float fdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ f3 normalize(f3 a) { return a * __builtin_amdgcn_rsqf(dot(a, a)); }
#define SR_INV_PI 0.318309886f
__device__ __forceinline__ f3 reflect(f3 x, f3 n) { return x - 2.0f * dot(n, x) * n; }
__device__ __forceinline__ f3 refract(f3 x, f3 n, float eta) {
  float dxn = dot(x, n);
  float k = 1.0f - eta * eta * (1.0f - dxn * dxn);
  f3 r = mk(0.f, 0.f, 0.f);
  if (k >= 0.0f) r = eta * x - (eta * dxn + sqrtf(k)) * n;
  return r;
}
__device__ __forceinline__ float tmax(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ float tmin(float a, float b) { return fminf(a, b); }
__device__ __forceinline__ float clampf(float x, float lo, float hi) { return tmin(tmax(x, lo), hi); }
__device__ __forceinline__ f3 xyz(float4 v) { return mk(v.x, v.y, v.z); }

// ------------------------------------------------------------------ RNG
// Philox-4x32-10 (Salmon et al., SC'11), standing in for ti.random().  A
// path's uniforms come in fixed blocks: block (segment k, slot s) is
// Philox(ctr = {pixel, block, band | stream << 8, k << 3 | s}, key = seed),
// its four words mapped to [0, 1) with 24 bits each.  Each sampling step of
// a segment owns one slot (SR_SLOT_*), so a handler generates its block once,
// at one place in the code, for every lane that reaches it.
__device__ __forceinline__ void philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1,
                                       uint32_t out[4]) {
  // Each round's two 32 x 32 -> 64-bit products are formed whole (one
  // v_mad_u64_u32 each) rather than as separate high and low multiplies:
  // the same bits for half the quarter-rate multiplies.
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

enum {
  SR_SLOT_EXT = 0,   // scatter_volume: free flight, event choice, 2 for HG or EEM
  SR_SLOT_NEE = 1,   // sample_direct_light: light pick, point on the light (2)
  SR_SLOT_SCAT = 2,  // scatter_surface: GGX half vector (2), Fresnel choice, subsurface distance
  SR_SLOT_OPQ = 3,   // scatter_surface, opaque: cosine direction (2), surface EEM (2)
  SR_SLOT_AUX = 4,   // camera jitter (segment 0, words 0-1); fluorescence direction (words 2-3)
  // subsurface random walk (SR_RENDER_SUBSURFACE), counter seg << 9 | step:
  SR_SLOT_WALK = 5,      // free flight, HG direction (2)
  SR_SLOT_WALK_NEE = 6,  // NEE at the exit point
  SR_SLOT_WALK_EXIT = 7  // cosine exit direction (2)
};
struct U4 {
  float u[4];
};
__device__ __forceinline__ float u24(uint32_t w) { return (float)(w >> 8) * (1.0f / 16777216.0f); }
__device__ __forceinline__ U4 draw4(uint32_t pixel, uint32_t block, uint32_t band_stream, uint32_t seg, uint32_t slot,
                                    uint32_t k0, uint32_t k1) {
  uint32_t o[4];
  philox(pixel, block, band_stream, (seg << 3) | slot, k0, k1, o);
  return U4{{u24(o[0]), u24(o[1]), u24(o[2]), u24(o[3])}};
}

// ------------------------------------------------------------------ scene view
struct SceneDev {
  const Node4* nodes;
  const float4* tris;   // [n_tris][3]
  const float4* shade;  // [n_tris][4], original face order
  const MatRec* mats;   // [M][B]
  const int* mat_eem;   // [M]
  const float* eem;     // [E][B][B]
  const LightRec* lights;  // [B][max_lights]
  const float4* lverts;    // [slots][3]
  const int* light_count;  // [B]
  const float* light_area; // [B]
  const float* lam_m;      // [B] band wavelength in metres (f32 as the reference forms it)
  const float* lam_nm;     // [B] WAVELENGTH_MIN + idx * STEP
  const LightNode* lnodes; // light tree in sibling pairs (root 0, node 1 unused); many-light NEE only
  const float* lpower;     // [B][n_lnodes] emitting area per band
  const LightKey* lkeys;   // [n_lkeys] emissive face -> leaf, sorted by face
  int n_lkeys, n_lnodes;
  int B, max_lights, M, n_tris;
  float cam_o[3], cam_u[3], cam_v[3], cam_w[3];
};

__device__ __forceinline__ MatRec get_mat(const SceneDev& s, int m, int b) { return s.mats[m * s.B + b]; }

// ------------------------------------------------------------------ traversal
struct HitD {
  float t;
  int prim;  // original face id, -1 on miss
  float u, v;
};

// Moller-Trumbore exactly as intersect_triangle (main.py:807-829), with the
// edges precomputed in f32 (bit-identical to v1 - v0 at trace time).  The
// candidate wins on smaller t, or equal t and smaller face id.
__device__ __forceinline__ void intersect_tri(f3 o, f3 d, float4 a, float4 b, float4 c, float& best_t, int& best_prim,
                                              float& bu, float& bv) {
#pragma clang fp contract(off)
  const f3 v0 = xyz(a), e1 = xyz(b), e2 = xyz(c);
  const int prim = __float_as_int(a.w);
  f3 pvec = cross(d, e2);
  float det = dot(e1, pvec);
  float inv_det = 1.0f / det;
  f3 tvec = o - v0;
  float u = dot(tvec, pvec) * inv_det;
  f3 qvec = cross(tvec, e1);
  float v = dot(d, qvec) * inv_det;
  float t = dot(e2, qvec) * inv_det;
  // bitwise, not short-circuit, tests and a select: no branch
  const bool ok = (fabsf(det) > 1e-8f) & (u >= 0.0f) & (u <= 1.0f) & (v >= 0.0f) & (u + v <= 1.0f) &
                  (t > SR_EPSILON) & ((t < best_t) | ((t == best_t) & (prim < best_prim)));
  best_t = ok ? t : best_t;
  best_prim = ok ? prim : best_prim;
  bu = ok ? u : bu;
  bv = ok ? v : bv;
}

// Entry i (0 <= i < SR_STACK_SIZE) of a lane's LDS stack column.
__device__ __forceinline__ int& stk_at(int* stk, int stride, int i) { return stk[i * stride]; }

// Tests the 4 children of node `cur` against the ray; returns the nearest
// hit child in `next` (SR_EMPTY when none) and pushes the others far-to-near.
// The near and far planes of each axis are fetched by address -- `ox`, `oy`,
// `oz` are the byte offsets of the near plane's float4 in the node line
// (lo_x at 0, hi_x at 16, lo_y at 32, ...: hi when the ray's direction is
// negative on that axis; the far plane is the other one, offset ^ 16) -- so
// no per-child selects are needed.  The child order
// comes from a branch-free sorting network; the result does not depend on
// the order (ties in t go to the smaller face id).
__device__ __forceinline__ void visit_node(const Node4* __restrict__ nodes, int cur, f3 inv, f3 oi, bool nx, bool ny,
                                           bool nz, float tmax_, int* stk, int stride, int& sp, int& next,
                                           bool& ovf) {
  // 32-bit byte offsets from the array base (kept in scalar registers): one
  // VGPR per address instead of a 64-bit pointer
  const char* base = reinterpret_cast<const char*>(nodes);
  const uint32_t nb = (uint32_t)cur * (uint32_t)sizeof(Node4);
  const uint32_t ox = nb + (nx ? 16u : 0u), oy = nb + (ny ? 48u : 32u), oz = nb + (nz ? 80u : 64u);
#define SR_LD4(off) (*reinterpret_cast<const float4*>(base + (size_t)(uint32_t)(off)))
  const float4 nxv = SR_LD4(ox), fxv = SR_LD4(ox ^ 16u);
  const float4 nyv = SR_LD4(oy), fyv = SR_LD4(oy ^ 16u);
  const float4 nzv = SR_LD4(oz), fzv = SR_LD4(oz ^ 16u);
#undef SR_LD4
  const int4 ch = *reinterpret_cast<const int4*>(base + (size_t)(nb + 96u));
  float tn[4], tf[4];
  const float* pnx = &nxv.x; const float* pfx = &fxv.x;
  const float* pny = &nyv.x; const float* pfy = &fyv.x;
  const float* pnz = &nzv.x; const float* pfz = &fzv.x;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    float a0 = fmaf(pnx[c], inv.x, -oi.x), a1 = fmaf(pny[c], inv.y, -oi.y), a2 = fmaf(pnz[c], inv.z, -oi.z);
    float b0 = fmaf(pfx[c], inv.x, -oi.x), b1 = fmaf(pfy[c], inv.y, -oi.y), b2 = fmaf(pfz[c], inv.z, -oi.z);
    tn[c] = fmaxf(fmaxf(a0, a1), fmaxf(a2, 0.0f));
    tf[c] = fminf(fminf(b0, b1), fminf(b2, tmax_));
  }
  int cref[4] = {ch.x, ch.y, ch.z, ch.w};
  float key[4];
  int hitc = 0;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    // conservative slack for the fma-rounded slab distances
    bool hit = tn[c] <= fmaf(tf[c], 1.0000005f, 1e-7f) && cref[c] != SR_EMPTY;
    key[c] = hit ? tn[c] : INFINITY;
    hitc += hit ? 1 : 0;
  }
#define SR_CSWAP(i, j)                                                   \
  {                                                                      \
    const bool sw = key[j] < key[i];                                     \
    const float ki = key[i], kj = key[j];                                \
    const int ri = cref[i], rj = cref[j];                                \
    key[i] = sw ? kj : ki; key[j] = sw ? ki : kj;                        \
    cref[i] = sw ? rj : ri; cref[j] = sw ? ri : rj;                      \
  }
  SR_CSWAP(0, 1) SR_CSWAP(2, 3) SR_CSWAP(0, 2) SR_CSWAP(1, 3) SR_CSWAP(1, 2)
#undef SR_CSWAP
  next = hitc > 0 ? cref[0] : SR_EMPTY;
  // Push the other hits far-to-near with three unconditional writes to
  // slots sp, sp + 1, sp + 2: slot sp + k takes the (np - k)-th nearest of
  // them while k < np, so the nearest ends on top; the slots at and above
  // the new top take copies that are never read.  No branch, whatever the
  // lane's hit count.  The builder bounds the depth so that a visit starts
  // with sp <= SR_STACK_SIZE - 3 (three entries per level above it); a
  // deeper stack is flagged (never written out of the lane's column).
  const int np = max(hitc - 1, 0);
  ovf = ovf || sp > SR_STACK_SIZE - 3;
  const int s0 = min(sp, SR_STACK_SIZE - 3);
  const int w0 = np == 3 ? cref[3] : (np == 2 ? cref[2] : cref[1]);
  const int w1 = np == 3 ? cref[2] : cref[1];
  stk_at(stk, stride, s0) = w0;
  stk_at(stk, stride, s0 + 1) = w1;
  stk_at(stk, stride, s0 + 2) = cref[1];
  sp = s0 + np;
}

// All SR_LEAF_MAX (4) triangle records from the leaf's first are fetched
// and tested, whatever the leaf's count: a leaf costs one memory round trip
// and no branch.  Records past the leaf's own faces are faces of the scene
// (the next leaves' in leaf order) or the SR_LEAF_MAX - 1 zero records at
// the end of the array (det = 0: never a hit), so they cannot change the
// walk's answer: a hit on such a face lies inside that face's own leaf box,
// which the walk visits unless a closer hit culls it.  The result does not
// depend on the test order (ties go to the smaller face id).
static_assert(SR_LEAF_MAX == 4, "visit_leaf reads four triangle records");
__device__ __forceinline__ void visit_leaf(const float4* __restrict__ tris, int ref, f3 o, f3 d, HitD& h) {
  const uint32_t r = (uint32_t)ref;
  const int first = (int)((r & ~SR_LEAF_BIT) >> 3);
  const float4* tp = tris + (size_t)first * 3;
  const float4 a0 = tp[0], b0 = tp[1], c0 = tp[2], a1 = tp[3], b1 = tp[4], c1 = tp[5];
  const float4 a2 = tp[6], b2 = tp[7], c2 = tp[8], a3 = tp[9], b3 = tp[10], c3 = tp[11];
  intersect_tri(o, d, a0, b0, c0, h.t, h.prim, h.u, h.v);
  intersect_tri(o, d, a1, b1, c1, h.t, h.prim, h.u, h.v);
  intersect_tri(o, d, a2, b2, c2, h.t, h.prim, h.u, h.v);
  intersect_tri(o, d, a3, b3, c3, h.t, h.prim, h.u, h.v);
}

// Closest hit over the 4-wide BVH, as a walk that can be suspended.  `stk`
// is this lane's LDS stack column (entry i at stk[i * stride]).  Children
// are visited nearest-first; boxes beyond the current best t are culled
// (ties kept: <=).
//
// Loop structure (while-while with speculative traversal): a lane walks inner
// nodes until it meets a leaf, parks that leaf and keeps walking nodes until
// every lane of the wave has a leaf (or is finished); then all lanes
// intersect their leaves together.  This keeps the wave in one of the two
// loop bodies instead of splitting it between nodes and leaves each step.
// `nodes` / `tris` are passed by value so they sit in scalar registers for
// the whole traversal (no reloads inside the loops).  Only hits with
// t < t_init are reported (t_init = 1e9, the reference's miss distance,
// for a plain closest-hit query); with no such hit the result is
// {t_init, -1}.
//
// t_near < t_init asks a narrower question: the closest hit with
// t <= t_near exactly, and otherwise only whether any hit with t < t_init
// exists (then the result is {nextafter(t_near), -2}).  This is what a
// path segment in a medium needs once its free-flight distance d is drawn:
// a surface farther than d + EPSILON is never reached, but whether one
// exists before ESCAPE still decides between a volume event and an escape
// (main.py:1210-1218).  The first far hit found shrinks the search to
// [0, t_near], so a dense medium's traversal stops almost at once.
//
// A walk is the state below, started by trace_begin and advanced by
// trace_resume, which the path loop calls once per iteration for every
// lane whose ray is unfinished: a wave walks only while enough of its lanes
// still walk, and the lanes whose rays ended meanwhile go on to shade and
// start new rays that join the walk (the lanes' ballots decide when to
// stop, so the wave's walks are compacted in time rather than in memory).
// The state is one lane's: next node, parked leaf, stack depth, best hit
// so far and the exact-hit limit.
struct TraceState {
  int cur, leaf, sp;
  HitD h;
  float t_near;
  bool far;  // a hit in (t_near, t_init) was seen
};

// The path kernels suspend a wave's walks once fewer than this many of its
// lanes still walk (32 of 16-48 measured best on the 1M-triangle heightfield
// and the volume scene; 16-20 clearly worse).
#ifndef SR_TRACE_KEEP
#define SR_TRACE_KEEP 32
#endif

__device__ __forceinline__ void trace_begin(TraceState& s, float t_init, float t_near) {
  s.cur = 0;  // root node; SR_EMPTY = nothing left
  s.leaf = SR_EMPTY;
  s.sp = 0;
  s.h.t = t_init;  // HitRecord(t=1e9), main.py:833
  s.h.prim = INT32_MIN;
  s.h.u = 0.f;
  s.h.v = 0.f;
  s.t_near = t_near;
  s.far = false;
}

// Walks the rays of the lanes with `live` set, in rounds of the
// while-while loop below, until fewer than min(keep, walking lanes at entry)
// lanes of the wave are still walking (keep <= 1: every walk ends).  A lane
// whose walk ended has its final hit in s.h and returns true; the others
// keep their state (and their stack column) for the next call.
__device__ __forceinline__ bool trace_resume(const Node4* __restrict__ nodes, const float4* __restrict__ tris, f3 o,
                                             f3 d, int* stk, int stride, uint32_t* overflow, TraceState& s, bool live,
                                             int keep) {
  const float tiny = 1e-12f;
  f3 dd = mk(fabsf(d.x) < tiny ? copysignf(tiny, d.x) : d.x, fabsf(d.y) < tiny ? copysignf(tiny, d.y) : d.y,
             fabsf(d.z) < tiny ? copysignf(tiny, d.z) : d.z);
  f3 inv = mk(1.0f / dd.x, 1.0f / dd.y, 1.0f / dd.z);
  f3 oi = mk(o.x * inv.x, o.y * inv.y, o.z * inv.z);
  // the near plane of an axis is the box's hi side when the direction is negative
  const bool nx = inv.x < 0.f, ny = inv.y < 0.f, nz = inv.z < 0.f;
  int cur = live ? s.cur : SR_EMPTY;
  int leaf = live ? s.leaf : SR_EMPTY;
  int sp = s.sp;
  HitD h = s.h;
  bool far = s.far;
  const float t_near = s.t_near;
  const float t_near_up = nextafterf(t_near, INFINITY);
  bool walking = cur != SR_EMPTY || leaf != SR_EMPTY;
  bool ovf = false;  // the lane's stack outgrew its column (flagged once, after the rounds)
  const int n0 = __popcll(__ballot(walking));
  const int stop = min(keep, n0);  // leave once fewer than this still walk
  int n = n0;
  while (n > 0) {
    // node loop
    while (cur != SR_EMPTY && cur >= 0) {
      int next;
      visit_node(nodes, cur, inv, oi, nx, ny, nz, h.t, stk, stride, sp, next, ovf);
      // The two entries below the top are read unconditionally (a slot
      // below 0 reads slot 0 and is not used): with no child to go on with,
      // the walk pops the top; when the node it goes on with is the first
      // leaf, it parks that leaf and pops (again).
      const int t1 = stk_at(stk, stride, max(sp - 1, 0)), t2 = stk_at(stk, stride, max(sp - 2, 0));
      const bool popped = next == SR_EMPTY;
      int s1 = sp;
      if (popped) {
        next = sp > 0 ? t1 : SR_EMPTY;
        s1 = max(sp - 1, 0);
      }
      cur = next;
      if (cur != SR_EMPTY && cur < 0 && leaf == SR_EMPTY) {  // park the first leaf, keep walking
        leaf = cur;
        cur = s1 > 0 ? (popped ? t2 : t1) : SR_EMPTY;
        s1 = max(s1 - 1, 0);
      }
      sp = s1;
      const bool waiting = cur != SR_EMPTY && cur >= 0 && leaf == SR_EMPTY;
      if (__ballot(waiting) == 0ull) break;
    }
    // leaf loop
    while (leaf != SR_EMPTY) {
      visit_leaf(tris, leaf, o, d, h);
      // a far hit: from now on only t <= t_near counts
      const bool fh = h.t > t_near && h.prim != INT32_MIN;
      far = far || fh;
      h.t = fh ? t_near_up : h.t;
      h.prim = fh ? INT32_MIN : h.prim;
      // the node walk stopped at a second leaf: test it next and pop
      // (SR_EMPTY is negative, so cur < 0 with cur != SR_EMPTY is a leaf)
      const int top = stk_at(stk, stride, max(sp - 1, 0));
      const bool again = cur != SR_EMPTY && cur < 0;
      leaf = again ? cur : SR_EMPTY;
      cur = again ? (sp > 0 ? top : SR_EMPTY) : cur;
      sp = again ? max(sp - 1, 0) : sp;
    }
    walking = cur != SR_EMPTY || leaf != SR_EMPTY;
    n = __popcll(__ballot(walking));
    if (n < stop) break;
  }
  if (ovf) atomicOr(overflow, 1u);
  if (live) {
    s.cur = cur;
    s.leaf = leaf;
    s.sp = sp;
    s.far = far;
    if (!walking && h.prim == INT32_MIN) h.prim = far ? -2 : -1;
    s.h = h;
  }
  return live && !walking;
}

// The closest hit, walked to the end in one call.
__device__ __forceinline__ HitD trace_closest(const Node4* __restrict__ nodes, const float4* __restrict__ tris, f3 o,
                                              f3 d, int* stk, int stride, uint32_t* overflow, float t_init = 1e9f,
                                              float t_near = INFINITY) {
  TraceState s;
  trace_begin(s, t_init, t_near);
  trace_resume(nodes, tris, o, d, stk, stride, overflow, s, true, 0);
  return s.h;
}

// HitRecord completion (main.py:860-863) from the face's shading record.
struct SurfHit {
  f3 normal, sn;  // geometric normal, shading normal (unflipped)
  int mat, front;
};
__device__ __forceinline__ SurfHit surface_of(const SceneDev& s, const HitD& h, f3 d) {
  SurfHit r;
  const float4* sh = s.shade + (size_t)h.prim * 4;
  float4 a = sh[0], b = sh[1], c = sh[2], e = sh[3];
  r.normal = xyz(a);
  r.mat = __float_as_int(b.w);
  r.sn = normalize((1.0f - h.u - h.v) * xyz(b) + h.u * xyz(c) + h.v * xyz(e));
  r.front = dot(d, r.normal) < 0.0f ? 1 : 0;
  return r;
}

// ------------------------------------------------------------------ sampling (main.py:682-792)
// sin(2 pi u) and cos(2 pi u).  The turn fraction is reduced exactly:
// r = u - rint(u), quadrant q = rint(4 r), f = r - q/4 in [-1/8, 1/8]; both
// functions are then the Taylor series of sin(2 pi f) (to f^9) and
// cos(2 pi f) (to f^10) -- truncation below 2e-9 on that interval, about
// 2 ulp after rounding -- rotated by q quarter turns.  The sampling angles
// of the reference are all 2 pi * uniform, so this replaces sinf/cosf of
// the rounded product with the same functions of the exact turn count.
__device__ __forceinline__ void sincos_turns(float u, float& s, float& c) {
  const float r = u - rintf(u);
  const float q = rintf(4.0f * r);
  const float f = fmaf(-0.25f, q, r);
  const float f2 = f * f;
  const float ps = fmaf(fmaf(fmaf(fmaf(42.058693f, f2, -76.70586f), f2, 81.60525f), f2, -41.3417f), f2, 6.2831855f) * f;
  const float pc =
      fmaf(fmaf(fmaf(fmaf(fmaf(-26.426256f, f2, 60.24464f), f2, -85.45682f), f2, 64.93939f), f2, -19.739208f), f2, 1.0f);
  const int qi = ((int)q) & 3;
  const float sa = (qi & 1) ? pc : ps;
  const float ca = (qi & 1) ? ps : pc;
  s = (qi & 2) ? -sa : sa;
  c = ((qi + 1) & 2) ? -ca : ca;
}
__device__ __forceinline__ void onb(f3 n, f3& t, f3& b) {
  f3 up = fabsf(n.y) < 0.999f ? mk(0, 1, 0) : mk(1, 0, 0);
  t = normalize(cross(up, n));
  b = cross(n, t);
}
__device__ __forceinline__ f3 random_in_unit_sphere(float u0, float u1) {
  float z = u0 * 2.0f - 1.0f;
  float rxy = fsqrt(tmax(0.0f, 1.0f - z * z));
  float sp, cp;
  sincos_turns(u1, sp, cp);  // phi = 2 pi u1
  return mk(rxy * cp, rxy * sp, z);
}
__device__ __forceinline__ f3 cosine_hemisphere(f3 n, float r1, float r2) {
  float sp, cp;
  sincos_turns(r1, sp, cp);  // phi = 2 pi r1
  float ct = fsqrt(1.0f - r2);
  float st = fsqrt(r2);
  f3 t, b;
  onb(n, t, b);
  return t * cp * st + b * sp * st + n * ct;
}
__device__ __forceinline__ float hg_eval(float c, float g) {
  float g2 = g * g;
  float den = 1.0f + g2 - 2.0f * g * c;
  return fdiv(1.0f - g2, 4.0f * SR_PI * den * fsqrt(den) + 1e-6f);
}
__device__ __forceinline__ f3 hg_sample(f3 w, float g, float u0, float u1) {
  float ct;
  if (fabsf(g) < 1e-4f) {
    ct = 2.0f * u0 - 1.0f;
  } else {
    float rs = fdiv(1.0f - g * g, 1.0f - g + 2.0f * g * u0);
    ct = fdiv(1.0f + g * g - rs * rs, 2.0f * g);
  }
  float st = fsqrt(tmax(0.0f, 1.0f - ct * ct));
  float sp, cp;
  sincos_turns(u1, sp, cp);  // phi = 2 pi u1
  f3 u, v;
  onb(w, u, v);
  return u * cp * st + v * sp * st + w * ct;
}

// ------------------------------------------------------------------ microfacet / Fresnel (main.py:866-978)
struct cplx {
  float re, im;
};
__device__ __forceinline__ cplx cmul(cplx a, cplx b) { return cplx{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ cplx cdiv(cplx a, cplx b) {
  float m = b.re * b.re + b.im * b.im;
  cplx p = cmul(a, cplx{b.re, -b.im});
  const float rm = frcp(m);
  return cplx{p.re * rm, p.im * rm};
}
// Principal square root as complex_sqrt (main.py:884-896).  With a
// correctly rounded sqrt mod >= |re|, so both radicands are >= 0; the clamps
// only guard against a less exact sqrt and change no correctly rounded result.
__device__ __forceinline__ cplx csqrt_ref(cplx z) {
  float mod = fsqrt(z.re * z.re + z.im * z.im);
  float re = fsqrt(tmax(0.0f, (mod + z.re) * 0.5f));
  float sg = z.im > 0.f ? 1.f : (z.im < 0.f ? -1.f : 0.f);
  float im = sg * fsqrt(tmax(0.0f, (mod - z.re) * 0.5f));
  if (z.im == 0.0f) im = 0.0f;
  return cplx{re, im};
}
__device__ __forceinline__ float ggx_D(float nh, float r) {
  float alpha = r * r;
  float a2 = alpha * alpha;
  float nh2 = nh * nh;
  float den = nh2 * (a2 - 1.0f) + 1.0f;
  return fdiv(a2, SR_PI * den * den);
}
__device__ __forceinline__ float ggx_G(f3 n, f3 v, f3 l, float r) {
  float nv = tmax(dot(n, v), 0.0f);
  float alpha = tmax(r * r, 1e-4f);
  float k = alpha * alpha * 0.5f;
  float g1v = fdiv(nv, nv * (1.0f - k) + k + 1e-7f);
  float nl = tmax(dot(n, l), 0.0f);
  float g1l = fdiv(nl, nl * (1.0f - k) + k + 1e-7f);
  return g1v * g1l;
}
__device__ __forceinline__ float fresnel_spectral(float ci, float ni, float ki, float nt, float kt) {
  float c = clampf(ci, -1.0f, 1.0f);
  cplx ein{ni * ni - ki * ki, 2 * ni * ki};
  cplx etr{nt * nt - kt * kt, 2 * nt * kt};
  // the reference forms cos and sin^2 of acos(c) (main.py:929-943): c, 1 - c^2
  float ct = c;
  float s2 = 1.0f - c * c;
  cplx left = cmul(csqrt_ref(ein), cplx{ct, 0.0f});
  cplx em = cmul(ein, cplx{s2, 0.0f});
  cplx right = csqrt_ref(cplx{etr.re - em.re, etr.im - em.im});
  cplx rs = cdiv(cplx{left.re - right.re, left.im - right.im}, cplx{left.re + right.re, left.im + right.im});
  cplx l2 = cmul(etr, left);
  cplx r2 = cmul(ein, right);
  cplx rp = cdiv(cplx{l2.re - r2.re, l2.im - r2.im}, cplx{l2.re + r2.re, l2.im + r2.im});
  return 0.5f * ((rs.re * rs.re + rs.im * rs.im) + (rp.re * rp.re + rp.im * rp.im));
}
__device__ __forceinline__ f3 ggx_sample_half(float x0, float x1, f3 n, float alpha) {
  float a2 = alpha * alpha;
  float sp, cp;
  sincos_turns(x0, sp, cp);  // phi = 2 pi x0
  float ct = fsqrt(fdiv(1.0f - x1, 1.0f + (a2 - 1.0f) * x1));
  float st = fsqrt(tmax(0.0f, 1.0f - ct * ct));  // ct may round one ulp above 1
  f3 t, b;
  onb(n, t, b);
  return normalize(t * (cp * st) + b * (sp * st) + n * ct);
}
// cos(phi_i - phi_o) * sin(alpha) * tan(beta) of oren_nayar (main.py:1071-1088),
// alpha = max(acos nl, acos nv), beta = min(...).  The larger angle has the
// smaller cosine, so sin(alpha) = sqrt(1 - cmin^2), tan(beta) =
// sqrt(1 - cmax^2) / cmax.  A cosine above 1 (acos -> NaN) drops out of the
// max/min as NaN operands do there; both above 1 gives NaN.  The term is
// only needed when both cosines exceed 1e-4 (otherwise cos(phi) = 0).
__device__ __forceinline__ float oren_nayar_angular(float nl, float nv, f3 l, f3 v, f3 n) {
  if (!(nl > 1e-4f && nv > 1e-4f)) return 0.0f;
  f3 pi = normalize(l - n * nl);
  f3 po = normalize(v - n * nv);
  const float cpd = tmax(0.0f, dot(pi, po));
  const float ca = nl > 1.0f ? nv : nl, cb = nv > 1.0f ? nl : nv;
  const float cmin = tmin(ca, cb), cmax = tmax(ca, cb);
  return cpd * fdiv(fsqrt(1.0f - cmin * cmin) * fsqrt(1.0f - cmax * cmax), cmax);
}
__device__ __forceinline__ float oren_nayar(float r, float nl, float nv, f3 l, f3 v, f3 n) {
  float sigma = r * SR_PI * 0.5f;
  float s2 = sigma * sigma;
  float A = 1.0f - fdiv(s2, 2.0f * (s2 + 0.33f));
  float B = fdiv(0.45f * s2, s2 + 0.09f);
  return A + B * oren_nayar_angular(nl, nv, l, v, n);
}
