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
__device__ __forceinline__ f3 normalize(f3 a) { return a / length(a); }
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
// Philox-4x32-10 (Salmon et al., SC'11).  Path stream: uniform n is word
// (n & 3) of Philox(ctr = {c0, c1, c2, n >> 2}, key = seed), mapped to
// [0, 1) with 24 bits.  Stands in for ti.random().
__device__ __forceinline__ void philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1,
                                       uint32_t out[4]) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
    uint32_t hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

struct Rng {
  uint32_t k0, k1, c0, c1, c2, n;
  uint32_t b0, b1, b2, b3;
  __device__ __forceinline__ void init(uint32_t key0, uint32_t key1, uint32_t a, uint32_t b, uint32_t c) {
    k0 = key0; k1 = key1; c0 = a; c1 = b; c2 = c; n = 0;
  }
  __device__ __forceinline__ float next() {
    if ((n & 3u) == 0u) {
      uint32_t o[4];
      philox(c0, c1, c2, n >> 2, k0, k1, o);
      b0 = o[0]; b1 = o[1]; b2 = o[2]; b3 = o[3];
    }
    uint32_t w = n & 3u;
    uint32_t x = w == 0 ? b0 : (w == 1 ? b1 : (w == 2 ? b2 : b3));
    ++n;
    return (float)(x >> 8) * (1.0f / 16777216.0f);
  }
};

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
  bool ok = fabsf(det) > 1e-8f && u >= 0.0f && u <= 1.0f && v >= 0.0f && u + v <= 1.0f && t > SR_EPSILON;
  ok = ok && (t < best_t || (t == best_t && prim < best_prim));
  if (ok) { best_t = t; best_prim = prim; bu = u; bv = v; }
}

// Tests the 4 children of node `cur` against the ray; returns the nearest
// hit child in `next` (SR_EMPTY when none) and pushes the others far-to-near.
__device__ __forceinline__ void visit_node(const Node4* __restrict__ nodes, int cur, f3 inv, f3 oi, bool nx, bool ny,
                                           bool nz, float tmax_, int* stk, int stride, int& sp, int& next,
                                           uint32_t* overflow) {
  const float4* np = reinterpret_cast<const float4*>(nodes + cur);
  float4 lx = np[0], hx = np[1], ly = np[2], hy = np[3], lz = np[4], hz = np[5];
  int4 ch = *reinterpret_cast<const int4*>(np + 6);
  float4 nxv = nx ? hx : lx, fxv = nx ? lx : hx;
  float4 nyv = ny ? hy : ly, fyv = ny ? ly : hy;
  float4 nzv = nz ? hz : lz, fzv = nz ? lz : hz;
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
    bool hit = tn[c] <= tf[c] * 1.0000005f + 1e-7f && cref[c] != SR_EMPTY;
    key[c] = hit ? tn[c] : INFINITY;
    hitc += hit ? 1 : 0;
  }
  if (hitc == 0) {
    next = SR_EMPTY;
    return;
  }
#define SR_CSWAP(i, j)                                                   \
  if (key[j] < key[i]) {                                                 \
    float tk = key[i]; key[i] = key[j]; key[j] = tk;                     \
    int tr = cref[i]; cref[i] = cref[j]; cref[j] = tr;                   \
  }
  SR_CSWAP(0, 1) SR_CSWAP(2, 3) SR_CSWAP(0, 2) SR_CSWAP(1, 3) SR_CSWAP(1, 2)
#undef SR_CSWAP
  next = cref[0];
#pragma unroll
  for (int c = 3; c >= 1; --c) {
    if (c < hitc) {
      if (sp < SR_STACK_SIZE) stk[(sp++) * stride] = cref[c];
      else atomicOr(overflow, 1u);
    }
  }
}

__device__ __forceinline__ void visit_leaf(const float4* __restrict__ tris, int ref, f3 o, f3 d, HitD& h) {
  const uint32_t r = (uint32_t)ref;
  const int first = (int)((r & ~SR_LEAF_BIT) >> 3);
  const int cnt = (int)(r & 7u);
  for (int k = 0; k < cnt; ++k) {
    const float4* tp = tris + (size_t)(first + k) * 3;
    intersect_tri(o, d, tp[0], tp[1], tp[2], h.t, h.prim, h.u, h.v);
  }
}

// Closest hit over the 4-wide BVH.  `stk` is this lane's LDS stack column
// (entry i at stk[i * stride]).  Children are visited nearest-first; boxes
// beyond the current best t are culled (ties kept: <=).
//
// Loop structure (while-while with speculative traversal): a lane walks inner
// nodes until it meets a leaf, parks that leaf and keeps walking nodes until
// every lane of the wave has a leaf (or is finished); then all lanes
// intersect their leaves together.  This keeps the wave in one of the two
// loop bodies instead of splitting it between nodes and leaves each step.
// `nodes` / `tris` are passed by value so they sit in scalar registers for
// the whole traversal (no reloads inside the loops).
__device__ __forceinline__ HitD trace_closest(const Node4* __restrict__ nodes, const float4* __restrict__ tris, f3 o,
                                              f3 d, int* stk, int stride, uint32_t* overflow) {
  HitD h;
  h.t = 1e9f;  // HitRecord(t=1e9), main.py:833
  h.prim = INT32_MIN;
  h.u = 0.f;
  h.v = 0.f;
  const float tiny = 1e-12f;
  f3 dd = mk(fabsf(d.x) < tiny ? copysignf(tiny, d.x) : d.x, fabsf(d.y) < tiny ? copysignf(tiny, d.y) : d.y,
             fabsf(d.z) < tiny ? copysignf(tiny, d.z) : d.z);
  f3 inv = mk(1.0f / dd.x, 1.0f / dd.y, 1.0f / dd.z);
  f3 oi = mk(o.x * inv.x, o.y * inv.y, o.z * inv.z);
  const bool nx = inv.x < 0.f, ny = inv.y < 0.f, nz = inv.z < 0.f;
  int sp = 0;
  int cur = 0;       // root node; SR_EMPTY = nothing left
  int leaf = SR_EMPTY;  // parked leaf
  while (cur != SR_EMPTY || leaf != SR_EMPTY) {
    // node loop
    while (cur != SR_EMPTY && cur >= 0) {
      int next;
      visit_node(nodes, cur, inv, oi, nx, ny, nz, h.t, stk, stride, sp, next, overflow);
      if (next == SR_EMPTY) next = sp > 0 ? stk[(--sp) * stride] : SR_EMPTY;
      cur = next;
      if (cur != SR_EMPTY && cur < 0 && leaf == SR_EMPTY) {  // park the first leaf, keep walking
        leaf = cur;
        cur = sp > 0 ? stk[(--sp) * stride] : SR_EMPTY;
      }
      const bool waiting = cur != SR_EMPTY && cur >= 0 && leaf == SR_EMPTY;
      if (__ballot(waiting) == 0ull) break;
    }
    // leaf loop
    while (leaf != SR_EMPTY) {
      visit_leaf(tris, leaf, o, d, h);
      leaf = SR_EMPTY;
      if (cur != SR_EMPTY && cur < 0) {
        leaf = cur;
        cur = sp > 0 ? stk[(--sp) * stride] : SR_EMPTY;
      }
    }
  }
  if (h.prim == INT32_MIN) h.prim = -1;
  return h;
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
__device__ __forceinline__ void onb(f3 n, f3& t, f3& b) {
  f3 up = fabsf(n.y) < 0.999f ? mk(0, 1, 0) : mk(1, 0, 0);
  t = normalize(cross(up, n));
  b = cross(n, t);
}
__device__ __forceinline__ f3 random_in_unit_sphere(Rng& rng) {
  float z = rng.next() * 2.0f - 1.0f;
  float rxy = sqrtf(tmax(0.0f, 1.0f - z * z));
  float phi = rng.next() * 2.0f * SR_PI;
  return mk(rxy * cosf(phi), rxy * sinf(phi), z);
}
__device__ __forceinline__ f3 cosine_hemisphere(f3 n, Rng& rng) {
  float r1 = rng.next();
  float r2 = rng.next();
  float phi = 2.0f * SR_PI * r1;
  float ct = sqrtf(1.0f - r2);
  float st = sqrtf(r2);
  f3 t, b;
  onb(n, t, b);
  return t * cosf(phi) * st + b * sinf(phi) * st + n * ct;
}
__device__ __forceinline__ float hg_eval(float c, float g) {
  float g2 = g * g;
  float den = 1.0f + g2 - 2.0f * g * c;
  return (1.0f - g2) / (4.0f * SR_PI * den * sqrtf(den) + 1e-6f);
}
__device__ __forceinline__ f3 hg_sample(f3 w, float g, Rng& rng) {
  float ct;
  if (fabsf(g) < 1e-4f) {
    ct = 2.0f * rng.next() - 1.0f;
  } else {
    float rs = (1.0f - g * g) / (1.0f - g + 2.0f * g * rng.next());
    ct = (1.0f + g * g - rs * rs) / (2.0f * g);
  }
  float st = sqrtf(tmax(0.0f, 1.0f - ct * ct));
  float phi = 2.0f * SR_PI * rng.next();
  f3 u, v;
  onb(w, u, v);
  return u * cosf(phi) * st + v * sinf(phi) * st + w * ct;
}

// ------------------------------------------------------------------ microfacet / Fresnel (main.py:866-978)
struct cplx {
  float re, im;
};
__device__ __forceinline__ cplx cmul(cplx a, cplx b) { return cplx{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ cplx cdiv(cplx a, cplx b) {
  float m = b.re * b.re + b.im * b.im;
  cplx p = cmul(a, cplx{b.re, -b.im});
  return cplx{p.re / m, p.im / m};
}
__device__ __forceinline__ cplx csqrt_ref(cplx z) {
  float mod = sqrtf(z.re * z.re + z.im * z.im);
  float re = sqrtf((mod + z.re) / 2.0f);
  float sg = z.im > 0.f ? 1.f : (z.im < 0.f ? -1.f : 0.f);
  float im = sg * sqrtf((mod - z.re) / 2.0f);
  if (z.im == 0.0f) im = 0.0f;
  return cplx{re, im};
}
__device__ __forceinline__ float ggx_D(float nh, float r) {
  float alpha = r * r;
  float a2 = alpha * alpha;
  float nh2 = nh * nh;
  float den = nh2 * (a2 - 1.0f) + 1.0f;
  return a2 / (SR_PI * den * den);
}
__device__ __forceinline__ float ggx_G(f3 n, f3 v, f3 l, float r) {
  float nv = tmax(dot(n, v), 0.0f);
  float alpha = tmax(r * r, 1e-4f);
  float k = alpha * alpha / 2.0f;
  float g1v = nv / (nv * (1.0f - k) + k + 1e-7f);
  float nl = tmax(dot(n, l), 0.0f);
  float g1l = nl / (nl * (1.0f - k) + k + 1e-7f);
  return g1v * g1l;
}
__device__ __forceinline__ float fresnel_spectral(float ci, float ni, float ki, float nt, float kt) {
  float c = clampf(ci, -1.0f, 1.0f);
  float ang = acosf(c);
  cplx ein{ni * ni - ki * ki, 2 * ni * ki};
  cplx etr{nt * nt - kt * kt, 2 * nt * kt};
  float ct = cosf(ang);
  float sn = sinf(ang);
  float s2 = sn * sn;
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
  float phi = 2.0f * SR_PI * x0;
  float ct = sqrtf((1.0f - x1) / (1.0f + (a2 - 1.0f) * x1));
  float st = sqrtf(1.0f - ct * ct);
  f3 t, b;
  onb(n, t, b);
  return normalize(t * (cosf(phi) * st) + b * (sinf(phi) * st) + n * ct);
}
__device__ __forceinline__ float oren_nayar(float r, float nl, float nv, f3 l, f3 v, f3 n) {
  float sigma = r * SR_PI / 2;
  float s2 = sigma * sigma;
  float A = 1.0f - s2 / (2.0f * (s2 + 0.33f));
  float B = 0.45f * s2 / (s2 + 0.09f);
  float ti = acosf(nl), to = acosf(nv);
  float al = tmax(ti, to), be = tmin(ti, to);
  float cpd = 0.0f;
  if (nl > 1e-4f && nv > 1e-4f) {
    f3 pi = normalize(l - n * nl);
    f3 po = normalize(v - n * nv);
    cpd = tmax(0.0f, dot(pi, po));
  }
  return A + B * cpd * sinf(al) * tanf(be);
}
