// sr_dual.h -- forward-mode dual numbers for the path-replay adjoint.
//
// Four derivative slots, one per differentiable quantity a single shading
// event can touch: the albedo of the shaded surface (at the active band),
// the GGX roughness of the designated rough material, the HG anisotropy g
// of the designated medium and the absorption coefficient sigma_a (at the
// active band) of the designated medium.
#pragma once
#include "sr_device.h"

enum { SL_ALB = 0, SL_ROUGH = 1, SL_G = 2, SL_SIG = 3, SL_N = 4 };

struct Dl {
  float v;
  float d[SL_N];
};

__device__ __forceinline__ Dl dc(float v) {
  Dl r;
  r.v = v;
#pragma unroll
  for (int i = 0; i < SL_N; ++i) r.d[i] = 0.f;
  return r;
}
__device__ __forceinline__ Dl dvar(float v, int slot, bool active) {
  Dl r = dc(v);
  if (active) r.d[slot] = 1.f;
  return r;
}
__device__ __forceinline__ Dl operator+(Dl a, Dl b) {
  Dl r;
  r.v = a.v + b.v;
#pragma unroll
  for (int i = 0; i < SL_N; ++i) r.d[i] = a.d[i] + b.d[i];
  return r;
}
__device__ __forceinline__ Dl operator-(Dl a, Dl b) {
  Dl r;
  r.v = a.v - b.v;
#pragma unroll
  for (int i = 0; i < SL_N; ++i) r.d[i] = a.d[i] - b.d[i];
  return r;
}
__device__ __forceinline__ Dl operator*(Dl a, Dl b) {
  Dl r;
  r.v = a.v * b.v;
#pragma unroll
  for (int i = 0; i < SL_N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
  return r;
}
__device__ __forceinline__ Dl operator/(Dl a, Dl b) {
  Dl r;
  r.v = a.v / b.v;
  const float ib = 1.0f / b.v;
#pragma unroll
  for (int i = 0; i < SL_N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
  return r;
}
__device__ __forceinline__ Dl operator+(Dl a, float b) { a.v += b; return a; }
__device__ __forceinline__ Dl operator+(float b, Dl a) { a.v += b; return a; }
__device__ __forceinline__ Dl operator-(Dl a, float b) { a.v -= b; return a; }
__device__ __forceinline__ Dl operator-(float b, Dl a) {
  Dl r;
  r.v = b - a.v;
#pragma unroll
  for (int i = 0; i < SL_N; ++i) r.d[i] = -a.d[i];
  return r;
}
__device__ __forceinline__ Dl operator*(Dl a, float b) {
  a.v *= b;
#pragma unroll
  for (int i = 0; i < SL_N; ++i) a.d[i] *= b;
  return a;
}
__device__ __forceinline__ Dl operator*(float b, Dl a) { return a * b; }
__device__ __forceinline__ Dl operator/(Dl a, float b) { return a * (1.0f / b); }
__device__ __forceinline__ Dl operator/(float a, Dl b) { return dc(a) / b; }
__device__ __forceinline__ Dl dsqrt(Dl a) {
  Dl r;
  r.v = sqrtf(a.v);
  const float k = r.v > 0.f ? 0.5f / r.v : 0.f;
#pragma unroll
  for (int i = 0; i < SL_N; ++i) r.d[i] = a.d[i] * k;
  return r;
}
__device__ __forceinline__ Dl dexp(Dl a) {
  Dl r;
  r.v = expf(a.v);
#pragma unroll
  for (int i = 0; i < SL_N; ++i) r.d[i] = a.d[i] * r.v;
  return r;
}
// d(log a) per slot (0 where a == 0)
__device__ __forceinline__ void dlog_into(const Dl& a, float out[SL_N]) {
  const float ia = a.v != 0.f ? 1.0f / a.v : 0.f;
#pragma unroll
  for (int i = 0; i < SL_N; ++i) out[i] = a.d[i] * ia;
}

// ------------------------------------------------ dual versions of the shading terms
__device__ __forceinline__ Dl ggx_D_d(float nh, Dl r) {  // main.py:900-907
  Dl alpha = r * r;
  Dl a2 = alpha * alpha;
  float nh2 = nh * nh;
  Dl den = (a2 - 1.0f) * nh2 + 1.0f;
  return a2 / ((den * den) * SR_PI);
}
__device__ __forceinline__ Dl ggx_G_d(f3 n, f3 v, f3 l, Dl r) {  // main.py:909-923
  float nv = tmax(dot(n, v), 0.0f);
  Dl alpha = r * r;
  if (alpha.v < 1e-4f) alpha = dc(1e-4f);
  Dl k = (alpha * alpha) * 0.5f;
  Dl g1v = dc(nv) / ((1.0f - k) * nv + k + 1e-7f);
  float nl = tmax(dot(n, l), 0.0f);
  Dl g1l = dc(nl) / ((1.0f - k) * nl + k + 1e-7f);
  return g1v * g1l;
}
__device__ __forceinline__ Dl oren_nayar_d(Dl r, float nl, float nv, f3 l, f3 v, f3 n) {
  Dl sigma = r * (SR_PI / 2);
  Dl s2 = sigma * sigma;
  Dl A = 1.0f - s2 / ((s2 + 0.33f) * 2.0f);
  Dl Bc = (s2 * 0.45f) / (s2 + 0.09f);
  return A + Bc * oren_nayar_angular(nl, nv, l, v, n);
}
__device__ __forceinline__ Dl hg_eval_d(float c, Dl g) {  // main.py:758-766
  Dl g2 = g * g;
  Dl den = 1.0f + g2 - g * (2.0f * c);
  return (1.0f - g2) / ((den * dsqrt(den)) * (4.0f * SR_PI) + 1e-6f);
}
// d log p / d g of the HG sampling density at cos c (sample_henyey_greenstein,
// main.py:768-792, draws exactly the HG distribution).
__device__ __forceinline__ float hg_score_g(float c, float g) {
  if (fabsf(g) < 1e-4f) return 3.0f * c;  // isotropic branch: derivative of HG at g = 0
  float den = 1.0f + g * g - 2.0f * g * c;
  return -2.0f * g / (1.0f - g * g) - 1.5f * (2.0f * g - 2.0f * c) / den;
}
// d log p / d r of the GGX half-vector density D_s(h) cos(h) with
// a2 = r^2 (importance_sample_ggx is called with alpha = roughness,
// main.py:968-978, 1019).
__device__ __forceinline__ float ggx_sample_score_r(float c, float r) {
  float a2 = r * r;
  float den = c * c * (a2 - 1.0f) + 1.0f;
  return (1.0f / a2 - 2.0f * c * c / den) * (2.0f * r);
}
