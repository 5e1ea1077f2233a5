// sr_path.hip -- the persistent spectral path loop (forward) and its
// path-replay adjoint, sharing one lane state machine.
//
// Work item = (pixel, spectral block).  A lane runs the item's B single-band
// paths one after another (cast_rays once per band, main.py:1502-1590),
// fetching items from a device queue in wave-sized batches.  Every loop
// iteration performs exactly one closest-hit traversal per active lane --
// either a path's extension ray (one path segment) or one step of its NEE
// shadow walk -- so lanes of a wave stay converged in the traversal, the
// dominant cost.  Per-lane state lives in VGPRs; the traversal stack in LDS.
//
// Adjoint (ADJ = true): each path is replayed twice with identical random
// numbers.  Pass 0 accumulates L = sum_k dL/dacc[band_k] * c_k over the
// path's contributions c_k.  Pass 1 repeats the path and, at every
// parameter-dependent step, adds to the gradient
//   (radiance still to come) * (d log factor / d theta + d log p / d theta)
// for throughput factors and sampling densities (score function), plus
// dL/dacc * d c_k / d theta for the local terms of each contribution
// (NEE bsdf/phase, MIS weights, shadow transmittance).  Derivatives of the
// local terms come from dual numbers (sr_dual.h); values and branch
// decisions always come from the forward code, so both passes follow the
// forward path exactly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "sr_dual.h"
#include "sr_kernels.h"

// PH_NEW: a path just ended; PH_NEXT: start the item's next band's path
enum { PH_NEW = 0, PH_EXT = 1, PH_SHD = 2, PH_REPLAY = 3, PH_PARK = 4, PH_FIN = 5, PH_NEXT = 6 };


struct Lane {
  int phase, band, pi, pj, blk;
  f3 pos, dir;
  float tp, sp;
  float dvol;  // the extension ray's free flight (scatter_volume), drawn before its trace
  int active, pl, vol, seg;
  float acc;
  int acc_band;
  // pending extension event (scatter_volume result)
  f3 en, esn, edir;
  int efront, emat, eaction, enew, nee_surf;
  // pending shadow walk (shadow_transmittance state); once the walk is
  // over, sdir keeps the NEE vertex (the light tree's MIS pdf at the next
  // emitter hit is taken there)
  f3 sdir;
  float smax, strav, stot, soff;
  int sprim, svol, sstep;
  // subsurface random walk (SR_RENDER_SUBSURFACE): 0 none, 1 walking inside
  // material wmat (step wstep), 2 at the exit point (NEE, then the exit lobe)
  int walk, wstep, wmat;
  // during the shadow walk (PH_SHD): the NEE light's area over its pick
  // probability (the band's emitting area for the CDF pick); then the NEE
  // contribution awaiting finish_segment (PH_FIN)
  float fdirect;
  uint32_t n_seg, n_trav, n_nf;
  // adjoint only
  int pass;
  float Lw, Eb, sp_dr, stot_dsig, hg_sc;
};

// The small tables of the lane's scene that every segment reads (get_mat,
// the band wavelengths, the NEE light records), seen where the launch keeps
// them: in LDS when every job's tables fit the block's share (copied there
// by the block, so these reads wait for LDS instead of the caches),
// otherwise in global memory.
struct Tabs {
  const MatRec* mats;      // [M][B]
  const float* lam_m;      // [B]
  const float* lam_nm;     // [B]
  const float* light_area; // [B]
  const int* light_count;  // [B]
  const int* mat_eem;      // [M]
  const LightRec* lights;  // [B][max_lights]
  const float4* lverts;    // [slots][3]
  int B;
  __device__ __forceinline__ MatRec mat(int m, int b) const { return mats[m * B + b]; }  // get_mat, main.py:980-996
};

__device__ __forceinline__ Tabs global_tabs(const SceneDev& s) {
  return Tabs{s.mats, s.lam_m, s.lam_nm, s.light_area, s.light_count, s.mat_eem, s.lights, s.lverts, s.B};
}
__device__ __forceinline__ Tabs lds_tabs(const SceneDev& s, const LdsTables& L, const char* base) {
  return Tabs{reinterpret_cast<const MatRec*>(base + L.mats),       reinterpret_cast<const float*>(base + L.lam_m),
              reinterpret_cast<const float*>(base + L.lam_nm),      reinterpret_cast<const float*>(base + L.light_area),
              reinterpret_cast<const int*>(base + L.light_count),   reinterpret_cast<const int*>(base + L.mat_eem),
              reinterpret_cast<const LightRec*>(base + L.lights),   reinterpret_cast<const float4*>(base + L.lverts),
              s.B};
}

// The block's copy of a job's tables, made by all its threads when the
// kernel starts (job 0) and each time the block's waves move to the next
// job of a batch; ends with a barrier.
__device__ __forceinline__ void copy_tables(const SceneDev& s, const LdsTables& L, char* base) {
  const int t = threadIdx.x, n = blockDim.x;
  auto copy16 = [&](uint32_t off, const void* src, uint32_t count) {
    const float4* a = reinterpret_cast<const float4*>(src);
    float4* d = reinterpret_cast<float4*>(base + off);
    for (uint32_t i = t; i < count; i += n) d[i] = a[i];
  };
  auto copy4 = [&](uint32_t off, const void* src, uint32_t count) {
    const int* a = reinterpret_cast<const int*>(src);
    int* d = reinterpret_cast<int*>(base + off);
    for (uint32_t i = t; i < count; i += n) d[i] = a[i];
  };
  copy16(L.mats, s.mats, L.n_mats * 2);
  copy16(L.lights, s.lights, L.n_lights);
  copy16(L.lverts, s.lverts, L.n_lverts);
  copy4(L.lam_m, s.lam_m, s.B);
  copy4(L.lam_nm, s.lam_nm, s.B);
  copy4(L.light_area, s.light_area, s.B);
  copy4(L.light_count, s.light_count, s.B);
  copy4(L.mat_eem, s.mat_eem, s.M);
  __syncthreads();
}

// The lane's current job (wave-uniform), its scene's tables, plus the
// adjoint configuration.
template <bool ADJ>
struct Ctx {
  const Job& a;
  const AdjMap& adj;
  float* lds_grad;
  Tabs t;
  float* rec;  // the job's path totals, or null -- a constant null in kernels built without them
};

// The lane's uniform block (segment `seg`, slot `slot`); the key is the job's
// seed (wave-uniform, so the key schedule runs on the scalar unit).
__device__ __forceinline__ U4 draws(const Job& a, const Lane& L, int seg, int slot) {
  return draw4((uint32_t)(L.pi * a.H + L.pj), (uint32_t)L.blk, (uint32_t)L.band | (a.stream << 8), (uint32_t)seg,
               (uint32_t)slot, a.seed_lo, a.seed_hi);
}

template <bool ADJ>
__device__ __forceinline__ float dacc(const Ctx<ADJ>& c, const Lane& L, int band) {
  return c.adj.d_accum[((size_t)L.pi * c.a.H + L.pj) * c.a.s.B + band];
}

template <bool ADJ>
__device__ __forceinline__ void emit_grad(const Ctx<ADJ>& c, int idx, float v) {
  if (idx >= 0 && v != 0.0f && isfinite(v)) atomicAdd(c.lds_grad + idx, v);
}

// L.acc_band: the band L.acc is summed for, with SR_ACC_SWITCHED set once
// a contribution of the path went to another band than the path's own and,
// in a launch that records path totals, SR_ACC_TOUCHED once the path added
// any (finite, nonzero) contribution.
#define SR_ACC_SWITCHED 0x100
#define SR_ACC_TOUCHED 0x200
#define SR_ACC_BAND_MASK 0xFF

template <bool ADJ>
__device__ __forceinline__ void flush_acc(Lane& L, const Ctx<ADJ>& c) {
  if (!ADJ && L.acc != 0.0f) {
    size_t idx = ((size_t)L.pi * c.a.H + L.pj) * c.a.s.B + (L.acc_band & SR_ACC_BAND_MASK);
    atomicAdd(c.a.sum + idx, L.acc);
  }
  L.acc = 0.0f;
}

// Index of the lane's path in Job::rec: [block - first_block][i][j][band].
template <bool ADJ>
__device__ __forceinline__ size_t rec_index(const Lane& L, const Ctx<ADJ>& c) {
  return (((size_t)(L.blk - c.a.first_block) * c.a.W + L.pi) * c.a.H + L.pj) * c.a.s.B + L.band;
}

// block_accumulator[i, j, active] += contribution (main.py:1589-1590);
// non-finite contributions are dropped and counted.
template <bool ADJ>
__device__ __forceinline__ void add_contrib(Lane& L, const Ctx<ADJ>& c, float v) {
  if (!isfinite(v)) {
    L.n_nf++;
    return;
  }
  if (v == 0.0f) return;
  if (ADJ) {
    float w = dacc(c, L, L.active) * v;
    if (L.pass == 0) L.Lw += w;
    else L.Eb += w;
    return;
  }
  if (L.active != (L.acc_band & SR_ACC_BAND_MASK)) {
    flush_acc<ADJ>(L, c);
    L.acc_band = L.active | SR_ACC_SWITCHED;
  }
  if (c.rec) L.acc_band |= SR_ACC_TOUCHED;
  L.acc += v;
}

template <bool ADJ>
__device__ __forceinline__ void end_path(Lane& L, const Ctx<ADJ>& c) {
  // The path's total for the adjoint: L.acc holds every contribution of
  // the path unless one went to another band (NaN).  A path that added no
  // contribution is recorded as -0.0, which a sum of nonzero contributions
  // started at +0.0 never is (x + -x rounds to +0.0).
  if (!ADJ && c.rec)
    c.rec[rec_index(L, c)] =
        (L.acc_band & SR_ACC_SWITCHED) ? NAN : ((L.acc_band & SR_ACC_TOUCHED) ? L.acc : -0.0f);
  flush_acc<ADJ>(L, c);
  if (ADJ && L.pass == 0) {
    L.pass = 1;  // replay the same path for the gradient pass
    L.phase = PH_REPLAY;
  } else {
    L.phase = PH_NEW;
  }
}

// Camera ray of cast_rays, main.py:1506-1529.
template <bool ADJ>
__device__ __forceinline__ void start_path(Lane& L, const Ctx<ADJ>& c) {
  const Job& a = c.a;
  const SceneDev& s = a.s;
  const U4 r = draws(a, L, 0, SR_SLOT_AUX);
  float u = fdiv((float)L.pi + r.u[0], (float)a.W);
  float v = fdiv((float)L.pj + r.u[1], (float)a.H);
  const float aspect = fdiv((float)a.W, (float)a.H);
  f3 cu = mk(s.cam_u[0], s.cam_u[1], s.cam_u[2]);
  f3 cv = mk(s.cam_v[0], s.cam_v[1], s.cam_v[2]);
  f3 cw = mk(s.cam_w[0], s.cam_w[1], s.cam_w[2]);
  L.dir = normalize(cu * (u * 2.0f - 1.0f) * aspect * 1.0f + cv * (v * 2.0f - 1.0f) * 1.0f - cw * 2.0f);
  L.pos = mk(s.cam_o[0], s.cam_o[1], s.cam_o[2]);
  L.tp = 1.0f;
  L.sp = 1.0f;
  L.active = L.band;
  L.pl = 0;
  L.vol = 0;
  L.seg = 0;
  L.acc = 0.0f;
  L.acc_band = L.band;
  L.walk = 0;
  L.wstep = 0;
  L.wmat = 0;
  if (ADJ) {
    if (L.pass == 0) {
      L.Lw = 0.0f;
      // A total recorded by the forward replaces the first pass: pass 0
      // would sum dL/dacc[band] x c_k over the same contributions c_k, all
      // in the path's own band.
      if (c.rec) {
        const float tot = c.rec[rec_index(L, c)];
        if (isfinite(tot)) {
          L.Lw = dacc(c, L, L.band) * tot;
          L.pass = 1;
          // A path that added no contribution adds nothing to the gradient:
          // each of its terms is (radiance still to come = 0) x (...) or the
          // derivative of a contribution that is 0 (skipped), so it is not
          // replayed at all.
          if (tot == 0.0f && signbit(tot)) {
            end_path<ADJ>(L, c);
            return;
          }
        }
      }
    }
    L.Eb = 0.0f;
    L.sp_dr = 0.0f;
  }
  if (a.depth > 0) {
    L.phase = PH_EXT;
  } else {
    end_path<ADJ>(L, c);
    if (L.phase == PH_REPLAY) L.phase = PH_NEW;
  }
}

// evaluate_bsdf, main.py:1360-1444 (its transmission branch is unreachable:
// both cosines are clamped to >= 1e-4 before NdotV * NdotL > 0 is tested).
// With `walk` (SR_RENDER_SUBSURFACE) an opaque material with sigma_s > 0 has
// no diffuse lobe here: that light enters the subsurface walk instead.
__device__ __forceinline__ void evaluate_bsdf(const Tabs& s, int vol_id, int surf_id, f3 V, f3 Lv, f3 N, int b,
                                              float& bsdf, float& pdf, bool walk = false) {
  MatRec vm = s.mat(vol_id, b);
  MatRec sm = s.mat(surf_id, b);
  float NdotL = tmax(1e-4f, dot(N, Lv));
  float NdotV = tmax(1e-4f, dot(N, V));
  bsdf = 0.0f;
  pdf = 0.0f;
  f3 H = normalize(V + Lv);
  if (dot(H, N) < 0.0f) H = -H;
  float VdotH = tmax(1e-4f, dot(V, H));
  float F = fresnel_spectral(VdotH, vm.n, vm.k, sm.n, sm.k);
  if (sm.r > 0.008f) {
    float NdotH = tmax(1e-4f, dot(N, H));
    float D = ggx_D(NdotH, sm.r);
    float G = ggx_G(N, V, Lv, sm.r);
    bsdf += fdiv(D * G * F, 4.0f * NdotL * NdotV);
    pdf += fdiv(D * NdotH * F, 4.0f * VdotH);
  }
  float absorb = fdiv(4.0f * SR_PI * sm.k * 1e9f, s.lam_nm[b]);
  float albedo = fdiv(sm.ss, sm.ss + absorb + 1e-6f);
  float on = oren_nayar(sm.r, NdotL, NdotV, Lv, V, N);
  float opq = (walk && sm.ss > 0.0f) ? 0.0f : 1.0f - (float)sm.tv;
  bsdf += (1.0f - F) * opq * (albedo * SR_INV_PI) * on;
  pdf += NdotL * 0.318309886184f * (1.0f - F) * opq;
}

// Dual version (derivatives only) of evaluate_bsdf with respect to the
// surface's albedo and, when it is the designated rough material, its
// roughness.  The Fresnel term's dependence on k is not differentiated (k of
// a diffuse albedo is ~1e-8 and moves F by less than 1e-7).
__device__ __forceinline__ void evaluate_bsdf_d(const Tabs& s, const AdjMap& m, int vol_id, int surf_id, f3 V,
                                                f3 Lv, f3 N, int b, Dl& bsdf, Dl& pdf) {
  MatRec vm = s.mat(vol_id, b);
  MatRec sm = s.mat(surf_id, b);
  Dl r = dvar(sm.r, SL_ROUGH, surf_id == m.rough_mat);
  float NdotL = tmax(1e-4f, dot(N, Lv));
  float NdotV = tmax(1e-4f, dot(N, V));
  bsdf = dc(0.0f);
  pdf = dc(0.0f);
  f3 H = normalize(V + Lv);
  if (dot(H, N) < 0.0f) H = -H;
  float VdotH = tmax(1e-4f, dot(V, H));
  float F = fresnel_spectral(VdotH, vm.n, vm.k, sm.n, sm.k);
  if (sm.r > 0.008f) {
    float NdotH = tmax(1e-4f, dot(N, H));
    Dl D = ggx_D_d(NdotH, r);
    Dl G = ggx_G_d(N, V, Lv, r);
    bsdf = bsdf + (D * G * F) / (4.0f * NdotL * NdotV);
    pdf = pdf + (D * (NdotH * F)) / (4.0f * VdotH);
  }
  float absorb = (4.0f * SR_PI * sm.k * 1e9f) / s.lam_nm[b];
  Dl albedo = dvar(sm.ss / (sm.ss + absorb + 1e-6f), SL_ALB, m.alb_off[surf_id] >= 0);
  Dl on = oren_nayar_d(r, NdotL, NdotV, Lv, V, N);
  float opq = 1.0f - (float)sm.tv;
  bsdf = bsdf + ((albedo * ((1.0f - F) * opq / SR_PI)) * on);
  pdf = pdf + NdotL * 0.318309886184f * (1.0f - F) * opq;
}

// A subsurface walk keeps its medium and its next flight in the shadow
// walk's fields, which are unused until the walk's exit NEE: L.stot =
// sigma_t, L.strav = sigma_s / sigma_t, L.soff = the HG g, L.smax = the
// current step's flight length.  The walk's steps are described at
// walk_step; it has no counterpart in the reference, so its arithmetic
// uses the hardware logarithm and reciprocal.
__device__ __forceinline__ float walk_flight(float u0, float sigma_t) {
  return -__logf(1.0f - u0) * frcp(sigma_t);
}

// scatter_surface, main.py:1000-1188, applied to the lane's ray as
// cast_rays does (main.py:1565-1574).  `direct` is this segment's NEE
// contribution (accumulated after the scatter), needed by the adjoint.
template <bool ADJ, bool SSS>
__device__ __forceinline__ void scatter_surface(Lane& L, const Ctx<ADJ>& c, float direct) {
  const Job& a = c.a;
  const SceneDev& s = a.s;
  if (SSS && L.walk == 2) {  // leaving a subsurface walk: Lambertian lobe about the outward normal
    const U4 u = draws(a, L, (L.seg << 9) | L.wstep, SR_SLOT_WALK_EXIT);
    const f3 S = cosine_hemisphere(L.en, u.u[0], u.u[1]);
    L.sp = 0.318309886184f * tmax(1e-4f, dot(L.en, S));
    L.dir = S;
    L.pos = L.pos + L.en * SR_EPSILON;
    L.walk = 0;
    return;
  }
  const int ab = L.active;
  MatRec sm = c.t.mat(L.emat, ab);
  MatRec vm = c.t.mat(L.vol, ab);
  const f3 V = -L.dir;
  const f3 N = L.esn;
  const float ni = vm.n, ki = vm.k, nt = sm.n, kt = sm.k;
  const U4 rs = draws(a, L, L.seg, SR_SLOT_SCAT);
  float x0 = rs.u[0];
  float x1 = rs.u[1];
  f3 H = ggx_sample_half(x0, x1, N, sm.r);
  float ia = tmax(1e-4f, dot(V, H));
  float F = fresnel_spectral(ia, ni, ki, nt, kt);
  float rnd = clampf(rs.u[2], 0.0f, 0.99f);
  int dr = 0;
  f3 ndir;
  float est, pdf;
  // branch id for the adjoint: 0 delta, 1 GGX reflection (max cos),
  // 2 GGX reflection on TIR (|cos|), 3 Oren-Nayar, 4 rough refraction
  int branch = 0;
  float on_NL = 0.f, on_NV = 0.f, on_albedo = 0.f;
  f3 on_S = mk(0.f, 0.f, 0.f);
  float rf_LdotH = 0.f, rf_VdotH = 0.f, rf_NdotH = 0.f, rf_NdotL = 0.f, rf_NdotV = 0.f, rf_dsq = 0.f, rf_t1 = 0.f;
  if (rnd < F) {
    ndir = reflect(L.dir, H);
    dr = 1;
    if (sm.r < 0.008f) { est = F; pdf = 1e9f; }
    else branch = 1;
  } else {
    float absorb = fdiv(4.0f * SR_PI * kt, c.t.lam_m[ab]);
    float dte = fdiv(-logf(-rs.u[3] + 1.0f + 1e-7f), sm.ss + absorb + 1e-6f);
    if (SSS && a.sss && !sm.tv && sm.ss > 0.0f) {
      // random-walk subsurface scattering (the TODO of main.py:1066) in
      // place of the single-scatter Oren-Nayar lobe: enter along a cosine-
      // distributed direction into the surface; walk_step does the rest
      const U4 ro = draws(a, L, L.seg, SR_SLOT_OPQ);
      ndir = cosine_hemisphere(-L.en, ro.u[0], ro.u[1]);
      est = 1.0f - F;
      pdf = 1.0f;
      dr = 1;  // the outside medium stays the path's medium
      branch = 5;
      L.walk = 1;
      L.wstep = 0;
      L.wmat = L.emat;
      L.stot = sm.ss + absorb;
      L.strav = fdiv(sm.ss, L.stot);
      L.soff = sm.g;
      L.smax = walk_flight(draws(a, L, L.seg << 9, SR_SLOT_WALK).u[0], L.stot);
    } else if (!sm.tv) {
      dr = 1;
      branch = 3;
      const U4 ro = draws(a, L, L.seg, SR_SLOT_OPQ);
      f3 S = cosine_hemisphere(N, ro.u[0], ro.u[1]);
      float NdotL = tmax(1e-4f, dot(N, S));
      float NdotV = tmax(1e-4f, dot(N, V));
      float albedo = fdiv(sm.ss, sm.ss + absorb + 1e-6f);
      float bsdf = (1.0f - F) * (albedo * SR_INV_PI) * oren_nayar(sm.r, NdotL, NdotV, S, V, N);
      pdf = 0.318309886184f * NdotL;
      on_NL = NdotL; on_NV = NdotV; on_albedo = albedo; on_S = S;
      const int eid = c.t.mat_eem[L.emat];
      if (eid > 0) {  // surface fluorescence reweights only (main.py:1090-1121)
        const float* row = s.eem + ((size_t)eid * s.B + ab) * s.B;
        float qy = 0.0f;
        for (int j = 0; j < s.B; ++j) qy = row[j] > qy ? row[j] : qy;
        float r3 = ro.u[2];
        if (qy > 0.0f && r3 < qy) {
          float norm = 0.0f;
          for (int j = 0; j < s.B; ++j) norm += row[j];
          float r4 = ro.u[3];
          float cum = 0.0f;
          int nw = ab;
          const float rn = frcp(norm);
          for (int j = 0; j < s.B; ++j) {
            cum += row[j] * rn;
            if (r4 < cum) { nw = j; break; }
          }
          MatRec fm = c.t.mat(L.emat, nw);
          float ab2 = fdiv(4.0f * SR_PI * fm.k, c.t.lam_m[nw]);
          bsdf *= expf(-ab2 * dte);
        }
      }
      est = fdiv(bsdf * NdotL, pdf);
      ndir = S;
    } else {
      f3 Rf = refract(L.dir, H, ni / nt);
      if (dot(Rf, Rf) < 1e-4f) {  // total internal reflection
        ndir = reflect(L.dir, H);
        dr = 1;
        if (sm.r < 0.008f) { est = F; pdf = 1e9f; }
        else branch = 2;
      } else {
        dr = 0;
        ndir = Rf;
        if (sm.r < 0.0018f) {
          est = 1.0f - F;
          pdf = 1e9f;
        } else {
          branch = 4;
          float LdotH = dot(Rf, H);
          float VdotH = dot(V, H);
          float NdotH = tmax(1e-4f, dot(N, H));
          float NdotL = tmax(1e-4f, fabsf(dot(N, Rf)));
          float NdotV = tmax(1e-4f, fabsf(dot(N, V)));
          float G = ggx_G(N, V, Rf, sm.r);
          float D = ggx_D(NdotH, sm.r);
          float den = ni * VdotH + nt * LdotH;
          float dsq = den * den;
          float t1 = fdiv(fabsf(LdotH * VdotH), NdotL * NdotV);
          float t2 = fdiv(nt * nt * (1.0f - F) * D * G, dsq);
          float bsdf = t1 * t2;
          pdf = fdiv(D * NdotH * (1.0f - F) * nt * nt * fabsf(LdotH), dsq);
          est = fdiv(bsdf * NdotL, pdf);
          rf_LdotH = LdotH; rf_VdotH = VdotH; rf_NdotH = NdotH; rf_NdotL = NdotL; rf_NdotV = NdotV;
          rf_dsq = dsq; rf_t1 = t1;
        }
      }
    }
  }
  float g_NdotL = 0.f, g_NdotH = 0.f, g_NdotV = 0.f, g_c = 0.f;
  if (branch == 1 || branch == 2) {  // GGX reflection lobe (main.py:1039-1053 and 1142-1156)
    float NdotL = tmax(1e-4f, dot(N, ndir));
    float NdotH = tmax(1e-4f, dot(N, H));
    float NdotV = tmax(1e-4f, dot(N, V));
    float D = ggx_D(NdotH, sm.r);
    float G = ggx_G(N, V, ndir, sm.r);
    float bsdf = fdiv(D * G * F, 4.0f * NdotL * NdotV);
    pdf = fdiv(D * NdotH, 4.0f * ia);
    float cc = branch == 2 ? fabsf(dot(N, ndir)) : tmax(0.0f, dot(N, ndir));
    est = fdiv(bsdf * cc, pdf);
    g_NdotL = NdotL; g_NdotH = NdotH; g_NdotV = NdotV; g_c = cc;
  }
  if (ADJ && L.pass == 1) {
    const AdjMap& m = c.adj;
    const bool rough = L.emat == m.rough_mat;
    const int aoff = m.alb_off[L.emat];
    // radiance that follows this scatter (excludes this segment's NEE term)
    float R = L.Lw - L.Eb - (isfinite(direct) ? dacc(c, L, ab) * direct : 0.0f);
    Dl r = dvar(sm.r, SL_ROUGH, rough);
    Dl est_d = dc(est);
    Dl pdf_d = dc(pdf);
    if (branch == 1 || branch == 2) {
      Dl D = ggx_D_d(g_NdotH, r);
      Dl G = ggx_G_d(N, V, ndir, r);
      pdf_d = (D * g_NdotH) / (4.0f * ia);
      est_d = (((D * G) * F) / (4.0f * g_NdotL * g_NdotV) * g_c) / pdf_d;
    } else if (branch == 3) {
      Dl alb = dvar(on_albedo, SL_ALB, aoff >= 0);
      Dl on = oren_nayar_d(r, on_NL, on_NV, on_S, V, N);
      est_d = ((alb * ((1.0f - F) / SR_PI)) * on) * on_NL / pdf;
    } else if (branch == 4) {
      Dl D = ggx_D_d(rf_NdotH, r);
      Dl G = ggx_G_d(N, V, ndir, r);
      Dl t2 = (D * G) * (nt * nt * (1.0f - F)) / rf_dsq;
      pdf_d = (D * (rf_NdotH * (1.0f - F) * nt * nt * fabsf(rf_LdotH))) / rf_dsq;
      est_d = ((t2 * rf_t1) * rf_NdotL) / pdf_d;
    }
    float dl[SL_N];
    dlog_into(est_d, dl);
    if (aoff >= 0) emit_grad(c, aoff + ab, R * dl[SL_ALB]);
    if (rough) {
      // half vector drawn from D_{a2 = r^2}(h) (n.h) regardless of the branch
      float sc = (m.flags & SR_ADJ_SKIP_ROUGH_SCORE) ? 0.0f : ggx_sample_score_r(tmax(1e-4f, dot(N, H)), sm.r);
      emit_grad(c, m.rough_idx, R * (dl[SL_ROUGH] + sc));
    }
    L.sp_dr = rough ? pdf_d.d[SL_ROUGH] : 0.0f;
  }
  L.sp = pdf;
  L.dir = ndir;
  L.tp *= est;
  float fdr = (float)dr;
  const float side = (SSS && branch == 5) ? -1.0f : (fdr - (1.0f - fdr));  // a walk starts inside
  L.pos = L.pos + L.en * SR_EPSILON * side;
  L.vol = L.emat * L.efront * (1 - dr) + L.vol * dr;
}

// End of one path-loop iteration (main.py:1565-1590).
template <bool ADJ, bool SSS>
__device__ __forceinline__ void finish_segment(Lane& L, const Ctx<ADJ>& c, float direct) {
  if (ADJ && L.pass == 1 && L.hg_sc != 0.0f) {
    float R = L.Lw - L.Eb - (isfinite(direct) ? dacc(c, L, L.active) * direct : 0.0f);
    emit_grad(c, c.adj.g_idx, R * L.hg_sc);
    L.hg_sc = 0.0f;
  }
  if (L.eaction == 1) {
    L.pl += 1;
    scatter_surface<ADJ, SSS>(L, c, direct);
  } else if (L.eaction == 2) {
    L.dir = L.edir;
    L.pl += 1;
  } else if (L.eaction == 3) {
    L.dir = L.edir;
    L.active = L.enew;
    L.pl += 1;
  }
  add_contrib<ADJ>(L, c, direct);
  L.seg += 1;
  if (L.seg >= c.a.depth) {
    end_path<ADJ>(L, c);
  } else if (c.a.end_dead && L.tp == 0.0f) {
    end_path<ADJ>(L, c);  // every later contribution is 0 (SR_RENDER_END_DEAD_PATHS)
  } else if (!isfinite(L.tp)) {
    // non-finite throughput (e.g. the 0/0 rough BTDF estimator on leaving a
    // rough dielectric, main.py:1176-1185 with n_i == n_t): end the path
    L.n_nf++;
    end_path<ADJ>(L, c);
  } else {
    L.phase = PH_EXT;
  }
}

// ------------------------------------------------------------------ light tree
// Many-light NEE (SR_RENDER_LIGHT_TREE).  A node's importance for a shading
// point x is its emitting area in the band times a conservative bound of
// the emitters' cosine towards x over the squared distance: the angle
// between the node's normal cone and the direction to x, reduced by the
// cone half angle (widened by a margin for f32 rounding when the tree is
// built) and by the angle the bounding sphere subtends, is clamped at 0;
// its cosine comes from angle-difference identities (no inverse
// trigonometry).  An importance of 0 means every emitter below faces away
// from x (cos_theta_light <= 0, main.py:1333).  The pick
// probability p of a light then replaces its area fraction: the estimator
// and the MIS pdfs use A_light / p where the reference uses the band's
// total emitting area (main.py:1351, 1355, 1556), which is the same thing
// for the area-CDF pick.
__device__ __forceinline__ float lt_importance(const LightNode& nd, float P, f3 x) {
  if (!(P > 0.0f)) return 0.0f;
  const f3 dv = x - mk(nd.c[0], nd.c[1], nd.c[2]);
  const float r2 = nd.r2;
  const float d2 = dot(dv, dv);
  float cb = 1.0f;  // inside the bounding sphere: no angular bound
  if (d2 > r2) {
    const float id = __builtin_amdgcn_rsqf(d2);
    const float ct = fminf(1.0f, fmaxf(-1.0f, dot(mk(nd.axis[0], nd.axis[1], nd.axis[2]), dv) * id));
    if (ct < nd.cos_o) {  // x outside the normal cone: theta - theta_o in (0, pi]
      const float st = fsqrt(fmaxf(0.0f, 1.0f - ct * ct));
      const float cd = ct * nd.cos_o + st * nd.sin_o;  // cos(theta - theta_o)
      const float sd = st * nd.cos_o - ct * nd.sin_o;  // sin(theta - theta_o)
      const float su = fminf(1.0f, nd.r * id), cu = fsqrt(fmaxf(0.0f, 1.0f - su * su));
      if (cd < cu) {  // beyond the sphere's angle too
        cb = cd * cu + sd * su;  // cos(theta - theta_o - theta_u)
        if (cb <= 0.0f) return 0.0f;
      }
    }
  }
  // distance to the centre, floored for points near the centre of a large node
  return P * cb * frcp(fmaxf(d2, 1e-3f * r2 + 1e-12f));
}

// The nodes are stored in sibling pairs: the children of an inner node sit
// at an even index f and f + 1 (root 0, node 1 unused), and the band's
// powers of a pair are adjacent too, so each level of a descent or of the
// pdf walk is one 128-byte node line plus one 8-byte power pair, both
// addressed by the pair index alone.
__device__ __forceinline__ float2 lt_power_pair(const SceneDev& s, int band, int f) {
  return *reinterpret_cast<const float2*>(s.lpower + (size_t)band * s.n_lnodes + f);
}

// Descends from the root with one uniform; returns the leaf (or -1 if no
// emitter below the root can face x) and its probability.
__device__ int lt_sample(const SceneDev& s, int band, f3 x, float u, float& p) {
  p = 1.0f;
  if (s.n_lkeys == 1) return 0;  // the root is the only leaf
  int f = 2;                     // the root's children
  // median splits keep the depth below 32; the bound only guarantees that
  // the loop ends whatever the arrays hold
  for (int level = 0; level < 64; ++level) {
    const LightNode nl = s.lnodes[f], nr = s.lnodes[f + 1];
    const float2 P = lt_power_pair(s, band, f);
    const float il = lt_importance(nl, P.x, x), ir = lt_importance(nr, P.y, x);
    const float sum = il + ir;
    if (!(sum > 0.0f)) return -1;
    const float rs = frcp(sum);
    const float pl = il * rs;
    int node, next, right;
    if (u < pl || !(ir > 0.0f)) {  // (il * rcp(il) may round below 1)
      u = fminf(fdiv(u, pl), 0x1.fffffep-1f);
      p *= pl;
      node = f;
      next = nl.left;
      right = nl.right;
    } else {
      const float pr = ir * rs;
      u = fminf(fdiv(u - pl, pr), 0x1.fffffep-1f);
      p *= pr;
      node = f + 1;
      next = nr.left;
      right = nr.right;
    }
    if (right < 0) return node;
    f = next;
  }
  return -1;
}

// Probability that lt_sample picks `leaf` from x: the walk up the parents,
// each step reading the pair that holds the current node.
__device__ float lt_pdf(const SceneDev& s, int band, f3 x, int leaf) {
  float p = 1.0f;
  int c = leaf;
  for (int level = 0; c > 0; ++level) {
    if (level >= 64) return 0.0f;  // (bounded like lt_sample)
    const int f = c & ~1;
    const LightNode nl = s.lnodes[f], nr = s.lnodes[f + 1];
    const float2 P = lt_power_pair(s, band, f);
    const float il = lt_importance(nl, P.x, x), ir = lt_importance(nr, P.y, x);
    const float sum = il + ir;
    if (!(sum > 0.0f)) return 0.0f;
    p *= (c == f ? il : ir) * frcp(sum);  // the pick probability exactly as lt_sample forms it
    c = c == f ? nl.parent : nr.parent;
  }
  return p;
}

__device__ __forceinline__ int lt_leaf_of(const SceneDev& s, int face) {
  int lo = 0, hi = s.n_lkeys;
  while (lo < hi) {
    const int m = (lo + hi) >> 1;
    if (s.lkeys[m].face < face) lo = m + 1;
    else hi = m;
  }
  return (lo < s.n_lkeys && s.lkeys[lo].face == face) ? s.lkeys[lo].leaf : -1;
}

// sample_direct_light, main.py:1296-1326, up to the shadow walk: pick a
// light and a point on it with the uniform block (seg, slot).
//
// The walk is not traced when its result cannot reach the image: with a
// throughput of exactly 0 the contribution tp * emission * factor * weight
// is 0 (or NaN, which is dropped) whatever the transmittance; and from a
// medium whose sigma_t = 4 pi k / lambda + sigma_s is dense (shadow_dense)
// the first step's attenuation exp(-sigma_t t), t > EPSILON, is below the
// 1e-6 that sample_direct_light requires (main.py:1329).  sigma_vol is the
// path medium's sigma_t in the active band (0 when unknown).
__device__ __forceinline__ bool shadow_dense(float sigma_t) {
  // exp(-sigma_t * t) <= exp(-14.5) = 5.0e-7 < 1e-6 for every t > EPSILON
  return sigma_t * SR_EPSILON >= 14.5f;
}

template <bool ADJ>
__device__ __forceinline__ void nee_setup(Lane& L, const Ctx<ADJ>& c, int surf, int seg, int slot,
                                          float sigma_vol) {
  const Job& a = c.a;
  const SceneDev& s = a.s;
  const int ab = L.active;
  L.nee_surf = surf;
  const int nl = c.t.light_count[ab];
  if (nl <= 0 || L.tp == 0.0f || shadow_dense(sigma_vol)) {
    L.fdirect = 0.0f;
    L.sdir = L.pos;
    L.phase = PH_FIN;
    return;
  }
  const U4 rn = draws(a, L, seg, slot);
  int lslot, lprim;
  if (a.nee_tree) {
    float p;
    const int leaf = lt_sample(s, ab, L.pos, rn.u[0], p);
    if (leaf < 0) {  // every light of the band faces away from this point
      L.fdirect = 0.0f;
      L.sdir = L.pos;
      L.phase = PH_FIN;
      return;
    }
    lslot = ~s.lnodes[leaf].left;
    lprim = s.lnodes[leaf].face;
    L.fdirect = s.shade[(size_t)lprim * 4].w / p;
  } else {
    const LightRec* lr = c.t.lights + (size_t)ab * s.max_lights;
    float r = rn.u[0];
    // first i with r < cdf[i] (main.py:1310-1313) by binary search over the
    // non-decreasing CDF: the same light as the linear scan, O(log n) reads
    // (0 if none, the scan's fallback)
    int lo = 0, hi = nl;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (r < lr[mid].cdf) hi = mid;
      else lo = mid + 1;
    }
    const int e = lo < nl ? lo : 0;
    const LightRec le = lr[e];
    lslot = le.slot;
    lprim = le.prim;
    L.fdirect = c.t.light_area[ab];
  }
  const float4* lv = c.t.lverts + (size_t)lslot * 3;
  f3 v0 = xyz(lv[0]), v1 = xyz(lv[1]), v2 = xyz(lv[2]);
  float bx = rn.u[1], by = rn.u[2];
  if (bx + by > 1.0f) { bx = 1.0f - bx; by = 1.0f - by; }
  f3 p = v0 + bx * (v1 - v0) + by * (v2 - v0);
  L.sdir = normalize(p - L.pos);
  L.smax = length(p - L.pos);
  L.sprim = lprim;
  L.svol = L.vol;
  L.strav = 0.0f;
  L.stot = 1.0f;
  L.soff = 0.0f;
  L.sstep = 0;
  if (ADJ) L.stot_dsig = 0.0f;
  L.phase = PH_SHD;
}

// One step of a subsurface random walk inside material wmat (sigma_s, the
// absorption 4 pi k / lambda and the HG g of its band, as the reference's
// opaque branch forms them, main.py:1057-1058): a free flight of length
// t_k; a boundary hit within it is the exit point, which acts as a
// Lambertian vertex (NEE, then a cosine lobe about the outward normal);
// otherwise an HG scatter with weight sigma_s / sigma_t.  Step k's uniform
// block (Philox slot SR_SLOT_WALK, counter seg << 9 | k) gives t_k and, for
// k > 0, the direction leaving scatter k - 1, so t_k is known before step
// k's trace: that trace only looks for hits up to t_k (t_init just above
// it), which cuts most of the traversal deep inside the medium, and "no hit
// reported" means "scatter at t_k".  A walk ends after kMaxWalk steps (in a
// closed mesh it always meets the boundary first in practice; a walker that
// has left an open mesh is ended by that limit).
constexpr int kMaxWalk = 256;

template <bool ADJ>
__device__ __forceinline__ void walk_step(Lane& L, const Ctx<ADJ>& c, const HitD& h) {
  const Job& a = c.a;
  const SceneDev& s = a.s;
  L.n_trav++;
  if (h.prim >= 0) {  // the boundary within the flight: the exit point
    L.pos = L.pos + L.dir * h.t;
    const SurfHit sh = surface_of(s, h, L.dir);
    const f3 no = dot(sh.normal, L.dir) > 0.0f ? sh.normal : -sh.normal;
    L.en = no;
    L.esn = no;
    L.efront = 1;
    L.eaction = 1;
    L.walk = 2;
    nee_setup<ADJ>(L, c, -2, (L.seg << 9) | L.wstep, SR_SLOT_WALK_NEE, 0.0f);
    return;
  }
  L.pos = L.pos + L.dir * L.smax;
  L.tp *= L.strav;
  if (++L.wstep >= kMaxWalk) {
    end_path<ADJ>(L, c);
    return;
  }
  const U4 u = draws(a, L, (L.seg << 9) | L.wstep, SR_SLOT_WALK);
  L.dir = hg_sample(L.dir, L.soff, u.u[1], u.u[2]);
  L.smax = walk_flight(u.u[0], L.stot);
  L.phase = PH_EXT;
}

// The lane's medium in its active band and its absorption 4 pi k / lambda.
template <bool ADJ>
__device__ __forceinline__ MatRec ext_medium(const Lane& L, const Ctx<ADJ>& c, float& sa) {
  const MatRec vm = c.t.mat(L.vol, L.active);
  sa = fdiv(4.0f * SR_PI * vm.k, c.t.lam_m[L.active]);
  return vm;
}
// Free-flight distance of scatter_volume (main.py:1203-1204), drawn from the
// segment's EXT block before its trace (kept in L.dvol): the trace then needs
// the exact closest hit only up to dvol + EPSILON (trace_closest's t_near).
template <bool ADJ>
__device__ __forceinline__ float ext_flight(const Lane& L, const Ctx<ADJ>& c) {
  float sa;
  const MatRec vm = ext_medium<ADJ>(L, c, sa);
  const U4 re4 = draws(c.a, L, L.seg, SR_SLOT_EXT);
  return fdiv(-logf(-re4.u[0] + 1.0f - 1e-7f), vm.ss + sa + 1e-6f);
}

// Extension ray came back: scatter_volume (main.py:1190-1292) and the
// cast_rays event dispatch (main.py:1540-1578) up to the NEE shadow walk.
template <bool ADJ, bool SSS>
__device__ __forceinline__ void on_extension(Lane& L, const Ctx<ADJ>& c, const HitD& h) {
  if (SSS && L.walk == 1) {
    walk_step<ADJ>(L, c, h);
    return;
  }
  const Job& a = c.a;
  const SceneDev& s = a.s;
  L.n_seg++;
  L.n_trav++;
  const int ab = L.active;
  float sa;
  const MatRec vm = ext_medium<ADJ>(L, c, sa);
  const float dvol = L.dvol;
  // h is exact for t <= dvol + EPSILON; beyond that only "some hit before
  // ESCAPE" (prim -2, t just above dvol + EPSILON) or "none" (t = ESCAPE)
  float ht = h.t;
  f3 hn = mk(0.f, 0.f, 0.f), hsn = mk(0.f, 0.f, 0.f);
  int hfront = 0, hmat = 0;
  if (h.prim >= 0) {
    SurfHit sh = surface_of(s, h, L.dir);
    hn = sh.normal;
    hsn = sh.sn;
    hfront = sh.front;
    hmat = sh.mat;
  }
  float travel = -1.0f;
  if (ht < SR_ESCAPE) {
    travel = ht;
    if (dvol + SR_EPSILON < travel) travel = dvol;
  }
  const float ff = (float)hfront;
  L.en = hn * ff - hn * (1.0f - ff);
  L.esn = hsn * ff - hsn * (1.0f - ff);
  L.efront = hfront;
  float T = expf(-(sa + vm.ss) * travel);
  float emit = fdiv((-T + 1.0f) * vm.em, sa + 1e-6f);
  float atten = T + emit;
  float est = 0.0f;
  float em_new = 0.0f;
  bool absorb_branch = false;
  L.eaction = 0;
  L.emat = 0;
  if (travel == ht) {
    est = atten / (1.0f + 1e-6f);
    L.eaction = 1;
    L.emat = hmat;
  } else if (travel >= 0.0f) {
    L.emat = -1;
    // a volume event: the rest of the segment's EXT block (its first word
    // drew dvol before the trace); surface hits and escapes never need it
    const U4 re4 = draws(a, L, L.seg, SR_SLOT_EXT);
    float re = re4.u[1];
    float st = fdiv(sa, vm.ss + sa + 1e-6f);
    if (re <= st) {
      absorb_branch = true;
      const int eid = c.t.mat_eem[L.vol];
      if (eid > 0) {
        const float* row = s.eem + ((size_t)eid * s.B + ab) * s.B;
        float qy = 0.0f;
        for (int j = 0; j < s.B; ++j) qy = row[j] > qy ? row[j] : qy;
        float rf = re4.u[2];
        if (qy > 0.0f && rf < qy) {
          L.eaction = 3;
          float norm = 0.0f;
          for (int j = 0; j < s.B; ++j) norm += row[j];
          float r2 = re4.u[3];
          float cum = 0.0f;
          int nw = ab;
          const float rn = frcp(norm);
          for (int j = 0; j < s.B; ++j) {
            cum += row[j] * rn;
            if (r2 < cum) { nw = j; break; }
          }
          L.enew = nw;
          const U4 rx = draws(a, L, L.seg, SR_SLOT_AUX);
          L.edir = random_in_unit_sphere(rx.u[2], rx.u[3]);
          em_new = c.t.mat(L.vol, nw).em;
          float em2 = fdiv((-(1.0f - T) + 1.0f) * em_new, sa + 1e-6f);
          est = (1.0f - T) + em2;
        } else {
          est = atten;
        }
      } else {
        est = atten / (1.0f + 1e-6f);
      }
    } else {
      L.eaction = 2;
      L.edir = hg_sample(L.dir, vm.g, re4.u[2], re4.u[3]);
      est = atten;
    }
  } else {
    end_path<ADJ>(L, c);  // escaped (main.py:1541-1543)
    return;
  }
  if (ADJ) L.hg_sc = 0.0f;
  if (ADJ && L.pass == 1) {
    const AdjMap& m = c.adj;
    const float R = L.Lw - L.Eb;  // everything from this segment on carries est
    if (L.vol == m.sig_mat) {
      Dl sad = dvar(sa, SL_SIG, true);
      Dl Td = dexp((sad + vm.ss) * (-travel));
      Dl attd = Td + ((1.0f - Td) * vm.em) / (sad + 1e-6f);
      Dl estd;
      const int eid = c.t.mat_eem[L.vol];
      if (L.eaction == 1) estd = attd / (1.0f + 1e-6f);
      else if (L.eaction == 3) estd = (1.0f - Td) + (Td * em_new) / (sad + 1e-6f);
      else if (L.eaction == 2 || eid > 0) estd = attd;
      else estd = attd / (1.0f + 1e-6f);
      float dl[SL_N];
      dlog_into(estd, dl);
      const float st = vm.ss + sa + 1e-6f;
      float score;
      if (L.eaction == 1) {
        score = -(ht - SR_EPSILON);  // P(no volume event before the surface)
      } else {
        score = 1.0f / st - travel;  // free-flight density
        score += absorb_branch ? (sa > 0.f ? 1.0f / sa - 1.0f / st : 0.0f) : -1.0f / st;
      }
      if (m.flags & SR_ADJ_SKIP_SIG_SCORE) score = 0.0f;
      if (m.flags & SR_ADJ_SKIP_SIG_FACTOR) dl[SL_SIG] = 0.0f;
      emit_grad(c, m.sig_off + ab, R * (dl[SL_SIG] + score));
    }
    // the HG direction only affects what follows this segment's NEE term:
    // its score is applied in finish_segment, once that term is known
    L.hg_sc = (L.eaction == 2 && L.vol == m.g_mat && !(m.flags & SR_ADJ_SKIP_HG_SCORE))
                  ? hg_score_g(dot(L.edir, L.dir), vm.g)
                  : 0.0f;
  }
  L.pos = L.pos + L.dir * travel;
  L.tp *= est;
  int surf = -1;
  if (L.eaction == 1) {
    MatRec smt = c.t.mat(L.emat, L.band);
    if (smt.se) {  // direct hit on an emitter (main.py:1550-1559)
      float d2 = travel * travel;
      float cl = tmax(0.0f, dot(hn, -L.dir));
      float lden = cl * c.t.light_area[L.active] + 1e-6f;
      if (a.nee_tree && cl > 0.0f && L.pl > 0) {
        // the light tree's pick probability at the previous NEE vertex; a
        // face the band's NEE cannot pick keeps the reference's weight
        const int leaf = lt_leaf_of(s, h.prim);
        if (leaf >= 0 && s.lpower[(size_t)L.active * s.n_lnodes + leaf] > 0.0f) {
          const float p = lt_pdf(s, L.active, L.sdir, leaf);
          if (p > 0.0f) lden = cl * (s.shade[(size_t)h.prim * 4].w / p) + 1e-6f;
        }
      }
      float pdfl = fdiv(d2 * tmin(1.0f, (float)L.pl), lden);
      float den = pdfl * pdfl + L.sp * L.sp + 1e-6f;
      float w = fdiv(L.sp * L.sp, den);
      float contrib = L.tp * smt.em * w;
      if (ADJ && L.pass == 1 && L.sp_dr != 0.0f && isfinite(contrib)) {
        float dlogw = 2.0f / L.sp - 2.0f * L.sp / den;
        emit_grad(c, c.adj.rough_idx, dacc(c, L, L.active) * contrib * dlogw * L.sp_dr);
      }
      add_contrib<ADJ>(L, c, contrib);
      end_path<ADJ>(L, c);
      return;
    }
    surf = L.emat;
  }
  nee_setup<ADJ>(L, c, surf, L.seg, SR_SLOT_NEE, sa + vm.ss);
}

// Shadow ray came back: one step of shadow_transmittance (main.py:1456-1498);
// on the last step, the rest of sample_direct_light (main.py:1329-1356).
template <bool ADJ, bool SSS>
__device__ __forceinline__ void on_shadow(Lane& L, const Ctx<ADJ>& c, const HitD& h) {
  const Job& a = c.a;
  const SceneDev& s = a.s;
  L.n_trav++;
  const int ab = L.active;
  MatRec vm = c.t.mat(L.svol, ab);
  float absorb = fdiv(4.0f * SR_PI * vm.k, c.t.lam_m[ab]);
  const bool sig_span = ADJ && L.svol == c.adj.sig_mat;
  L.stot *= expf(-(absorb + vm.ss) * h.t);
  if (sig_span) L.stot_dsig -= h.t;
  bool done = false;
  if (h.t >= (L.smax - L.strav) || h.prim == L.sprim) {
    L.stot *= expf(-(absorb + vm.ss) * (L.smax - L.strav));
    if (sig_span) L.stot_dsig -= (L.smax - L.strav);
    done = true;
  } else {
    SurfHit sh = surface_of(s, h, L.sdir);
    MatRec sm = c.t.mat(sh.mat, ab);
    float n_i = vm.n;
    float n_t = sh.front ? sm.n : c.t.mat(0, ab).n;
    const float ff = (float)sh.front;
    f3 inc = sh.sn * ff - sh.sn * (1.0f - ff);
    f3 refr = refract(L.sdir, inc, n_i / n_t);
    float rt = (dot(L.sdir, refr) - 1.0f) * 37.9743f;
    float af = frcp(1.0f + rt * rt * rt * rt * (3.0f - 2.0f * rt));
    L.stot *= af;
    if (L.stot < 1e-6f || !sm.tv) {
      L.stot = 0.0f;
      done = true;
    } else {
      L.strav += h.t + SR_EPSILON;
      L.soff = h.t + SR_EPSILON;
      L.svol = sh.mat * sh.front;
      L.sstep += 1;
      if (L.sstep >= SR_SCATTER_BOUNCES) {
        done = true;
      } else if (L.smax - L.strav >= 0.0f && !(ADJ && L.svol == c.adj.sig_mat)) {
        // The next step would start in medium svol.  If it is dense, that
        // step's exp(-sigma_t t) (t > EPSILON) and, when the light is
        // reached, exp(-sigma_t (smax - strav)) <= 1 leave stot <= 5e-7: no
        // contribution either way (main.py:1463-1469, 1329), so it is not traced.
        const MatRec nv = c.t.mat(L.svol, ab);
        if (shadow_dense(fdiv(4.0f * SR_PI * nv.k, c.t.lam_m[ab]) + nv.ss)) {
          L.stot = 0.0f;
          done = true;
        }
      }
    }
  }
  if (!done) return;
  const float aeff = L.fdirect;
  float direct = 0.0f;
  if (L.stot > 1e-6f) {
    const float4* shl = s.shade + (size_t)L.sprim * 4;
    f3 ln = xyz(shl[0]);
    int lmat = __float_as_int(shl[1].w);
    float cl = tmax(0.0f, dot(ln, -L.sdir));
    float d2 = L.smax * L.smax;
    float gt = fdiv(cl, d2 + 1e-4f);
    float bsdf, cosv, pdf;
    if (SSS && L.nee_surf == -2) {  // exit point of a subsurface walk: Lambertian
      cosv = tmax(0.0f, dot(L.esn, L.sdir));
      bsdf = SR_INV_PI;
      pdf = cosv * SR_INV_PI;
    } else if (L.nee_surf != -1) {
      cosv = tmax(0.0f, dot(L.esn, L.sdir));
      evaluate_bsdf(c.t, L.vol, L.nee_surf, -L.dir, L.sdir, L.esn, ab, bsdf, pdf, SSS && c.a.sss);
    } else {
      cosv = tmax(0.0f, dot(L.dir, L.sdir));
      bsdf = hg_eval(cosv, c.t.mat(L.vol, ab).g);
      pdf = bsdf;
    }
    float factor = bsdf * cosv * L.stot * gt * aeff;
    float emis = c.t.mat(lmat, L.band).em;
    float pdfl = fdiv(d2, tmax(1e-6f, cl) * aeff);
    float w = fdiv(pdfl * pdfl, pdfl * pdfl + pdf * pdf + 1e-6f);
    direct = L.tp * emis * factor * w;
    if (ADJ && L.pass == 1 && isfinite(direct) && direct != 0.0f) {
      const AdjMap& m = c.adj;
      Dl bd, pd;
      if (L.nee_surf != -1) {
        evaluate_bsdf_d(c.t, m, L.vol, L.nee_surf, -L.dir, L.sdir, L.esn, ab, bd, pd);
      } else {
        bd = hg_eval_d(cosv, dvar(c.t.mat(L.vol, ab).g, SL_G, L.vol == m.g_mat));
        pd = bd;
      }
      Dl wd = dc(pdfl * pdfl) / ((pd * pd) + (pdfl * pdfl + 1e-6f));
      float lb[SL_N], lw[SL_N];
      dlog_into(bd, lb);
      dlog_into(wd, lw);
      // accumulated into the band after this segment's update (main.py:1586-1590)
      const int band_acc = L.eaction == 3 ? L.enew : ab;
      const float dv = dacc(c, L, band_acc) * direct;
      if (L.nee_surf != -1 && m.alb_off[L.nee_surf] >= 0)
        emit_grad(c, m.alb_off[L.nee_surf] + ab, dv * (lb[SL_ALB] + lw[SL_ALB]));
      if (L.nee_surf == m.rough_mat) emit_grad(c, m.rough_idx, dv * (lb[SL_ROUGH] + lw[SL_ROUGH]));
      if (L.nee_surf == -1 && L.vol == m.g_mat && !(m.flags & SR_ADJ_SKIP_G_NEE))
        emit_grad(c, m.g_idx, dv * (lb[SL_G] + lw[SL_G]));
      if (m.sig_mat >= 0 && !(m.flags & SR_ADJ_SKIP_SIG_FACTOR)) emit_grad(c, m.sig_off + ab, dv * L.stot_dsig);
    }
  }
  // finish_segment runs once per loop iteration for every lane that reaches
  // it (from either handler), so its large scatter_surface body is shared
  L.fdirect = direct;
  L.sdir = L.pos;
  L.phase = PH_FIN;
}

// A work item is a run of up to SR_BANDS_PER_ITEM single-band paths of one
// pixel and block: (pixel, block, band group).  Within a block the items run
// tile by tile (8x8 pixels) and, in a tile, group by group, so the 64
// consecutive items a wave claims are one tile's pixels (coherent camera
// rays); an image whose sides are not multiples of 8 runs group by group
// over whole columns.  A lane runs its item's bands back to back (the same
// pixel, so its rays stay close to the last ones), and items are small
// enough that the persistent grid still balances at the end of a launch.
__device__ __forceinline__ uint32_t band_groups(const Job& a) {
  return ((uint32_t)a.s.B + SR_BANDS_PER_ITEM - 1) / SR_BANDS_PER_ITEM;
}
__device__ __forceinline__ void item_to_path(const Job& a, uint32_t item, int& pi, int& pj, int& blk, int& band,
                                             int& band_end) {
  const uint32_t npix = (uint32_t)a.W * (uint32_t)a.H;
  const uint32_t G = band_groups(a);
  const uint32_t per_block = npix * G;
  blk = a.first_block + (int)(item / per_block);
  const uint32_t r = item % per_block;
  uint32_t g;
  if (a.tiled) {
    const uint32_t tile = r / (64u * G), rem = r % (64u * G);
    const uint32_t w = rem & 63u;
    g = rem >> 6;
    const uint32_t tx = tile % (uint32_t)(a.W >> 3), ty = tile / (uint32_t)(a.W >> 3);
    pi = (int)(tx * 8 + (w & 7u));
    pj = (int)(ty * 8 + (w >> 3));
  } else {
    g = r / npix;
    const uint32_t p = r % npix;
    pi = (int)(p / (uint32_t)a.H);
    pj = (int)(p % (uint32_t)a.H);
  }
  band = (int)(g * SR_BANDS_PER_ITEM);
  band_end = min(band + SR_BANDS_PER_ITEM, a.s.B);
}

// Item k of job J's queue q (k < q_len[q] * n_blocks): the queue's share of
// every spectral block, block after block.
__device__ __forceinline__ uint32_t queue_item(const Job& J, int q, uint32_t k) {
  const uint32_t len = J.q_len[q];
  const uint32_t b = k / len;
  return b * J.per_block + J.q_off[q] + (k - b * len);
}

// Per-wave flush of the lane counters into job j's slot.
__device__ __forceinline__ void flush_counters(const PathArgs& P, int j, Lane& L, int lane) {
  unsigned long long seg = L.n_seg, trav = L.n_trav, nf = L.n_nf;
  for (int off = 32; off >= 1; off >>= 1) {
    seg += __shfl_xor(seg, off);
    trav += __shfl_xor(trav, off);
    nf += __shfl_xor(nf, off);
  }
  if (lane == 0 && (seg | trav | nf)) {
    atomicAdd(P.counters + 4 * j + 0, seg);
    atomicAdd(P.counters + 4 * j + 1, trav);
    atomicAdd(P.counters + 4 * j + 2, nf);
  }
  L.n_seg = L.n_trav = L.n_nf = 0;
}

// The walk limits of the lane's ray: extension rays report the exact
// closest hit up to the free flight (t_near) and otherwise only whether a
// surface lies before ESCAPE; shadow steps may start their search at the
// light when the medium cannot attenuate.
template <bool ADJ, bool SSS>
__device__ __forceinline__ void ray_limits(Lane& L, const Ctx<ADJ>& c, const PathArgs& P, float& t_init,
                                           float& t_near) {
  t_init = 1e9f;
  t_near = INFINITY;
  if (L.phase == PH_EXT) {
    if (SSS && L.walk == 1) {
      t_init = nextafterf(L.smax, INFINITY);  // hits at t <= the flight
    } else {
      // a path segment: hits at or beyond ESCAPE mean "escaped"
      // (main.py:1210), and past the free flight only their existence counts
      t_init = SR_ESCAPE;
      L.dvol = ext_flight<ADJ>(L, c);
      t_near = L.dvol + SR_EPSILON;
    }
  } else {
    // A shadow step through a medium with sigma_a + sigma_s = 0 needs only
    // the hits before the light: any hit at t >= smax - strav ends the walk
    // as "reached" with an attenuation of exactly 1 whatever its t
    // (main.py:1463-1468), so the search starts at that distance.  (Not for
    // the adjoint's sigma_a material: its span derivative uses t.)  (k = 0
    // and sigma_s = 0 make that sum exactly 0.)
    const SceneDev& s = c.a.s;
    const float4 mv = *reinterpret_cast<const float4*>(c.t.mats + L.svol * s.B + L.active);
    if (mv.y == 0.0f && mv.w == 0.0f && !(ADJ && L.svol == P.adj.sig_mat)) t_init = L.smax - L.strav;
  }
}

// REFILL picks the wave's traversal discipline (see the loop body); LT
// reads the scene's small tables from their LDS copy at `tab` (a single-job
// launch) instead of global memory.
// REC: the job's path totals (Job::rec) are written (forward) or read
// (adjoint); without it the kernel has no trace of them.
template <bool ADJ, bool SSS, bool REFILL, bool LT, bool REC = false>
__device__ __forceinline__ void path_loop(const PathArgs& P, int* stk, float* lds_grad, const char* tab) {
  const int stride = blockDim.x;
  const int lane = (int)__lane_id();
  // wave-uniform job cursor: a wave takes items from one job until that
  // job's queues are dry and none of its lanes still runs a path of it.
  // Items come in chunks of P.chunk per atomic (a wave-local range
  // [lbase, lend) of queue lq).  The wave claims from its XCD's queue
  // first (q0: blocks are dealt to the XCDs in turn), then, each time the
  // queue it claims from runs dry, from the next one; gdry once all are.
  const int q0 = (int)(blockIdx.x % SR_QUEUES);
  int cur = 0;
  int ext_streak = 0;
  bool gdry = false;
  int qn = 0, lq = 0;
  uint32_t lbase = 0, lend = 0;
  Lane L;
  L.phase = PH_PARK;  // waiting for a work item
  L.band = 0;
  L.pi = L.pj = L.blk = 0;
  L.acc = 0.0f;
  L.acc_band = 0;
  L.n_seg = L.n_trav = L.n_nf = 0;
  L.pass = 0;
  L.Lw = L.Eb = L.sp_dr = L.stot_dsig = L.hg_sc = 0.0f;
  L.fdirect = 0.0f;
  L.dvol = 0.0f;
  bool has_item = false;
  int band_end = 0;  // the lane's item covers bands [.., band_end)
  int tr = 0;  // the lane's ray: 0 not started, 1 walking, 2 hit known
  TraceState ts;
  trace_begin(ts, 0.0f, 0.0f);
  // The job table is written only by the host: read it through the constant
  // address space so every field access is a (cached) scalar load.
  typedef __attribute__((address_space(4))) const Job ConstJob;
  const ConstJob* jobs = (const ConstJob*)P.jobs;
  while (true) {
    cur = __builtin_amdgcn_readfirstlane(cur);
    const Job& J = *(const Job*)(jobs + cur);
    const Ctx<ADJ> c{J, P.adj, lds_grad, LT ? lds_tabs(J.s, J.tab, tab) : global_tabs(J.s),
                     REC ? J.rec : nullptr};
    if (L.phase == PH_NEW) {  // a path ended: the item's next band, if any
      L.band += 1;
      L.pass = 0;
      if (L.band < band_end) L.phase = PH_NEXT;
    }
    {  // items whose last path ended: one count per wave, the job's completion stamped by the last
      const bool fin = L.phase == PH_NEW && has_item;
      const unsigned long long fmask = __ballot(fin);
      if (fmask) {
        const int leader = __ffsll((long long)fmask) - 1;
        if (lane == leader) {
          const uint32_t k = (uint32_t)__popcll(fmask);
          if (atomicAdd(P.done_items + cur, k) + k == J.n_items) P.clocks[1 + cur] = wall_clock64();
        }
      }
      if (L.phase == PH_NEW) {
        L.pass = 0;
        has_item = false;
        L.phase = PH_PARK;
      }
    }
    const bool need = L.phase == PH_PARK;
    const unsigned long long mask = __ballot(need);
    if (mask && !(gdry && lbase >= lend)) {
      const uint32_t n = (uint32_t)__popcll(mask);
      const uint32_t left = lend - lbase;
      uint32_t nb = 0, nend = 0;
      const int nq = (q0 + qn) % SR_QUEUES;
      if (left < n && !gdry) {
        const uint32_t qitems = J.q_len[nq] * (uint32_t)J.n_blocks;
        const int leader = __ffsll((long long)mask) - 1;
        uint32_t got = 0;
        if (lane == leader) got = atomicAdd(P.queues + cur * SR_QUEUES + nq, P.chunk);
        nb = __builtin_amdgcn_readlane(got, leader);
        nend = nb >= qitems ? nb : min(nb + P.chunk, qitems);
        if (nb + P.chunk >= qitems) {
          qn += 1;
          gdry = qn >= SR_QUEUES;
        }
      }
      if (need) {
        const uint32_t r =
            __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
        uint32_t item = 0xFFFFFFFFu;
        if (r < left) item = queue_item(J, lq, lbase + r);
        else if (nb + (r - left) < nend) item = queue_item(J, nq, nb + (r - left));
        if (item != 0xFFFFFFFFu) {
          item_to_path(J, item, L.pi, L.pj, L.blk, L.band, band_end);
          L.pass = 0;
          has_item = true;
          L.phase = PH_NEXT;
        }
      }
      if (left >= n) {
        lbase += n;
      } else {
        lbase = min(nb + (n - left), nend);
        lend = nend;
        lq = nq;
      }
    }
    if (L.phase == PH_NEXT || L.phase == PH_REPLAY) start_path<ADJ>(L, c);
    if (__ballot(L.phase != PH_PARK) == 0ull) {
      // every lane is parked: the job's queue is dry for this wave and none
      // of its lanes still runs a path of it -- move the wave to the next job
      if (!(gdry && lbase >= lend)) continue;
      flush_counters(P, cur, L, lane);
      if (cur + 1 >= P.n_jobs) break;
      cur += 1;
      if (LT) {
        // The block's waves move to the next job together: each arrives here
        // once it is done with the last one, and the job's tables then
        // replace the last job's in LDS (copy_tables ends with a barrier).
        __syncthreads();
        const Job& Jn = *(const Job*)(jobs + cur);
        copy_tables(Jn.s, Jn.tab, const_cast<char*>(tab));
      }
      gdry = false;
      qn = lq = 0;
      lbase = lend = 0;
      continue;
    }
    bool do_ext;
    if (REFILL) {
      // Rays.  A lane in PH_EXT or PH_SHD owns one ray: tr is 0 before its
      // walk starts, 1 while it walks the BVH, 2 once its hit is known.
      // Every walking lane of the wave walks together, extension and shadow
      // rays alike, until fewer than SR_TRACE_KEEP of them still walk; a
      // walk that outlasts the iteration keeps its state (and its LDS stack
      // column) and goes on in the next, beside the rays that the finished
      // lanes start after their handlers.
      if ((L.phase == PH_EXT || L.phase == PH_SHD) && tr == 0) {
        float t_init, t_near;
        ray_limits<ADJ, SSS>(L, c, P, t_init, t_near);
        trace_begin(ts, t_init, t_near);
        tr = 1;
      }
      if (__ballot(tr == 1)) {
        const bool ext = L.phase == PH_EXT;
        const f3 ro = ext ? L.pos : L.pos + L.sdir * L.soff;
        const f3 rd = ext ? L.dir : L.sdir;
        if (trace_resume(J.s.nodes, J.s.tris, ro, rd, stk, stride, P.overflow, ts, tr == 1, SR_TRACE_KEEP)) tr = 2;
      }
      // Every lane whose ray is done shades in this iteration: the
      // extension handler (scatter_volume + NEE set-up) for the extension
      // rays, then the shadow-walk handler (walk step) for the shadow rays,
      // then the end of segment (incl. scatter_surface) for the lanes whose
      // walk is over.  Waiting a ready group out for a larger one (the
      // walk-to-the-end loop below) only delays lanes here: the walks do
      // not wait for the handlers, so a lane held back is a lane idle.
      if (__ballot(tr == 2 || L.phase == PH_FIN) == 0ull) continue;
      // A handler pass takes as long whatever the number of lanes it
      // serves.  With the scene tables in LDS (LT) the passes are short and
      // every ready lane shades now; with them in global memory (many
      // lights: the per-band light records outgrow the LDS share) each pass
      // waits on those loads, and the smaller ready group waits for the next
      // iteration unless it has 8 lanes (128 and 1024 lamps over a floor:
      // 10-15% less time per block, the 1282-light emitting sphere
      // unchanged; the LDS-table scenes lose 0.2-0.3% with the wait).
      bool run_e = true, run_s = true;
      if (!LT) {
        const int n_er = __popcll(__ballot(tr == 2 && L.phase == PH_EXT));
        const int n_sr = __popcll(__ballot(tr == 2 && L.phase == PH_SHD));
        run_e = n_er >= 8 || n_er >= n_sr;
        run_s = n_sr >= 8 || n_sr > n_er;
      }
      if (run_e && tr == 2 && L.phase == PH_EXT) {
        tr = 0;
        on_extension<ADJ, SSS>(L, c, ts.h);
      }
      if (run_s && tr == 2 && L.phase == PH_SHD) {
        tr = 0;
        on_shadow<ADJ, SSS>(L, c, ts.h);
      }
      do_ext = false;  // (every lane at PH_FIN ends its segment below)
    } else {
      // One handler group per iteration, chosen for the whole wave:
      // extension rays (trace + scatter_volume + NEE set-up) or shadow-walk
      // steps (trace + walk step + end of segment), each ray walked to its
      // end.  Lanes of the other group wait one iteration.  Every path
      // alternates EXT -> SHD -> FIN -> EXT, so the lanes fall into step.
      const int n_ext = __popcll(__ballot(L.phase == PH_EXT));
      const int n_shd = __popcll(__ballot(L.phase == PH_SHD || L.phase == PH_FIN));
      do_ext = n_ext > n_shd && !(ext_streak >= 4 && n_shd > 0);
      ext_streak = do_ext ? ext_streak + 1 : 0;
      if (L.phase == (do_ext ? PH_EXT : PH_SHD)) {
        float t_init, t_near;
        ray_limits<ADJ, SSS>(L, c, P, t_init, t_near);
        const f3 ro = do_ext ? L.pos : L.pos + L.sdir * L.soff;
        const f3 rd = do_ext ? L.dir : L.sdir;
        const HitD h = trace_closest(J.s.nodes, J.s.tris, ro, rd, stk, stride, P.overflow, t_init, t_near);
        if (do_ext) on_extension<ADJ, SSS>(L, c, h);
        else on_shadow<ADJ, SSS>(L, c, h);
      }
    }
    if (!do_ext && L.phase == PH_FIN) finish_segment<ADJ, SSS>(L, c, L.fdirect);
  }
}

// Traversal discipline and occupancy of the three path kernels, chosen by
// measurement on the 1M-triangle heightfield and the volume scenes.  The
// walk state a suspended ray keeps (TraceState, ~9 registers) does not fit
// beside the shading state at 4 waves per SIMD (128 VGPRs: the spills cost
// more than the refill gains), so the kernels run at 3 waves per SIMD
// (168 VGPRs; 12 waves, three 256-lane blocks and their 40-entry LDS stacks
// per CU).  The subsurface kernel keeps the walk-to-the-end loop: its
// random walk gains nothing from the refill.
#ifndef SR_REFILL_PATH
#define SR_REFILL_PATH true
#endif
#ifndef SR_REFILL_SSS
#define SR_REFILL_SSS false
#endif
#ifndef SR_REFILL_ADJ
#define SR_REFILL_ADJ true
#endif
#ifndef SR_WAVES_SSS
#define SR_WAVES_SSS 3
#endif
#ifndef SR_WAVES_ADJ
#define SR_WAVES_ADJ 3
#endif

// LDS of a launch: the lanes' traversal stacks (SR_STACK_SIZE entries per
// lane, column-major), then the scene's small tables (tab.bytes, a single-job
// launch), then the adjoint's gradient table.
template <bool LT, bool REC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3))) void sr_path_kernel(PathArgs a) {
  extern __shared__ int lds_dyn[];
  char* tab = reinterpret_cast<char*>(lds_dyn + SR_STACK_SIZE * blockDim.x);
  if (threadIdx.x == 0) atomicMin(a.clocks, wall_clock64());
  if (LT) copy_tables(a.jobs[0].s, a.jobs[0].tab, tab);
  path_loop<false, false, SR_REFILL_PATH, LT, REC>(a, lds_dyn + threadIdx.x, nullptr, tab);
}

// The same path loop with random-walk subsurface scattering (SR_RENDER_SUBSURFACE).
template <bool LT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(SR_WAVES_SSS, SR_WAVES_SSS))) void sr_path_sss_kernel(PathArgs a) {
  extern __shared__ int lds_dyn[];
  char* tab = reinterpret_cast<char*>(lds_dyn + SR_STACK_SIZE * blockDim.x);
  if (threadIdx.x == 0) atomicMin(a.clocks, wall_clock64());
  if (LT) copy_tables(a.jobs[0].s, a.jobs[0].tab, tab);
  path_loop<false, true, SR_REFILL_SSS, LT>(a, lds_dyn + threadIdx.x, nullptr, tab);
}

template <bool LT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(SR_WAVES_ADJ, SR_WAVES_ADJ))) void sr_adjoint_kernel(PathArgs a) {
  extern __shared__ int lds_dyn[];
  char* tab = reinterpret_cast<char*>(lds_dyn + SR_STACK_SIZE * blockDim.x);
  float* lds_grad = reinterpret_cast<float*>(tab + (LT ? a.tab_bytes : 0u));
  for (int i = threadIdx.x; i < a.adj.n_params; i += blockDim.x) lds_grad[i] = 0.0f;
  if (LT) copy_tables(a.jobs[0].s, a.jobs[0].tab, tab);  // (ends with the barrier the gradient table needs)
  else __syncthreads();
  path_loop<true, false, SR_REFILL_ADJ, LT, true>(a, lds_dyn + threadIdx.x, lds_grad, tab);
  __syncthreads();
  for (int i = threadIdx.x; i < a.adj.n_params; i += blockDim.x) {
    float v = lds_grad[i];
    if (v != 0.0f) atomicAdd(a.adj.grad + i, v * a.adj.scale);
  }
}

__global__ void sr_fill_u32_kernel(uint32_t* __restrict__ p, size_t n, uint32_t v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t step = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += step) p[i] = v;
}

size_t sr_path_lds_bytes(int block, bool adjoint, uint32_t table_bytes) {
  return (size_t)SR_STACK_SIZE * block * sizeof(int) + table_bytes + (adjoint ? SR_MAX_PARAMS * sizeof(float) : 0);
}

size_t sr_lds_table_budget(int block, bool adjoint) {
  const size_t share = (size_t)SR_LDS_TABLE_BUDGET * (size_t)block / 256;
  const size_t grads = adjoint ? SR_MAX_PARAMS * sizeof(float) : 0;
  return share > grads ? share - grads : 0;
}

LdsTables sr_lds_tables(int n_mats, int B, int n_lights, int n_slots, int block, bool adjoint) {
  LdsTables L;
  std::memset(&L, 0, sizeof(L));
  auto a16 = [](uint64_t x) { return (x + 15) & ~(uint64_t)15; };
  uint64_t o = 0;
  L.mats = (uint32_t)o;        o = a16(o + (uint64_t)n_mats * B * sizeof(MatRec));
  L.lights = (uint32_t)o;      o = a16(o + (uint64_t)n_lights * sizeof(LightRec));
  L.lverts = (uint32_t)o;      o = a16(o + (uint64_t)n_slots * 3 * sizeof(float4));
  L.lam_m = (uint32_t)o;       o = a16(o + (uint64_t)B * 4);
  L.lam_nm = (uint32_t)o;      o = a16(o + (uint64_t)B * 4);
  L.light_area = (uint32_t)o;  o = a16(o + (uint64_t)B * 4);
  L.light_count = (uint32_t)o; o = a16(o + (uint64_t)B * 4);
  L.mat_eem = (uint32_t)o;     o = a16(o + (uint64_t)n_mats * 4);
  if (o > sr_lds_table_budget(block, adjoint)) {
    std::memset(&L, 0, sizeof(L));
    return L;
  }
  L.bytes = (uint32_t)o;
  L.n_mats = (uint32_t)(n_mats * B);
  L.n_lights = (uint32_t)n_lights;
  L.n_lverts = (uint32_t)(n_slots * 3);
  return L;
}

static int grid_for(const void* fn, int block, size_t lds, int* grid_out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return -1;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return -1;
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, block, lds) != hipSuccess) return -1;
  if (per_cu < 1) per_cu = 1;
  *grid_out = per_cu * prop.multiProcessorCount;
  return 0;
}

// A persistent grid that fits both instantiations of a kernel (tables in
// global memory, or in LDS up to the whole table budget).
static int grid_both(const void* fn_g, const void* fn_l, int block, bool adjoint, int* grid_out) {
  const size_t table_max = sr_lds_table_budget(block, adjoint);
  int g0 = 0, g1 = 0;
  if (grid_for(fn_g, block, sr_path_lds_bytes(block, adjoint), &g0)) return -1;
  if (grid_for(fn_l, block, sr_path_lds_bytes(block, adjoint, (uint32_t)table_max), &g1)) return -1;
  *grid_out = std::min(g0, g1);
  return 0;
}

int sr_path_grid_bs(int block, int* grid_out) {
  return grid_both((const void*)sr_path_kernel<false, false>, (const void*)sr_path_kernel<true, false>, block, false,
                   grid_out);
}
int sr_path_grid(int* grid_out) { return sr_path_grid_bs(256, grid_out); }
int sr_adjoint_grid(int* grid_out) {
  return grid_both((const void*)sr_adjoint_kernel<false>, (const void*)sr_adjoint_kernel<true>, 256, true, grid_out);
}

int sr_launch_path(const PathArgs& a, int grid, int block, hipStream_t st) {
  const size_t lds = sr_path_lds_bytes(block, false, a.tab_bytes);
  if (a.record) {  // a single job writing its path totals
    if (a.tab_bytes) hipLaunchKernelGGL(HIP_KERNEL_NAME(sr_path_kernel<true, true>), dim3(grid), dim3(block), lds, st, a);
    else hipLaunchKernelGGL(HIP_KERNEL_NAME(sr_path_kernel<false, true>), dim3(grid), dim3(block), lds, st, a);
  } else {
    if (a.tab_bytes) hipLaunchKernelGGL(HIP_KERNEL_NAME(sr_path_kernel<true, false>), dim3(grid), dim3(block), lds, st, a);
    else hipLaunchKernelGGL(HIP_KERNEL_NAME(sr_path_kernel<false, false>), dim3(grid), dim3(block), lds, st, a);
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int sr_path_sss_grid_bs(int block, int* grid_out) {
  return grid_both((const void*)sr_path_sss_kernel<false>, (const void*)sr_path_sss_kernel<true>, block, false,
                   grid_out);
}
int sr_path_sss_grid(int* grid_out) { return sr_path_sss_grid_bs(256, grid_out); }

int sr_launch_path_sss(const PathArgs& a, int grid, int block, hipStream_t st) {
  const size_t lds = sr_path_lds_bytes(block, false, a.tab_bytes);
  if (a.tab_bytes) hipLaunchKernelGGL(sr_path_sss_kernel<true>, dim3(grid), dim3(block), lds, st, a);
  else hipLaunchKernelGGL(sr_path_sss_kernel<false>, dim3(grid), dim3(block), lds, st, a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int sr_launch_adjoint(const PathArgs& a, int grid, int block, hipStream_t st) {
  const size_t lds = sr_path_lds_bytes(block, true, a.tab_bytes);
  if (a.tab_bytes) hipLaunchKernelGGL(sr_adjoint_kernel<true>, dim3(grid), dim3(block), lds, st, a);
  else hipLaunchKernelGGL(sr_adjoint_kernel<false>, dim3(grid), dim3(block), lds, st, a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int sr_launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t st) {
  if (n == 0) return 0;
  size_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(sr_fill_u32_kernel, dim3((unsigned)blocks), dim3(256), 0, st, p, n, v);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
