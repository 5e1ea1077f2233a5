// sr_kernels.hip -- HIP kernels for gfx950: closest-hit batch trace, the
// persistent spectral path loop, the block-mean update and the CMF resolve.
#include <hip/hip_runtime.h>

#include "sr_device.h"
#include "sr_kernels.h"

// ====================================================================== trace
// sr_trace_closest: one lane per ray (the parity entry point).
__global__ __launch_bounds__(256) void sr_trace_kernel(SceneDev s, int n, const float* __restrict__ orig,
                                                       const float* __restrict__ dir, float* __restrict__ out_t,
                                                       int* __restrict__ out_prim, float* __restrict__ out_uv,
                                                       uint32_t* overflow) {
  __shared__ int lds_stack[SR_STACK_SIZE * 256];
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  f3 o = mk(orig[r * 3], orig[r * 3 + 1], orig[r * 3 + 2]);
  f3 d = mk(dir[r * 3], dir[r * 3 + 1], dir[r * 3 + 2]);
  HitD h = trace_closest(s, o, d, lds_stack + threadIdx.x, blockDim.x, overflow);
  out_t[r] = h.t;
  out_prim[r] = h.prim;
  out_uv[r * 2] = h.prim >= 0 ? h.u : 0.0f;
  out_uv[r * 2 + 1] = h.prim >= 0 ? h.v : 0.0f;
}

// ====================================================================== path loop
enum { PH_NEW = 0, PH_EXT = 1, PH_SHD = 2, PH_DONE = 3 };

struct Lane {
  int phase, band, pi, pj, blk;
  Rng rng;
  f3 pos, dir;
  float tp, sp;
  int active, pl, vol, seg;
  float acc;
  int acc_band;
  // pending extension event (scatter_volume result)
  f3 en, esn, edir;
  int efront, emat, eaction, enew, nee_surf;
  // pending shadow walk (shadow_transmittance state)
  f3 sdir;
  float smax, strav, stot, soff;
  int sprim, svol, sstep;
  uint32_t n_seg, n_trav, n_nf;
};

__device__ __forceinline__ void flush_acc(Lane& L, const PathArgs& a) {
  if (L.acc != 0.0f) {
    size_t idx = ((size_t)L.pi * a.H + L.pj) * a.s.B + L.acc_band;
    atomicAdd(a.sum + idx, L.acc);
  }
  L.acc = 0.0f;
}

// block_accumulator[i, j, active] += c  (main.py:1589-1590); non-finite
// contributions are dropped and counted.
__device__ __forceinline__ void add_contrib(Lane& L, const PathArgs& a, float c) {
  if (!isfinite(c)) { L.n_nf++; return; }
  if (c == 0.0f) return;
  if (L.active != L.acc_band) {
    flush_acc(L, a);
    L.acc_band = L.active;
  }
  L.acc += c;
}

__device__ __forceinline__ void end_path(Lane& L, const PathArgs& a) {
  flush_acc(L, a);
  L.phase = PH_NEW;
}

// Camera ray of cast_rays, main.py:1506-1529.
__device__ __forceinline__ void start_path(Lane& L, const PathArgs& a) {
  const SceneDev& s = a.s;
  L.rng.init(a.seed_lo, a.seed_hi, (uint32_t)(L.pi * a.H + L.pj), (uint32_t)L.blk,
             (uint32_t)L.band | (a.stream << 8));
  float u = ((float)L.pi + L.rng.next()) / (float)a.W;
  float v = ((float)L.pj + L.rng.next()) / (float)a.H;
  const float aspect = (float)a.W / (float)a.H;
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
  L.phase = a.depth > 0 ? PH_EXT : PH_NEW;
}

// evaluate_bsdf, main.py:1360-1444 (its transmission branch is unreachable:
// both cosines are clamped to >= 1e-4 before NdotV * NdotL > 0 is tested).
__device__ __forceinline__ void evaluate_bsdf(const SceneDev& s, int vol_id, int surf_id, f3 V, f3 Lv, f3 N, int b,
                                              float& bsdf, float& pdf) {
  MatRec vm = get_mat(s, vol_id, b);
  MatRec sm = get_mat(s, surf_id, b);
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
    bsdf += (D * G * F) / (4.0f * NdotL * NdotV);
    pdf += (D * NdotH * F) / (4.0f * VdotH);
  }
  float absorb = (4.0f * SR_PI * sm.k * 1e9f) / s.lam_nm[b];
  float albedo = sm.ss / (sm.ss + absorb + 1e-6f);
  float on = oren_nayar(sm.r, NdotL, NdotV, Lv, V, N);
  float opq = 1.0f - (float)sm.tv;
  bsdf += (1.0f - F) * opq * (albedo / SR_PI) * on;
  pdf += NdotL * 0.318309886184f * (1.0f - F) * opq;
}

// scatter_surface, main.py:1000-1188, applied to the lane's ray as
// cast_rays does (main.py:1565-1574).
__device__ __forceinline__ void scatter_surface(Lane& L, const PathArgs& a) {
  const SceneDev& s = a.s;
  const int ab = L.active;
  MatRec sm = get_mat(s, L.emat, ab);
  MatRec vm = get_mat(s, L.vol, ab);
  const f3 V = -L.dir;
  const f3 N = L.esn;
  const float ni = vm.n, ki = vm.k, nt = sm.n, kt = sm.k;
  float x0 = L.rng.next();
  float x1 = L.rng.next();
  f3 H = ggx_sample_half(x0, x1, N, sm.r);
  float ia = tmax(1e-4f, dot(V, H));
  float F = fresnel_spectral(ia, ni, ki, nt, kt);
  float rnd = clampf(L.rng.next(), 0.0f, 0.99f);
  int dr = 0;
  f3 ndir;
  float est, pdf;
  bool ggx_reflect = false;
  float cos_abs = 0.0f;
  if (rnd < F) {
    ndir = reflect(L.dir, H);
    dr = 1;
    if (sm.r < 0.008f) { est = F; pdf = 1e9f; }
    else { ggx_reflect = true; cos_abs = 0.0f; }
  } else {
    float absorb = (4.0f * SR_PI * kt) / s.lam_m[ab];
    float dte = -logf(-L.rng.next() + 1.0f + 1e-7f) / (sm.ss + absorb + 1e-6f);
    if (!sm.tv) {
      dr = 1;
      f3 S = cosine_hemisphere(N, L.rng);
      float NdotL = tmax(1e-4f, dot(N, S));
      float NdotV = tmax(1e-4f, dot(N, V));
      float albedo = sm.ss / (sm.ss + absorb + 1e-6f);
      float bsdf = (1.0f - F) * (albedo / SR_PI) * oren_nayar(sm.r, NdotL, NdotV, S, V, N);
      pdf = 0.318309886184f * NdotL;
      const int eid = s.mat_eem[L.emat];
      if (eid > 0) {
        const float* row = s.eem + ((size_t)eid * s.B + ab) * s.B;
        float qy = 0.0f;
        for (int j = 0; j < s.B; ++j) qy = row[j] > qy ? row[j] : qy;
        float r3 = L.rng.next();
        if (qy > 0.0f && r3 < qy) {
          float norm = 0.0f;
          for (int j = 0; j < s.B; ++j) norm += row[j];
          float r4 = L.rng.next();
          float cum = 0.0f;
          int nw = ab;
          for (int j = 0; j < s.B; ++j) {
            cum += row[j] / norm;
            if (r4 < cum) { nw = j; break; }
          }
          MatRec fm = get_mat(s, L.emat, nw);
          float ab2 = (4.0f * SR_PI * fm.k) / s.lam_m[nw];
          bsdf *= expf(-ab2 * dte);
        }
      }
      est = (bsdf * NdotL) / pdf;
      ndir = S;
    } else {
      f3 Rf = refract(L.dir, H, ni / nt);
      if (dot(Rf, Rf) < 1e-4f) {
        ndir = reflect(L.dir, H);
        dr = 1;
        if (sm.r < 0.008f) { est = F; pdf = 1e9f; }
        else { ggx_reflect = true; cos_abs = 1.0f; }
      } else {
        dr = 0;
        ndir = Rf;
        if (sm.r < 0.0018f) {
          est = 1.0f - F;
          pdf = 1e9f;
        } else {
          float LdotH = dot(Rf, H);
          float VdotH = dot(V, H);
          float NdotH = tmax(1e-4f, dot(N, H));
          float NdotL = tmax(1e-4f, fabsf(dot(N, Rf)));
          float NdotV = tmax(1e-4f, fabsf(dot(N, V)));
          float G = ggx_G(N, V, Rf, sm.r);
          float D = ggx_D(NdotH, sm.r);
          float den = ni * VdotH + nt * LdotH;
          float dsq = den * den;
          float t1 = fabsf(LdotH * VdotH) / (NdotL * NdotV);
          float t2 = (nt * nt * (1.0f - F) * D * G) / dsq;
          float bsdf = t1 * t2;
          pdf = (D * NdotH * (1.0f - F) * nt * nt * fabsf(LdotH)) / dsq;
          est = (bsdf * NdotL) / pdf;
        }
      }
    }
  }
  if (ggx_reflect) {  // GGX reflection lobe (main.py:1039-1053 and 1142-1156)
    float NdotL = tmax(1e-4f, dot(N, ndir));
    float NdotH = tmax(1e-4f, dot(N, H));
    float NdotV = tmax(1e-4f, dot(N, V));
    float D = ggx_D(NdotH, sm.r);
    float G = ggx_G(N, V, ndir, sm.r);
    float bsdf = (D * G * F) / (4.0f * NdotL * NdotV);
    pdf = (D * NdotH) / (4.0f * ia);
    float c = cos_abs != 0.0f ? fabsf(dot(N, ndir)) : tmax(0.0f, dot(N, ndir));
    est = (bsdf * c) / pdf;
  }
  L.sp = pdf;
  L.dir = ndir;
  L.tp *= est;
  float fdr = (float)dr;
  L.pos = L.pos + L.en * SR_EPSILON * (fdr - (1.0f - fdr));
  L.vol = L.emat * L.efront * (1 - dr) + L.vol * dr;
}

// End of one path-loop iteration (main.py:1565-1590).
__device__ __forceinline__ void finish_segment(Lane& L, const PathArgs& a, float direct) {
  if (L.eaction == 1) {
    L.pl += 1;
    scatter_surface(L, a);
  } else if (L.eaction == 2) {
    L.dir = L.edir;
    L.pl += 1;
  } else if (L.eaction == 3) {
    L.dir = L.edir;
    L.active = L.enew;
    L.pl += 1;
  }
  add_contrib(L, a, direct);
  L.seg += 1;
  if (L.seg >= a.depth) {
    end_path(L, a);
  } else if (!isfinite(L.tp)) {
    // non-finite throughput (e.g. the 0/0 rough BTDF estimator on leaving a
    // rough dielectric, main.py:1176-1185 with n_i == n_t): end the path
    L.n_nf++;
    end_path(L, a);
  } else {
    L.phase = PH_EXT;
  }
}

// Extension ray came back: scatter_volume (main.py:1190-1292) and the
// cast_rays event dispatch (main.py:1540-1578) up to the NEE shadow walk.
__device__ __forceinline__ void on_extension(Lane& L, const PathArgs& a, const HitD& h) {
  const SceneDev& s = a.s;
  L.n_seg++;
  L.n_trav++;
  const int ab = L.active;
  MatRec vm = get_mat(s, L.vol, ab);
  float sa = (4.0f * SR_PI * vm.k) / s.lam_m[ab];
  float dvol = -logf(-L.rng.next() + 1.0f - 1e-7f) / (vm.ss + sa + 1e-6f);
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
  float emit = ((-T + 1.0f) * vm.em) / (sa + 1e-6f);
  float atten = T + emit;
  float est = 0.0f;
  L.eaction = 0;
  L.emat = 0;
  if (travel == ht) {
    est = atten / (1.0f + 1e-6f);
    L.eaction = 1;
    L.emat = hmat;
  } else if (travel >= 0.0f) {
    L.emat = -1;
    float re = L.rng.next();
    float st = sa / (vm.ss + sa + 1e-6f);
    if (re <= st) {
      const int eid = s.mat_eem[L.vol];
      if (eid > 0) {
        const float* row = s.eem + ((size_t)eid * s.B + ab) * s.B;
        float qy = 0.0f;
        for (int j = 0; j < s.B; ++j) qy = row[j] > qy ? row[j] : qy;
        float rf = L.rng.next();
        if (qy > 0.0f && rf < qy) {
          L.eaction = 3;
          float norm = 0.0f;
          for (int j = 0; j < s.B; ++j) norm += row[j];
          float r2 = L.rng.next();
          float cum = 0.0f;
          int nw = ab;
          for (int j = 0; j < s.B; ++j) {
            cum += row[j] / norm;
            if (r2 < cum) { nw = j; break; }
          }
          L.enew = nw;
          L.edir = random_in_unit_sphere(L.rng);
          float em2 = ((-(1.0f - T) + 1.0f) * get_mat(s, L.vol, nw).em) / (sa + 1e-6f);
          est = (1.0f - T) + em2;
        } else {
          est = atten;
        }
      } else {
        est = atten / (1.0f + 1e-6f);
      }
    } else {
      L.eaction = 2;
      L.edir = hg_sample(L.dir, vm.g, L.rng);
      est = atten;
    }
  } else {
    end_path(L, a);  // escaped (main.py:1541-1543)
    return;
  }
  L.pos = L.pos + L.dir * travel;
  L.tp *= est;
  int surf = -1;
  if (L.eaction == 1) {
    MatRec smt = get_mat(s, L.emat, L.band);
    if (smt.se) {  // direct hit on an emitter (main.py:1550-1559)
      float d2 = travel * travel;
      float cl = tmax(0.0f, dot(hn, -L.dir));
      float pdfl = (d2 * tmin(1.0f, (float)L.pl)) / (cl * s.light_area[L.active] + 1e-6f);
      float w = (L.sp * L.sp) / (pdfl * pdfl + L.sp * L.sp + 1e-6f);
      add_contrib(L, a, L.tp * smt.em * w);
      end_path(L, a);
      return;
    }
    surf = L.emat;
  }
  // sample_direct_light, main.py:1296-1326 (up to the shadow walk)
  L.nee_surf = surf;
  const int nl = s.light_count[ab];
  if (nl <= 0) {
    finish_segment(L, a, 0.0f);
    return;
  }
  const LightRec* lr = s.lights + (size_t)ab * s.max_lights;
  float r = L.rng.next();
  int e = 0;
  for (int i = 0; i < nl; ++i) {
    if (r < lr[i].cdf) { e = i; break; }
  }
  LightRec le = lr[e];
  const float4* lv = s.lverts + (size_t)le.slot * 3;
  f3 v0 = xyz(lv[0]), v1 = xyz(lv[1]), v2 = xyz(lv[2]);
  float bx = L.rng.next(), by = L.rng.next();
  if (bx + by > 1.0f) { bx = 1.0f - bx; by = 1.0f - by; }
  f3 p = v0 + bx * (v1 - v0) + by * (v2 - v0);
  L.sdir = normalize(p - L.pos);
  L.smax = length(p - L.pos);
  L.sprim = le.prim;
  L.svol = L.vol;
  L.strav = 0.0f;
  L.stot = 1.0f;
  L.soff = 0.0f;
  L.sstep = 0;
  L.phase = PH_SHD;
}

// Shadow ray came back: one step of shadow_transmittance (main.py:1456-1498);
// on the last step, the rest of sample_direct_light (main.py:1329-1356).
__device__ __forceinline__ void on_shadow(Lane& L, const PathArgs& a, const HitD& h) {
  const SceneDev& s = a.s;
  L.n_trav++;
  const int ab = L.active;
  MatRec vm = get_mat(s, L.svol, ab);
  float absorb = (4.0f * SR_PI * vm.k) / s.lam_m[ab];
  L.stot *= expf(-(absorb + vm.ss) * h.t);
  bool done = false;
  if (h.t >= (L.smax - L.strav) || h.prim == L.sprim) {
    L.stot *= expf(-(absorb + vm.ss) * (L.smax - L.strav));
    done = true;
  } else {
    SurfHit sh = surface_of(s, h, L.sdir);
    MatRec sm = get_mat(s, sh.mat, ab);
    float n_i = vm.n;
    float n_t = sh.front ? sm.n : get_mat(s, 0, ab).n;
    const float ff = (float)sh.front;
    f3 inc = sh.sn * ff - sh.sn * (1.0f - ff);
    f3 refr = refract(L.sdir, inc, n_i / n_t);
    float rt = (dot(L.sdir, refr) - 1.0f) * 37.9743f;
    float af = 1.0f / (1.0f + rt * rt * rt * rt * (3.0f - 2.0f * rt));
    L.stot *= af;
    if (L.stot < 1e-6f || !sm.tv) {
      L.stot = 0.0f;
      done = true;
    } else {
      L.strav += h.t + SR_EPSILON;
      L.soff = h.t + SR_EPSILON;
      L.svol = sh.mat * sh.front;
      L.sstep += 1;
      if (L.sstep >= SR_SCATTER_BOUNCES) done = true;
    }
  }
  if (!done) return;
  float direct = 0.0f;
  if (L.stot > 1e-6f) {
    const float4* shl = s.shade + (size_t)L.sprim * 4;
    f3 ln = xyz(shl[0]);
    int lmat = __float_as_int(shl[1].w);
    float cl = tmax(0.0f, dot(ln, -L.sdir));
    float d2 = L.smax * L.smax;
    float gt = cl / (d2 + 1e-4f);
    float bsdf, cosv, pdf;
    if (L.nee_surf != -1) {
      cosv = tmax(0.0f, dot(L.esn, L.sdir));
      evaluate_bsdf(s, L.vol, L.nee_surf, -L.dir, L.sdir, L.esn, ab, bsdf, pdf);
    } else {
      cosv = tmax(0.0f, dot(L.dir, L.sdir));
      bsdf = hg_eval(cosv, get_mat(s, L.vol, ab).g);
      pdf = bsdf;
    }
    float factor = bsdf * cosv * L.stot * gt * s.light_area[ab];
    float emis = get_mat(s, lmat, L.band).em;
    float pdfl = d2 / (tmax(1e-6f, cl) * s.light_area[ab]);
    float w = (pdfl * pdfl) / (pdfl * pdfl + pdf * pdf + 1e-6f);
    direct = L.tp * emis * factor * w;
  }
  finish_segment(L, a, direct);
}

__device__ __forceinline__ void item_to_pixel(const PathArgs& a, uint32_t item, int& pi, int& pj, int& blk) {
  const uint32_t npix = (uint32_t)a.W * (uint32_t)a.H;
  blk = a.first_block + (int)(item / npix);
  uint32_t r = item % npix;
  if (a.tiled) {
    uint32_t tile = r >> 6, w = r & 63u;
    uint32_t tx = tile % (uint32_t)(a.W >> 3), ty = tile / (uint32_t)(a.W >> 3);
    pi = (int)(tx * 8 + (w & 7u));
    pj = (int)(ty * 8 + (w >> 3));
  } else {
    pi = (int)(r / (uint32_t)a.H);
    pj = (int)(r % (uint32_t)a.H);
  }
}

__global__ __launch_bounds__(256) void sr_path_kernel(PathArgs a) {
  __shared__ int lds_stack[SR_STACK_SIZE * 256];
  int* stk = lds_stack + threadIdx.x;
  const int stride = blockDim.x;
  const int lane = (int)__lane_id();
  Lane L;
  L.phase = PH_NEW;
  L.band = a.s.B;
  L.pi = L.pj = L.blk = 0;
  L.acc = 0.0f;
  L.acc_band = 0;
  L.n_seg = L.n_trav = L.n_nf = 0;
  while (true) {
    bool need = false;
    if (L.phase == PH_NEW) {
      L.band += 1;
      need = L.band >= a.s.B;
    }
    const unsigned long long mask = __ballot(need);
    if (mask) {
      const int leader = __ffsll((long long)mask) - 1;
      uint32_t base = 0;
      if (lane == leader) base = atomicAdd(a.queue, (uint32_t)__popcll(mask));
      base = __shfl(base, leader);
      if (need) {
        uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
        uint32_t item = base + rank;
        if (item >= a.n_items) {
          L.phase = PH_DONE;
        } else {
          item_to_pixel(a, item, L.pi, L.pj, L.blk);
          L.band = 0;
        }
      }
    }
    if (L.phase == PH_NEW) start_path(L, a);
    if (__ballot(L.phase != PH_DONE) == 0ull) break;
    if (L.phase == PH_EXT || L.phase == PH_SHD) {
      const bool ext = L.phase == PH_EXT;
      f3 ro = ext ? L.pos : L.pos + L.sdir * L.soff;
      f3 rd = ext ? L.dir : L.sdir;
      HitD h = trace_closest(a.s, ro, rd, stk, stride, a.overflow);
      if (ext) on_extension(L, a, h);
      else on_shadow(L, a, h);
    }
  }
  // per-wave counter reduction, one atomic per wave
  unsigned long long seg = L.n_seg, trav = L.n_trav, nf = L.n_nf;
  for (int off = 32; off >= 1; off >>= 1) {
    seg += __shfl_xor(seg, off);
    trav += __shfl_xor(trav, off);
    nf += __shfl_xor(nf, off);
  }
  if (lane == 0) {
    atomicAdd(a.counters + 0, seg);
    atomicAdd(a.counters + 1, trav);
    atomicAdd(a.counters + 2, nf);
  }
}

// ====================================================================== block mean
// average_and_reset_block (main.py:1592-1607) for n blocks at once:
// mean = (mean * first + sum) / (first + n); sum = 0.
__global__ void sr_finalize_kernel(float* __restrict__ mean, float* __restrict__ sum, size_t n, float first,
                                   float total) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t step = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += step) {
    mean[i] = (mean[i] * first + sum[i]) / total;
    sum[i] = 0.0f;
  }
}

__global__ void sr_fill_kernel(float* __restrict__ p, size_t n, float v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t step = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += step) p[i] = v;
}

// ====================================================================== resolve
// color_rays, main.py:1609-1629: CMF projection, XYZ -> RGB, Reinhard,
// gamma, nearest upsample from the render grid to the display grid.
__global__ void sr_resolve_kernel(const float* __restrict__ acc, int W, int H, int B, ResolveParams p,
                                  float* __restrict__ out, int OW, int OH) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= OW * OH) return;
  const int i = idx / OH, j = idx % OH;
  const int ri = (int)((float)i / (float)OW * (float)W);
  const int rj = (int)((float)j / (float)OH * (float)H);
  const float* a = acc + ((size_t)ri * H + rj) * B;
  float x = 0.f, y = 0.f, z = 0.f;
  for (int k = 0; k < B; ++k) {
    float v = a[k];
    x += v * p.cmf[k * 3 + 0];
    y += v * p.cmf[k * 3 + 1];
    z += v * p.cmf[k * 3 + 2];
  }
  const float sc = p.step / p.norm;
  x *= sc; y *= sc; z *= sc;
  float rgb[3];
  for (int r = 0; r < 3; ++r) {
    float c = p.T[r * 3 + 0] * x + p.T[r * 3 + 1] * y + p.T[r * 3 + 2] * z;
    float e = p.exposure * c;
    float m = e / (e + 1.0f);
    rgb[r] = powf(m, 1.0f / p.gamma);
  }
  float* o = out + (size_t)idx * 3;
  o[0] = rgb[0]; o[1] = rgb[1]; o[2] = rgb[2];
}

// ====================================================================== launchers
int sr_launch_trace(const SceneDev& s, int n, const float* o, const float* d, float* t, int* prim, float* uv,
                    uint32_t* overflow, hipStream_t st) {
  if (n <= 0) return 0;
  const int bs = 256;
  hipLaunchKernelGGL(sr_trace_kernel, dim3((n + bs - 1) / bs), dim3(bs), 0, st, s, n, o, d, t, prim, uv, overflow);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int sr_path_grid(int* grid_out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return -1;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return -1;
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, sr_path_kernel, 256, 0) != hipSuccess) return -1;
  if (per_cu < 1) per_cu = 1;
  *grid_out = per_cu * prop.multiProcessorCount;
  return 0;
}

int sr_launch_path(const PathArgs& a, int grid, hipStream_t st) {
  hipLaunchKernelGGL(sr_path_kernel, dim3(grid), dim3(256), 0, st, a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int sr_launch_finalize(float* mean, float* sum, size_t n, int first, int nb, hipStream_t st) {
  if (n == 0) return 0;
  size_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(sr_finalize_kernel, dim3((unsigned)blocks), dim3(256), 0, st, mean, sum, n, (float)first,
                     (float)(first + nb));
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int sr_launch_fill(float* p, size_t n, float v, hipStream_t st) {
  if (n == 0) return 0;
  size_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(sr_fill_kernel, dim3((unsigned)blocks), dim3(256), 0, st, p, n, v);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int sr_launch_resolve(const float* acc, int W, int H, int B, const ResolveParams& p, float* out, int OW, int OH,
                      hipStream_t st) {
  int n = OW * OH;
  if (n <= 0) return 0;
  hipLaunchKernelGGL(sr_resolve_kernel, dim3((n + 255) / 256), dim3(256), 0, st, acc, W, H, B, p, out, OW, OH);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
