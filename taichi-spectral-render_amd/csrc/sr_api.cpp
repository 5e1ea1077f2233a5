// sr_api.cpp -- C ABI of the spectral path tracer (host runtime).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/spectral_render.h"
#include "sr_kernels.h"

int sr_build_bvh4(const float* v, int n, std::vector<Node4>& nodes, std::vector<int>& tri_order);
int sr_build_light_tree(int n, const float* lverts, const int* slots, const int* slot_mat, const unsigned char* se,
                        int B, std::vector<LightNode>& nodes, std::vector<float>& power, std::vector<LightKey>& keys);

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(expr)                                                                         \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess) return fail(SR_EHIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

// every SR_RENDER_* flag sr_render, sr_render_noalloc and sr_batch_submit accept
#define SR_RENDER_FLAGS (SR_RENDER_LIGHT_TREE | SR_RENDER_SUBSURFACE | SR_RENDER_END_DEAD_PATHS)

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  int ensure(size_t n) {
    if (n <= bytes && p) return 0;
    release();
    if (n == 0) n = 16;
    if (hipMalloc(&p, n) != hipSuccess) { p = nullptr; return fail(SR_ENOMEM, "hipMalloc(%zu) failed", n); }
    bytes = n;
    return 0;
  }
  template <class T>
  int upload(const T* src, size_t count, hipStream_t st) {
    int rc = ensure(count * sizeof(T));
    if (rc) return rc;
    if (count) HIPCHK(hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, st));
    return 0;
  }
};

struct V3h {
  float x, y, z;
};
inline V3h sub(V3h a, V3h b) { return V3h{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3h crs(V3h a, V3h b) { return V3h{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float dt(V3h a, V3h b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3h nrm(V3h a) {
  float l = std::sqrt(dt(a, a));
  return V3h{a.x / l, a.y / l, a.z / l};
}

}  // namespace

struct sr_handle {
  int B = 0;
  float wl_min = 0, wl_max = 0, step = 0;
  int device = 0;
  hipStream_t own = nullptr, st = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t ev_switch = nullptr;  // orders work across a change of stream (sr_set_stream)
  bool timed = false;
  // Bumped whenever something a launched job table captures changes: the
  // scene arrays (upload), the camera, the material tables, the block-sum
  // buffer.  A batch re-fills its job table when a handle's generation moved.
  uint64_t gen = 0;
  // scene
  bool has_scene = false;
  int n_tris = 0, n_mats = 0, n_eem = 0, n_nodes = 0, depth = 0, max_lights = 0, n_slots = 0;
  int n_lnodes = 0, n_lkeys = 0;
  DevBuf nodes, tris, shade, mats, mat_eem, eem, lights, lverts, light_count, light_area, lam_m, lam_nm;
  DevBuf lnodes, lpower, lkeys;  // light tree
  DevBuf rgb_tmp;                // resolved render-grid image for an upsampled display
  std::vector<float> h_light_area;
  std::vector<int> h_light_count;
  std::vector<unsigned char> h_se;  // surface_emission [M][B] the light lists were built from
  float cam_o[3], cam_u[3], cam_v[3], cam_w[3];
  // work
  DevBuf sum, work, jobs, overflow, tmp, alb_off;
  // where the last path launch left its counters and overflow flag (the
  // handle's work area, or the device scratch of sr_render_noalloc)
  void* last_work = nullptr;
  uint32_t* last_overflow = nullptr;
  Job job_host;  // staging copy of the one-job table (kept alive for the async copy)
  int grid = 0, adj_grid = 0, sss_grid = 0;
};

static void fill_scene_dev(const sr_handle* h, SceneDev& s) {
  s.nodes = (const Node4*)h->nodes.p;
  s.tris = (const float4*)h->tris.p;
  s.shade = (const float4*)h->shade.p;
  s.mats = (const MatRec*)h->mats.p;
  s.mat_eem = (const int*)h->mat_eem.p;
  s.eem = (const float*)h->eem.p;
  s.lights = (const LightRec*)h->lights.p;
  s.lverts = (const float4*)h->lverts.p;
  s.light_count = (const int*)h->light_count.p;
  s.light_area = (const float*)h->light_area.p;
  s.lam_m = (const float*)h->lam_m.p;
  s.lam_nm = (const float*)h->lam_nm.p;
  s.lnodes = h->n_lnodes > 0 ? (const LightNode*)h->lnodes.p : nullptr;
  s.lpower = h->n_lnodes > 0 ? (const float*)h->lpower.p : nullptr;
  s.lkeys = h->n_lkeys > 0 ? (const LightKey*)h->lkeys.p : nullptr;
  s.n_lkeys = h->n_lkeys;
  s.n_lnodes = h->n_lnodes;
  s.B = h->B;
  s.max_lights = std::max(h->max_lights, 1);
  s.M = h->n_mats;
  s.n_tris = h->n_tris;
  for (int k = 0; k < 3; ++k) {
    s.cam_o[k] = h->cam_o[k];
    s.cam_u[k] = h->cam_u[k];
    s.cam_v[k] = h->cam_v[k];
    s.cam_w[k] = h->cam_w[k];
  }
}

static void set_camera_host(sr_handle* h, const float* cam) {
  V3h o{cam[0], cam[1], cam[2]}, la{cam[3], cam[4], cam[5]}, up{cam[6], cam[7], cam[8]};
  V3h w = nrm(sub(o, la));  // update_camera, main.py:633-640
  V3h u = nrm(crs(up, w));
  V3h v = crs(w, u);
  h->cam_o[0] = o.x; h->cam_o[1] = o.y; h->cam_o[2] = o.z;
  h->cam_u[0] = u.x; h->cam_u[1] = u.y; h->cam_u[2] = u.z;
  h->cam_v[0] = v.x; h->cam_v[1] = v.y; h->cam_v[2] = v.z;
  h->cam_w[0] = w.x; h->cam_w[1] = w.y; h->cam_w[2] = w.z;
}

static std::vector<unsigned char> surface_emission_flags(const int* mat_i32, int n_mats, int B) {
  std::vector<unsigned char> se((size_t)n_mats * B);
  for (int m = 0; m < n_mats; ++m)
    for (int b = 0; b < B; ++b) se[(size_t)m * B + b] = mat_i32[((size_t)m * 2 + 1) * B + b] ? 1 : 0;
  return se;
}

static int upload_materials(sr_handle* h, const float* mat_f32, const int* mat_i32, const int* mat_eem, int n_mats,
                            const float* eem, int n_eem) {
  const int B = h->B;
  if (n_mats <= 0) return fail(SR_EINVAL, "scene needs at least one material (material 0 is the ambient medium)");
  if (!mat_f32 || !mat_i32 || !mat_eem) return fail(SR_EINVAL, "material arrays must not be NULL");
  if (n_eem < 0 || (n_eem > 0 && !eem)) return fail(SR_EINVAL, "bad EEM table");
  std::vector<MatRec> mr((size_t)n_mats * B);
  for (int m = 0; m < n_mats; ++m) {
    int e = mat_eem[m];
    if (e < 0 || (e > 0 && e >= n_eem)) return fail(SR_EINVAL, "material %d has eem_id %d outside [0, %d)", m, e, n_eem);
    for (int b = 0; b < B; ++b) {
      MatRec& r = mr[(size_t)m * B + b];
      const float* f = mat_f32 + (size_t)m * 6 * B;
      const int* i = mat_i32 + (size_t)m * 2 * B;
      r.n = f[0 * B + b];
      r.k = f[1 * B + b];
      r.r = f[2 * B + b];
      r.ss = f[3 * B + b];
      r.em = f[4 * B + b];
      r.g = f[5 * B + b];
      r.tv = i[0 * B + b];
      r.se = i[1 * B + b];
    }
  }
  int rc;
  if ((rc = h->mats.upload(mr.data(), mr.size(), h->st))) return rc;
  if ((rc = h->mat_eem.upload(mat_eem, (size_t)n_mats, h->st))) return rc;
  std::vector<float> et((size_t)std::max(n_eem, 1) * B * B, 0.0f);
  if (n_eem > 0) std::memcpy(et.data(), eem, sizeof(float) * (size_t)n_eem * B * B);
  if ((rc = h->eem.upload(et.data(), et.size(), h->st))) return rc;
  h->n_mats = n_mats;
  h->n_eem = n_eem;
  HIPCHK(hipStreamSynchronize(h->st));
  return 0;
}

struct sr_batch {
  int device = 0;
  hipStream_t st = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  struct Spec {
    sr_handle* h;
    int W, H, first, nb, depth;
    uint64_t seed;
    uint32_t stream;
    float* accum;
    int flags;     // SR_RENDER_*
    uint64_t gen;  // h->gen when the job table was last filled
  };
  std::vector<Spec> specs;
  std::vector<Job> host_jobs;
  uint32_t tab_bytes = 0;  // LDS table region of the launch (0: the jobs' tables stay in global memory)
  DevBuf jobs, work, overflow;
  int block = 256, grid_pct = 100;
  uint32_t chunk = 64;
  bool dirty = true;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool launched = false;
  double clock_khz = 100000.0;
};

// Work area of a launch over n jobs: queues[n][SR_QUEUES], done[n] (u32),
// then clocks[n + 1] and counters[4 n] (u64, 8-byte aligned).
static constexpr size_t work_u32_bytes(int n) { return ((size_t)4 * (SR_QUEUES + 1) * n + 7) / 8 * 8; }
static constexpr size_t work_bytes(int n) { return work_u32_bytes(n) + 8 * (size_t)(n + 1) + 32 * (size_t)n; }
struct WorkView {
  uint32_t* queues;
  uint32_t* done;
  unsigned long long* clocks;
  unsigned long long* counters;
};
static WorkView work_view(void* base, int n) {
  WorkView w;
  w.queues = (uint32_t*)base;
  w.done = w.queues + (size_t)SR_QUEUES * n;
  w.clocks = (unsigned long long*)((char*)base + work_u32_bytes(n));
  w.counters = w.clocks + (n + 1);
  return w;
}
// A job's work queue is a u32 that every wave advances by `chunk` per claim,
// with at most one claim per wave past the end: the counter must not wrap
// (each of a job's queues holds at most its n_items).
static bool queue_overflows(uint32_t n_items, int grid, int block, uint32_t chunk) {
  const unsigned long long waves = ((unsigned long long)grid * (unsigned long long)block + 63) / 64;
  return (unsigned long long)n_items + (waves + 1) * (unsigned long long)chunk >= (1ull << 32);
}

// Graph-safe reset of the work area (kernels only, no memset nodes).
static int reset_work(void* base, int n, hipStream_t st) {
  WorkView w = work_view(base, n);
  if (sr_launch_fill_u32(w.queues, (size_t)(SR_QUEUES + 1) * n, 0u, st)) return -1;
  if (sr_launch_fill_u32((uint32_t*)w.clocks, (size_t)2 * (n + 1), 0xFFFFFFFFu, st)) return -1;
  if (sr_launch_fill_u32((uint32_t*)w.counters, (size_t)8 * n, 0u, st)) return -1;
  return 0;
}

extern "C" {

const char* sr_last_error(void) { return g_err.c_str(); }

int sr_version(void) { return 100; }

int sr_create(int bands, float wl_min, float wl_max, sr_handle** out) {
  g_err.clear();
  if (!out) return fail(SR_EINVAL, "out is NULL");
  *out = nullptr;
  if (bands < 2 || bands > SR_MAX_BANDS) return fail(SR_EINVAL, "bands=%d outside [2, %d]", bands, SR_MAX_BANDS);
  if (!(wl_max > wl_min) || !(wl_min > 0.f)) return fail(SR_EINVAL, "need 0 < wl_min < wl_max");
  sr_handle* h = new sr_handle();
  h->B = bands;
  h->wl_min = wl_min;
  h->wl_max = wl_max;
  h->step = (float)(((double)wl_max - (double)wl_min) / (double)(bands - 1));
  hipError_t e = hipGetDevice(&h->device);
  if (e == hipSuccess) e = hipStreamCreate(&h->own);  // blocking: ordered with the null stream
  if (e == hipSuccess) e = hipEventCreate(&h->ev0);
  if (e == hipSuccess) e = hipEventCreate(&h->ev1);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&h->ev_switch, hipEventDisableTiming);
  if (e != hipSuccess) {
    delete h;
    return fail(SR_EHIP, "HIP init failed: %s", hipGetErrorString(e));
  }
  h->st = h->own;
  if (sr_path_grid(&h->grid) != 0 || sr_adjoint_grid(&h->adj_grid) != 0) {
    delete h;
    return fail(SR_EHIP, "occupancy query failed");
  }
  int rc = h->work.ensure(work_bytes(1));
  if (!rc) rc = h->jobs.ensure(sizeof(Job));
  if (!rc) rc = h->overflow.ensure(64);
  if (rc) {
    delete h;
    return rc;
  }
  (void)hipMemset(h->work.p, 0, h->work.bytes);
  (void)hipMemset(h->overflow.p, 0, h->overflow.bytes);
  *out = h;
  return 0;
}

int sr_destroy(sr_handle* h) {
  if (!h) return 0;
  (void)hipSetDevice(h->device);
  if (h->st) (void)hipStreamSynchronize(h->st);
  DevBuf* bufs[] = {&h->nodes, &h->tris, &h->shade, &h->mats, &h->mat_eem, &h->eem, &h->lights, &h->lverts,
                    &h->light_count, &h->light_area, &h->lam_m, &h->lam_nm, &h->lnodes, &h->lpower, &h->lkeys,
                    &h->sum, &h->work, &h->jobs, &h->overflow, &h->tmp, &h->alb_off, &h->rgb_tmp};
  for (DevBuf* b : bufs) b->release();
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  if (h->ev_switch) (void)hipEventDestroy(h->ev_switch);
  if (h->own) (void)hipStreamDestroy(h->own);
  delete h;
  return 0;
}

int sr_set_stream(sr_handle* h, void* stream) {
  if (!h) return fail(SR_EINVAL, "null handle");
  hipStream_t ns = stream ? (hipStream_t)stream : h->own;
  if (ns == h->st) return 0;
  // Everything already enqueued on the old stream (it shares the handle's
  // block sums, work counters and staging buffers) completes before
  // anything enqueued on the new one: an event, no host wait.
  HIPCHK(hipSetDevice(h->device));
  HIPCHK(hipEventRecord(h->ev_switch, h->st));
  HIPCHK(hipStreamWaitEvent(ns, h->ev_switch, 0));
  h->st = ns;
  return 0;
}

// Light tree over the light slots (light_tree.cpp), the same for both scene
// builds: lv [n_slots][3][4] f32, slot faces and their materials.
static int upload_light_tree(sr_handle* h, int n_slots, const float* lv, const int* slots, const int* slot_mat,
                             const int* mat_i32, int n_mats, hipStream_t st) {
  const int B = h->B;
  std::vector<unsigned char> se((size_t)n_mats * B);
  for (int m = 0; m < n_mats; ++m)
    for (int b = 0; b < B; ++b) se[(size_t)m * B + b] = mat_i32[((size_t)m * 2 + 1) * B + b] ? 1 : 0;
  std::vector<LightNode> ln;
  std::vector<float> lp;
  std::vector<LightKey> lk;
  sr_build_light_tree(n_slots, lv, slots, slot_mat, se.data(), B, ln, lp, lk);
  int rc;
  if ((rc = h->lnodes.upload(ln.data(), ln.size(), st))) return rc;
  if ((rc = h->lpower.upload(lp.data(), lp.size(), st))) return rc;
  if ((rc = h->lkeys.upload(lk.data(), lk.size(), st))) return rc;
  h->n_lnodes = (int)ln.size();
  h->n_lkeys = (int)lk.size();
  return 0;
}

int sr_upload_scene(sr_handle* h, const float* tris_v, const float* tris_n, const int* tri_mat, int n_tris,
                    const float* mat_f32, const int* mat_i32, const int* mat_eem, int n_mats, const float* eem,
                    int n_eem, const float* cam) {
  g_err.clear();
  if (!h) return fail(SR_EINVAL, "null handle");
  if (n_tris < 0 || (n_tris > 0 && (!tris_v || !tris_n || !tri_mat))) return fail(SR_EINVAL, "bad triangle arrays");
  if (n_tris >= (1 << 27)) return fail(SR_EINVAL, "n_tris=%d exceeds 2^27", n_tris);
  if (!cam) return fail(SR_EINVAL, "cam is NULL");
  HIPCHK(hipSetDevice(h->device));
  const int B = h->B;
  for (int f = 0; f < n_tris; ++f)
    if (tri_mat[f] < 0 || tri_mat[f] >= n_mats)
      return fail(SR_EINVAL, "face %d has material %d outside [0, %d)", f, tri_mat[f], n_mats);
  int rc = upload_materials(h, mat_f32, mat_i32, mat_eem, n_mats, eem, n_eem);
  if (rc) return rc;

  // per-face normal and area in f32 (main.py:526, 535-536)
  std::vector<float> area((size_t)n_tris);
  std::vector<float4> shade((size_t)std::max(n_tris, 1) * 4);
  for (int f = 0; f < n_tris; ++f) {
    const float* p = tris_v + (size_t)f * 9;
    const float* q = tris_n + (size_t)f * 9;
    V3h v0{p[0], p[1], p[2]}, v1{p[3], p[4], p[5]}, v2{p[6], p[7], p[8]};
    V3h n = crs(sub(v1, v0), sub(v2, v0));
    float a = std::sqrt(dt(n, n)) / 2.0f;
    if (a > 1e-6f) n = V3h{n.x / (a * 2.0f), n.y / (a * 2.0f), n.z / (a * 2.0f)};
    area[f] = a;
    int mat = tri_mat[f];
    float matf;
    std::memcpy(&matf, &mat, 4);
    shade[(size_t)f * 4 + 0] = make_float4(n.x, n.y, n.z, a);
    shade[(size_t)f * 4 + 1] = make_float4(q[0], q[1], q[2], matf);
    shade[(size_t)f * 4 + 2] = make_float4(q[3], q[4], q[5], 0.f);
    shade[(size_t)f * 4 + 3] = make_float4(q[6], q[7], q[8], 0.f);
  }

  // per-band light lists (main.py:590-627): faces whose material emits at
  // band b, in face order; CDF of area fractions accumulated in f64 and
  // stored f32, last entry forced to 1.
  std::vector<std::vector<int>> per_band((size_t)B);
  std::vector<int> slot_of((size_t)n_tris, -1);
  std::vector<int> slots;
  int max_lights = 0;
  for (int b = 0; b < B; ++b) {
    for (int f = 0; f < n_tris; ++f) {
      if (mat_i32[((size_t)tri_mat[f] * 2 + 1) * B + b]) per_band[b].push_back(f);
    }
    max_lights = std::max(max_lights, (int)per_band[b].size());
  }
  // (the path kernels copy these records to LDS as one 16-byte unit each)
  static_assert(sizeof(LightRec) == 16, "LightRec is 16 bytes");
  std::vector<LightRec> lr((size_t)B * std::max(max_lights, 1));
  std::vector<int> lcount((size_t)B, 0);
  std::vector<float> larea((size_t)B, 0.0f);
  for (int b = 0; b < B; ++b) {
    const std::vector<int>& ids = per_band[b];
    lcount[b] = (int)ids.size();
    if (ids.empty()) continue;
    double tot = 0.0;
    for (int f : ids) tot += (double)area[f];
    larea[b] = (float)tot;
    double c = 0.0;
    for (size_t i = 0; i < ids.size(); ++i) {
      int f = ids[i];
      c += (double)area[f] / tot;
      if (slot_of[f] < 0) {
        slot_of[f] = (int)slots.size();
        slots.push_back(f);
      }
      LightRec& L = lr[(size_t)b * std::max(max_lights, 1) + i];
      L.cdf = (float)c;
      L.prim = f;
      L.slot = slot_of[f];
      L.pad = 0;
    }
    lr[(size_t)b * std::max(max_lights, 1) + ids.size() - 1].cdf = 1.0f;
  }
  std::vector<float4> lv(std::max<size_t>(slots.size(), 1) * 3);
  for (size_t s = 0; s < slots.size(); ++s) {
    const float* p = tris_v + (size_t)slots[s] * 9;
    for (int k = 0; k < 3; ++k) lv[s * 3 + k] = make_float4(p[3 * k], p[3 * k + 1], p[3 * k + 2], 0.f);
  }

  // 4-wide BVH and triangle records in leaf order
  std::vector<Node4> nodes;
  std::vector<int> order;
  int depth = sr_build_bvh4(tris_v, n_tris, nodes, order);
  if (depth < 0) return fail(SR_EINVAL, "BVH build failed");
  if (3 * depth + 1 > SR_STACK_SIZE) return fail(SR_EOVERFLOW, "BVH depth %d exceeds the traversal stack", depth);
  // SR_LEAF_MAX - 1 zero records past the end: a leaf's four records are read at once
  std::vector<float4> tr((order.size() + SR_LEAF_MAX - 1) * 3, make_float4(0.f, 0.f, 0.f, 0.f));
  for (size_t i = 0; i < order.size(); ++i) {
    int f = order[i];
    const float* p = tris_v + (size_t)f * 9;
    float pf;
    std::memcpy(&pf, &f, 4);
    V3h v0{p[0], p[1], p[2]}, v1{p[3], p[4], p[5]}, v2{p[6], p[7], p[8]};
    V3h e1 = sub(v1, v0), e2 = sub(v2, v0);
    tr[i * 3 + 0] = make_float4(v0.x, v0.y, v0.z, pf);
    tr[i * 3 + 1] = make_float4(e1.x, e1.y, e1.z, 0.f);
    tr[i * 3 + 2] = make_float4(e2.x, e2.y, e2.z, 0.f);
  }
  if (nodes.empty()) {  // empty scene: one node with four empty slots
    Node4 t;
    std::memset(&t, 0, sizeof(t));
    for (int s = 0; s < 4; ++s) {
      t.lo_x[s] = t.lo_y[s] = t.lo_z[s] = 3.0e38f;
      t.hi_x[s] = t.hi_y[s] = t.hi_z[s] = -3.0e38f;
      t.child[s] = SR_EMPTY;
    }
    nodes.push_back(t);
  }

  // band wavelengths as the reference forms them (main.py:1057, 1394)
  std::vector<float> lm((size_t)B), ln((size_t)B);
  for (int b = 0; b < B; ++b) {
    lm[b] = (float)((double)h->wl_min * 1e-9) + (float)b * h->step * 1e-9f;
    ln[b] = h->wl_min + (float)b * h->step;
  }

  hipStream_t st = h->st;
  if ((rc = h->nodes.upload(nodes.data(), nodes.size(), st))) return rc;
  if ((rc = h->tris.upload(tr.data(), tr.size(), st))) return rc;
  if ((rc = h->shade.upload(shade.data(), shade.size(), st))) return rc;
  if ((rc = h->lights.upload(lr.data(), lr.size(), st))) return rc;
  if ((rc = h->lverts.upload(lv.data(), lv.size(), st))) return rc;
  {
    std::vector<int> smat(slots.size());
    for (size_t s = 0; s < slots.size(); ++s) smat[s] = tri_mat[slots[s]];
    if ((rc = upload_light_tree(h, (int)slots.size(), (const float*)lv.data(), slots.data(), smat.data(), mat_i32,
                                n_mats, st)))
      return rc;
  }
  if ((rc = h->light_count.upload(lcount.data(), lcount.size(), st))) return rc;
  if ((rc = h->light_area.upload(larea.data(), larea.size(), st))) return rc;
  if ((rc = h->lam_m.upload(lm.data(), lm.size(), st))) return rc;
  if ((rc = h->lam_nm.upload(ln.data(), ln.size(), st))) return rc;
  HIPCHK(hipStreamSynchronize(st));
  h->n_tris = n_tris;
  h->n_nodes = (int)nodes.size();
  h->depth = depth;
  h->max_lights = max_lights;
  h->n_slots = (int)slots.size();
  h->h_light_area = larea;
  h->h_light_count = lcount;
  h->h_se = surface_emission_flags(mat_i32, n_mats, B);
  set_camera_host(h, cam);
  h->gen++;
  h->has_scene = true;
  return 0;
}

int sr_set_camera(sr_handle* h, const float* cam) {
  if (!h || !cam) return fail(SR_EINVAL, "null argument");
  set_camera_host(h, cam);
  h->gen++;
  return 0;
}

int sr_update_materials(sr_handle* h, const float* mat_f32, const int* mat_i32, const int* mat_eem, int n_mats,
                        const float* eem, int n_eem) {
  g_err.clear();
  if (!h) return fail(SR_EINVAL, "null handle");
  if (!h->has_scene) return fail(SR_ESTATE, "no scene uploaded");
  if (n_mats != h->n_mats) return fail(SR_EINVAL, "material count changed (%d -> %d)", h->n_mats, n_mats);
  if (!mat_i32) return fail(SR_EINVAL, "material arrays must not be NULL");
  // the light lists, their CDFs, the emitting areas and the light tree were
  // built from surface_emission: a change there needs a new sr_upload_scene
  const std::vector<unsigned char> se = surface_emission_flags(mat_i32, n_mats, h->B);
  for (size_t i = 0; i < se.size(); ++i)
    if (se[i] != h->h_se[i])
      return fail(SR_EINVAL, "surface_emission of material %d, band %d changed: re-upload the scene",
                  (int)(i / (size_t)h->B), (int)(i % (size_t)h->B));
  HIPCHK(hipSetDevice(h->device));
  int rc = upload_materials(h, mat_f32, mat_i32, mat_eem, n_mats, eem, n_eem);
  if (!rc) h->gen++;
  return rc;
}

// Device scene compile: the same records as sr_upload_scene (main.py:509-631),
// built on the GPU from device triangle arrays (sr_build.hip).
int sr_upload_scene_gpu(sr_handle* h, const float* dev_tris_v, const float* dev_tris_n, const int* dev_tri_mat,
                        int n_tris, const float* mat_f32, const int* mat_i32, const int* mat_eem, int n_mats,
                        const float* eem, int n_eem, const float* cam) {
  g_err.clear();
  if (!h) return fail(SR_EINVAL, "null handle");
  if (n_tris < 0 || (n_tris > 0 && (!dev_tris_v || !dev_tris_n || !dev_tri_mat)))
    return fail(SR_EINVAL, "bad triangle arrays");
  if (n_tris >= (1 << 27)) return fail(SR_EINVAL, "n_tris=%d exceeds 2^27", n_tris);
  if (!cam) return fail(SR_EINVAL, "cam is NULL");
  if (n_mats <= 0 || !mat_i32) return fail(SR_EINVAL, "no materials");
  HIPCHK(hipSetDevice(h->device));
  int rc = upload_materials(h, mat_f32, mat_i32, mat_eem, n_mats, eem, n_eem);
  if (rc) return rc;
  DeviceSceneOut o;
  char err[256];
  rc = sr_build_scene_device(dev_tris_v, dev_tris_n, dev_tri_mat, n_tris, mat_i32, n_mats, h->B, h->st, &o, err,
                             sizeof(err));
  if (rc) {
    sr_free_scene_out(&o);
    return fail(rc, "%s", err);
  }
  if (3 * o.depth + 1 > SR_STACK_SIZE) {
    sr_free_scene_out(&o);
    return fail(SR_EOVERFLOW, "BVH depth %d exceeds the traversal stack", o.depth);
  }
  const int B = h->B;
  std::vector<float> lm((size_t)B), ln((size_t)B);
  for (int b = 0; b < B; ++b) {
    lm[b] = (float)((double)h->wl_min * 1e-9) + (float)b * h->step * 1e-9f;
    ln[b] = h->wl_min + (float)b * h->step;
  }
  if ((rc = h->lam_m.upload(lm.data(), lm.size(), h->st)) || (rc = h->lam_nm.upload(ln.data(), ln.size(), h->st))) {
    sr_free_scene_out(&o);
    return rc;
  }
  // the handle takes over the device records
  auto adopt = [](DevBuf& d, void*& p, size_t bytes) {
    d.release();
    d.p = p;
    d.bytes = bytes;
    p = nullptr;
  };
  adopt(h->nodes, o.nodes, o.nodes_bytes);
  adopt(h->tris, o.tris, o.tris_bytes);
  adopt(h->shade, o.shade, o.shade_bytes);
  adopt(h->lights, o.lights, o.lights_bytes);
  adopt(h->lverts, o.lverts, o.lverts_bytes);
  adopt(h->light_count, o.light_count, o.light_count_bytes);
  adopt(h->light_area, o.light_area, o.light_area_bytes);
  if ((rc = upload_light_tree(h, o.n_slots, o.h_lverts.data(), o.h_slots.data(), o.h_slot_mat.data(), mat_i32, n_mats,
                              h->st)))
    return rc;
  HIPCHK(hipStreamSynchronize(h->st));
  h->n_tris = n_tris;
  h->n_nodes = o.n_nodes;
  h->depth = o.depth;
  h->max_lights = o.max_lights;
  h->n_slots = o.n_slots;
  h->h_light_area = o.h_light_area;
  h->h_light_count = o.h_light_count;
  h->h_se = surface_emission_flags(mat_i32, n_mats, h->B);
  set_camera_host(h, cam);
  h->gen++;
  h->has_scene = true;
  return 0;
}

int sr_scene_info(sr_handle* h, int* n_nodes, int* bvh_depth, int* max_lights, long long* device_bytes) {
  if (!h) return fail(SR_EINVAL, "null handle");
  if (n_nodes) *n_nodes = h->n_nodes;
  if (bvh_depth) *bvh_depth = h->depth;
  if (max_lights) *max_lights = h->max_lights;
  if (device_bytes) {
    const DevBuf* bufs[] = {&h->nodes, &h->tris, &h->shade, &h->mats, &h->mat_eem, &h->eem, &h->lights,
                            &h->lverts, &h->light_count, &h->light_area, &h->lam_m, &h->lam_nm};
    long long t = 0;
    for (const DevBuf* b : bufs) t += (long long)b->bytes;
    *device_bytes = t;
  }
  return 0;
}

// Read-back of the compiled records (the reference's scene fields,
// main.py:1700-1709); sizes are the logical ones (a buffer may be larger).
int sr_scene_array(sr_handle* h, int which, void* dst, size_t capacity, size_t* bytes) {
  g_err.clear();
  if (!h) return fail(SR_EINVAL, "null handle");
  if (!h->has_scene) return fail(SR_ESTATE, "no scene uploaded");
  const size_t B = (size_t)h->B;
  const DevBuf* b = nullptr;
  size_t n = 0;
  switch (which) {
    case SR_ARRAY_NODES: b = &h->nodes; n = (size_t)h->n_nodes * sizeof(Node4); break;
    case SR_ARRAY_TRIS: b = &h->tris; n = ((size_t)h->n_tris + SR_LEAF_MAX - 1) * 3 * sizeof(float4); break;
    case SR_ARRAY_SHADE: b = &h->shade; n = (size_t)std::max(h->n_tris, 1) * 4 * sizeof(float4); break;
    case SR_ARRAY_LIGHTS: b = &h->lights; n = B * (size_t)std::max(h->max_lights, 1) * sizeof(LightRec); break;
    case SR_ARRAY_LVERTS: b = &h->lverts; n = (size_t)std::max(h->n_slots, 1) * 3 * sizeof(float4); break;
    case SR_ARRAY_LIGHT_COUNT: b = &h->light_count; n = B * sizeof(int); break;
    case SR_ARRAY_LIGHT_AREA: b = &h->light_area; n = B * sizeof(float); break;
    case SR_ARRAY_LIGHT_TREE: b = &h->lnodes; n = (size_t)h->n_lnodes * sizeof(LightNode); break;
    case SR_ARRAY_LIGHT_POWER: b = &h->lpower; n = (size_t)h->n_lnodes * B * sizeof(float); break;
    case SR_ARRAY_LIGHT_KEYS: b = &h->lkeys; n = (size_t)h->n_lkeys * sizeof(LightKey); break;
    default: return fail(SR_EINVAL, "unknown scene array %d", which);
  }
  if (n > b->bytes) return fail(SR_ESTATE, "scene array %d holds %zu bytes, expected %zu", which, b->bytes, n);
  if (bytes) *bytes = n;
  if (!dst || n == 0) return 0;
  if (capacity < n) return fail(SR_EINVAL, "capacity %zu < %zu bytes", capacity, n);
  HIPCHK(hipSetDevice(h->device));
  HIPCHK(hipMemcpyAsync(dst, b->p, n, hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  return 0;
}

int sr_trace_closest(sr_handle* h, int n, const float* orig, const float* dir, float* out_t, int* out_prim,
                     float* out_uv, int flags) {
  g_err.clear();
  if (!h) return fail(SR_EINVAL, "null handle");
  if (!h->has_scene) return fail(SR_ESTATE, "no scene uploaded");
  if (n < 0) return fail(SR_EINVAL, "n < 0");
  if (n == 0) return 0;
  if (!orig || !dir || !out_t || !out_prim || !out_uv) return fail(SR_EINVAL, "null array");
  HIPCHK(hipSetDevice(h->device));
  SceneDev s;
  fill_scene_dev(h, s);
  hipStream_t st = h->st;
  HIPCHK(hipMemsetAsync(h->overflow.p, 0, 4, st));
  if (flags & SR_DEVICE_PTRS) {
    if (sr_launch_trace(s, n, orig, dir, out_t, out_prim, out_uv, (uint32_t*)h->overflow.p, st))
      return fail(SR_EHIP, "trace launch failed");
    return 0;
  }
  const size_t nb = (size_t)n;
  size_t need = nb * (3 + 3 + 1 + 1 + 2) * 4;
  int rc = h->tmp.ensure(need);
  if (rc) return rc;
  char* base = (char*)h->tmp.p;
  float* d_o = (float*)base;
  float* d_d = d_o + nb * 3;
  float* d_t = d_d + nb * 3;
  int* d_p = (int*)(d_t + nb);
  float* d_uv = (float*)(d_p + nb);
  HIPCHK(hipMemcpyAsync(d_o, orig, nb * 12, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(d_d, dir, nb * 12, hipMemcpyHostToDevice, st));
  if (sr_launch_trace(s, n, d_o, d_d, d_t, d_p, d_uv, (uint32_t*)h->overflow.p, st))
    return fail(SR_EHIP, "trace launch failed");
  HIPCHK(hipMemcpyAsync(out_t, d_t, nb * 4, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(out_prim, d_p, nb * 4, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(out_uv, d_uv, nb * 8, hipMemcpyDeviceToHost, st));
  uint32_t ovf = 0;
  HIPCHK(hipMemcpyAsync(&ovf, h->overflow.p, 4, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (ovf) return fail(SR_EOVERFLOW, "traversal stack overflow");
  return 0;
}

static int check_render_args(sr_handle* h, int width, int height, int first_block, int n_blocks, int max_depth,
                             uint32_t stream) {
  if (!h) return fail(SR_EINVAL, "null handle");
  if (!h->has_scene) return fail(SR_ESTATE, "no scene uploaded");
  if (width <= 0 || height <= 0) return fail(SR_EINVAL, "bad resolution %dx%d", width, height);
  if (first_block < 0 || n_blocks < 0) return fail(SR_EINVAL, "bad block range");
  if (max_depth < 0) return fail(SR_EINVAL, "max_depth < 0");
  if (stream > 0xFFFFFFu) return fail(SR_EINVAL, "stream id must fit 24 bits");
  // work items: pixels x blocks x band groups
  const unsigned long long items =
      (unsigned long long)width * height * (unsigned long long)n_blocks * sr_items_per_pixel(h->B);
  if (items >= 0x7FFFFFFFull) return fail(SR_EINVAL, "W*H*n_blocks*ceil(bands/%d) must be < 2^31 per call",
                                          SR_BANDS_PER_ITEM);
  return 0;
}

static void fill_job(sr_handle* h, Job& j, int width, int height, uint64_t seed, uint32_t stream, int first_block,
                     int n_blocks, int max_depth) {
  std::memset(&j, 0, sizeof(j));
  fill_scene_dev(h, j.s);
  j.sum = (float*)h->sum.p;
  j.n_items = (uint32_t)((unsigned long long)width * height * (unsigned long long)n_blocks * sr_items_per_pixel(h->B));
  j.W = width;
  j.H = height;
  j.depth = max_depth;
  j.first_block = first_block;
  j.n_blocks = n_blocks;
  j.tiled = (width % 8 == 0 && height % 8 == 0) ? 1 : 0;
  // the XCD queues split each spectral block's items in eighths, on tile-group bounds when tiled
  j.per_block = (uint32_t)((unsigned long long)width * height * sr_items_per_pixel(h->B));
  const uint64_t unit = j.tiled ? 64 : 1, units = j.per_block / unit;
  for (int x = 0; x < SR_QUEUES; ++x) {
    const uint64_t lo = (uint64_t)x * units / SR_QUEUES * unit, hi = (uint64_t)(x + 1) * units / SR_QUEUES * unit;
    j.q_off[x] = (uint32_t)lo;
    j.q_len[x] = (uint32_t)(hi - lo);
  }
  j.seed_lo = (uint32_t)seed;
  j.seed_hi = (uint32_t)(seed >> 32);
  j.stream = stream;
}

static int ensure_sum(sr_handle* h, int width, int height, hipStream_t st) {
  const size_t n = (size_t)width * height * h->B;
  if (h->sum.bytes < n * 4) {
    int rc = h->sum.ensure(n * 4);
    if (rc) return rc;
    h->gen++;
    if (sr_launch_fill((float*)h->sum.p, n, 0.0f, st)) return fail(SR_EHIP, "fill launch failed");
  }
  return 0;
}

// Device buffers of one path launch: block sums, the one-job table, the
// work area and the overflow flag.
struct LaunchBufs {
  float* sum;
  void* jobs;
  void* work;
  uint32_t* overflow;
};

static int prepare_path_args(sr_handle* h, PathArgs& a, const LaunchBufs& lb, int width, int height, uint64_t seed,
                             uint32_t stream, int first_block, int n_blocks, int max_depth, int flags = 0,
                             bool adjoint = false, float* rec = nullptr) {
  hipStream_t st = h->st;
  fill_job(h, h->job_host, width, height, seed, stream, first_block, n_blocks, max_depth);
  h->job_host.sum = lb.sum;
  h->job_host.rec = rec;
  h->job_host.nee_tree = ((flags & SR_RENDER_LIGHT_TREE) && h->n_lnodes > 0) ? 1 : 0;
  h->job_host.sss = (flags & SR_RENDER_SUBSURFACE) ? 1 : 0;
  h->job_host.end_dead = (flags & SR_RENDER_END_DEAD_PATHS) ? 1 : 0;
  // the job's small tables go to LDS when they fit (single launches run 256-lane blocks)
  h->job_host.tab = sr_lds_tables(h->n_mats, h->B, h->B * std::max(h->max_lights, 1), h->n_slots, 256, adjoint);
  HIPCHK(hipMemcpyAsync(lb.jobs, &h->job_host, sizeof(Job), hipMemcpyHostToDevice, st));
  if (reset_work(lb.work, 1, st)) return fail(SR_EHIP, "reset launch failed");
  if (sr_launch_fill_u32(lb.overflow, 1, 0u, st)) return fail(SR_EHIP, "reset launch failed");
  std::memset(&a, 0, sizeof(a));
  WorkView w = work_view(lb.work, 1);
  a.jobs = (const Job*)lb.jobs;
  a.n_jobs = 1;
  a.queues = w.queues;
  a.done_items = w.done;
  a.clocks = w.clocks;
  a.counters = w.counters;
  a.overflow = lb.overflow;
  a.chunk = 64;
  a.tab_bytes = h->job_host.tab.bytes;
  a.record = rec != nullptr ? 1u : 0u;
  a.adj.rough_mat = a.adj.g_mat = a.adj.sig_mat = -2;
  a.adj.rough_idx = a.adj.g_idx = a.adj.sig_off = -1;
  h->last_work = lb.work;
  h->last_overflow = lb.overflow;
  return 0;
}

static LaunchBufs handle_bufs(sr_handle* h) {
  return LaunchBufs{(float*)h->sum.p, h->jobs.p, h->work.p, (uint32_t*)h->overflow.p};
}

// The path launch and the fold of its blocks into dev_accum (sr_render and
// sr_render_noalloc; the arguments are checked by the caller).
static int render_with(sr_handle* h, const LaunchBufs& lb, int width, int height, uint64_t seed, uint32_t stream,
                       int first_block, int n_blocks, int max_depth, int flags, float* dev_accum,
                       float* rec = nullptr) {
  PathArgs a;
  int rc = prepare_path_args(h, a, lb, width, height, seed, stream, first_block, n_blocks, max_depth, flags, false,
                             rec);
  if (rc) return rc;
  hipStream_t st = h->st;
  const bool sss = (flags & SR_RENDER_SUBSURFACE) != 0;
  if (sss && h->sss_grid == 0 && sr_path_sss_grid(&h->sss_grid)) return fail(SR_EHIP, "occupancy query failed");
  int grid = sss ? h->sss_grid : h->grid;
  const unsigned long long waves_needed = ((unsigned long long)h->job_host.n_items + 255) / 256;
  if ((unsigned long long)grid > waves_needed) grid = (int)waves_needed;
  if (queue_overflows(h->job_host.n_items, grid, 256, a.chunk)) return fail(SR_EINVAL, "work queue would wrap");
  HIPCHK(hipEventRecord(h->ev0, st));
  if ((sss ? sr_launch_path_sss(a, grid, 256, st) : sr_launch_path(a, grid, 256, st)))
    return fail(SR_EHIP, "path kernel launch failed");
  HIPCHK(hipEventRecord(h->ev1, st));
  h->timed = true;
  const size_t n = (size_t)width * height * h->B;
  if (sr_launch_finalize(dev_accum, lb.sum, n, first_block, n_blocks, st))
    return fail(SR_EHIP, "finalize launch failed");
  return 0;
}

int sr_render(sr_handle* h, int width, int height, uint64_t seed, uint32_t stream, int first_block, int n_blocks,
              int max_depth, int flags, float* dev_accum) {
  g_err.clear();
  if (flags & ~SR_RENDER_FLAGS) return fail(SR_EINVAL, "unknown render flags 0x%x", flags);
  int rc = check_render_args(h, width, height, first_block, n_blocks, max_depth, stream);
  if (rc) return rc;
  if (!dev_accum) return fail(SR_EINVAL, "dev_accum is NULL");
  HIPCHK(hipSetDevice(h->device));
  if (n_blocks == 0) return 0;
  if ((rc = ensure_sum(h, width, height, h->st))) return rc;
  return render_with(h, handle_bufs(h), width, height, seed, stream, first_block, n_blocks, max_depth, flags,
                     dev_accum);
}

int sr_render_record(sr_handle* h, int width, int height, uint64_t seed, uint32_t stream, int first_block,
                     int n_blocks, int max_depth, int flags, float* dev_accum, float* dev_path_totals) {
  g_err.clear();
  if (!h) return fail(SR_EINVAL, "null handle");
  if (flags & ~SR_RENDER_END_DEAD_PATHS)
    return fail(SR_EINVAL, "sr_render_record takes only SR_RENDER_END_DEAD_PATHS (flags 0x%x): the adjoint "
                           "replays the area-CDF light pick without the subsurface walk", flags);
  int rc = check_render_args(h, width, height, first_block, n_blocks, max_depth, stream);
  if (rc) return rc;
  if (!dev_accum || !dev_path_totals) return fail(SR_EINVAL, "dev_accum or dev_path_totals is NULL");
  HIPCHK(hipSetDevice(h->device));
  if (n_blocks == 0) return 0;
  if ((rc = ensure_sum(h, width, height, h->st))) return rc;
  return render_with(h, handle_bufs(h), width, height, seed, stream, first_block, n_blocks, max_depth, flags,
                     dev_accum, dev_path_totals);
}

// ---------------------------------------------------------------- allocation-free render
// sr_render_noalloc keeps every buffer of its launch -- the block sums, the
// one-job table, the work area and the overflow flag -- in this file-scope
// device array (one copy per GPU, resident while the library is loaded).
// Host code never uses the array's name as an address: each device's
// address comes from hipGetSymbolAddress on that device's first call.
namespace {
constexpr size_t kScratchSumBytes = (size_t)SR_NOALLOC_MAX_ELEMS * sizeof(float);
constexpr size_t kScratchJobOff = kScratchSumBytes;
constexpr size_t kScratchWorkOff = kScratchJobOff + (sizeof(Job) + 255) / 256 * 256;
constexpr size_t kScratchOvfOff = kScratchWorkOff + 256;
static_assert(work_bytes(1) <= 256, "a one-job work area fits its scratch slot");
constexpr size_t kScratchBytes = kScratchOvfOff + 256;
static_assert(sizeof(Job) % 8 == 0, "the job table is copied whole");
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// The calling thread's current device must be `device`.
static int scratch_for_device(int device, unsigned char** out) {
  if (device < 0 || device >= kMaxDevices) return fail(SR_EINVAL, "device %d outside [0, %d)", device, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(SR_EHIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
    g_scratch_ptr[device] = p;
  }
  *out = (unsigned char*)g_scratch_ptr[device];
  return 0;
}

int sr_render_noalloc(sr_handle* h, int width, int height, uint64_t seed, uint32_t stream, int first_block,
                      int n_blocks, int max_depth, int flags, float* dev_accum) {
  g_err.clear();
  if (flags & ~SR_RENDER_FLAGS) return fail(SR_EINVAL, "unknown render flags 0x%x", flags);
  int rc = check_render_args(h, width, height, first_block, n_blocks, max_depth, stream);
  if (rc) return rc;
  if (!dev_accum) return fail(SR_EINVAL, "dev_accum is NULL");
  const unsigned long long n = (unsigned long long)width * height * (unsigned long long)h->B;
  if (n > SR_NOALLOC_MAX_ELEMS)
    return fail(SR_EINVAL, "width * height * bands = %llu exceeds the %u elements sr_render_noalloc holds", n,
                (unsigned)SR_NOALLOC_MAX_ELEMS);
  HIPCHK(hipSetDevice(h->device));
  if (n_blocks == 0) return 0;
  unsigned char* base = nullptr;
  if ((rc = scratch_for_device(h->device, &base))) return rc;
  LaunchBufs lb{(float*)base, base + kScratchJobOff, base + kScratchWorkOff, (uint32_t*)(base + kScratchOvfOff)};
  // the block sums start at zero (a fill kernel: no memset node if captured)
  if (sr_launch_fill(lb.sum, (size_t)n, 0.0f, h->st)) return fail(SR_EHIP, "fill launch failed");
  return render_with(h, lb, width, height, seed, stream, first_block, n_blocks, max_depth, flags, dev_accum);
}

static int adjoint_with(sr_handle* h, int width, int height, uint64_t seed, uint32_t stream, int first_block,
                        int n_blocks, int max_depth, const float* dev_d_accum, const int* alb_off, int rough_mat,
                        int rough_idx, int g_mat, int g_idx, int sig_mat, int sig_off, int n_params, int flags,
                        const float* rec, float* dev_grad) {
  int rc = check_render_args(h, width, height, first_block, n_blocks, max_depth, stream);
  if (rc) return rc;
  if (!dev_d_accum || !dev_grad || !alb_off) return fail(SR_EINVAL, "null array");
  if (n_params <= 0 || n_params > SR_MAX_PARAMS) return fail(SR_EINVAL, "n_params=%d outside [1, %d]", n_params,
                                                              SR_MAX_PARAMS);
  const int B = h->B, M = h->n_mats;
  for (int m = 0; m < M; ++m)
    if (alb_off[m] >= 0 && alb_off[m] + B > n_params) return fail(SR_EINVAL, "alb_off[%d] out of range", m);
  if (rough_mat >= M || g_mat >= M || sig_mat >= M) return fail(SR_EINVAL, "parameter material out of range");
  if (rough_mat >= 0 && (rough_idx < 0 || rough_idx >= n_params)) return fail(SR_EINVAL, "bad rough_idx");
  if (g_mat >= 0 && (g_idx < 0 || g_idx >= n_params)) return fail(SR_EINVAL, "bad g_idx");
  if (sig_mat >= 0 && (sig_off < 0 || sig_off + B > n_params)) return fail(SR_EINVAL, "bad sig_off");
  HIPCHK(hipSetDevice(h->device));
  if (n_blocks == 0) return 0;
  if ((rc = ensure_sum(h, width, height, h->st))) return rc;
  PathArgs a;
  // Paths are replayed only until their throughput is exactly 0: every
  // gradient term after that point is R x (...) or a contribution's own
  // derivative, and R (the radiance still to come) and the contributions are
  // exactly 0 there, so the gradient is the same whether or not they run.
  if ((rc = prepare_path_args(h, a, handle_bufs(h), width, height, seed, stream, first_block, n_blocks, max_depth,
                               SR_RENDER_END_DEAD_PATHS, true, const_cast<float*>(rec))))
    return rc;
  if ((rc = h->alb_off.upload(alb_off, (size_t)M, h->st))) return rc;
  a.adj.d_accum = dev_d_accum;
  a.adj.grad = dev_grad;
  a.adj.alb_off = (const int*)h->alb_off.p;
  a.adj.rough_mat = rough_mat < 0 ? -2 : rough_mat;
  a.adj.rough_idx = rough_idx;
  a.adj.g_mat = g_mat < 0 ? -2 : g_mat;
  a.adj.g_idx = g_idx;
  a.adj.sig_mat = sig_mat < 0 ? -2 : sig_mat;
  a.adj.sig_off = sig_off;
  a.adj.n_params = n_params;
  a.adj.flags = flags;
  a.adj.scale = 1.0f / (float)(first_block + n_blocks);
  hipStream_t st = h->st;
  int grid = h->adj_grid;
  const unsigned long long waves_needed = ((unsigned long long)h->job_host.n_items + 255) / 256;
  if ((unsigned long long)grid > waves_needed) grid = (int)waves_needed;
  if (queue_overflows(h->job_host.n_items, grid, 256, a.chunk)) return fail(SR_EINVAL, "work queue would wrap");
  HIPCHK(hipEventRecord(h->ev0, st));
  if (sr_launch_adjoint(a, grid, 256, st)) return fail(SR_EHIP, "adjoint kernel launch failed");
  HIPCHK(hipEventRecord(h->ev1, st));
  h->timed = true;
  HIPCHK(hipStreamSynchronize(st));  // alb_off staging buffer is reused by the next call
  return 0;
}

int sr_render_adjoint(sr_handle* h, int width, int height, uint64_t seed, uint32_t stream, int first_block,
                      int n_blocks, int max_depth, const float* dev_d_accum, const int* alb_off, int rough_mat,
                      int rough_idx, int g_mat, int g_idx, int sig_mat, int sig_off, int n_params, int flags,
                      float* dev_grad) {
  g_err.clear();
  return adjoint_with(h, width, height, seed, stream, first_block, n_blocks, max_depth, dev_d_accum, alb_off,
                      rough_mat, rough_idx, g_mat, g_idx, sig_mat, sig_off, n_params, flags, nullptr, dev_grad);
}

int sr_render_adjoint_record(sr_handle* h, int width, int height, uint64_t seed, uint32_t stream, int first_block,
                             int n_blocks, int max_depth, const float* dev_d_accum, const int* alb_off,
                             int rough_mat, int rough_idx, int g_mat, int g_idx, int sig_mat, int sig_off,
                             int n_params, int flags, const float* dev_path_totals, float* dev_grad) {
  g_err.clear();
  if (!h) return fail(SR_EINVAL, "null handle");
  if (!dev_path_totals) return fail(SR_EINVAL, "dev_path_totals is NULL");
  return adjoint_with(h, width, height, seed, stream, first_block, n_blocks, max_depth, dev_d_accum, alb_off,
                      rough_mat, rough_idx, g_mat, g_idx, sig_mat, sig_off, n_params, flags, dev_path_totals,
                      dev_grad);
}

int sr_render_normals(sr_handle* h, int width, int height, int out_w, int out_h, float* dev_rgb) {
  g_err.clear();
  if (!h) return fail(SR_EINVAL, "null handle");
  if (!h->has_scene) return fail(SR_ESTATE, "no scene uploaded");
  if (!dev_rgb) return fail(SR_EINVAL, "null array");
  if (width <= 0 || height <= 0 || out_w <= 0 || out_h <= 0) return fail(SR_EINVAL, "bad size");
  if ((long long)out_w * out_h >= (1LL << 31)) return fail(SR_EINVAL, "display grid too large");
  HIPCHK(hipSetDevice(h->device));
  if ((long long)width * height >= (1LL << 31)) return fail(SR_EINVAL, "render grid too large");
  SceneDev s;
  fill_scene_dev(h, s);
  float* tmp = nullptr;
  if (out_w != width || out_h != height) {
    int rc = h->rgb_tmp.ensure((size_t)width * height * 3 * sizeof(float));
    if (rc) return rc;
    tmp = (float*)h->rgb_tmp.p;
  }
  if (sr_launch_normals(s, width, height, out_w, out_h, dev_rgb, tmp, (uint32_t*)h->overflow.p, h->st))
    return fail(SR_EHIP, "normals launch failed");
  return 0;
}

int sr_to_rgba8(sr_handle* h, const float* dev_rgb, int64_t n_pixels, void* dev_rgba) {
  g_err.clear();
  if (!h) return fail(SR_EINVAL, "null handle");
  if (!dev_rgb || !dev_rgba) return fail(SR_EINVAL, "null array");
  if (n_pixels < 0) return fail(SR_EINVAL, "bad size");
  HIPCHK(hipSetDevice(h->device));
  if (sr_launch_rgba8(dev_rgb, (size_t)n_pixels, dev_rgba, h->st)) return fail(SR_EHIP, "rgba8 launch failed");
  return 0;
}

int sr_resolve(sr_handle* h, const float* dev_accum, int width, int height, const float* cmf, const float* xyz2rgb,
               float norm, float exposure, float gamma, int out_w, int out_h, float* dev_rgb) {
  g_err.clear();
  if (!h) return fail(SR_EINVAL, "null handle");
  if (!dev_accum || !dev_rgb || !cmf || !xyz2rgb) return fail(SR_EINVAL, "null array");
  if (width <= 0 || height <= 0 || out_w <= 0 || out_h <= 0) return fail(SR_EINVAL, "bad size");
  HIPCHK(hipSetDevice(h->device));
  ResolveParams p;
  std::memset(&p, 0, sizeof(p));
  std::memcpy(p.cmf, cmf, sizeof(float) * 3 * h->B);
  std::memcpy(p.T, xyz2rgb, sizeof(float) * 9);
  p.step = h->step;
  p.norm = norm;
  p.exposure = exposure;
  p.gamma = gamma;
  float* tmp = nullptr;
  if (out_w != width || out_h != height) {
    int rc = h->rgb_tmp.ensure((size_t)width * height * 3 * sizeof(float));
    if (rc) return rc;
    tmp = (float*)h->rgb_tmp.p;
  }
  if (sr_launch_resolve(dev_accum, width, height, h->B, p, dev_rgb, out_w, out_h, tmp, h->st))
    return fail(SR_EHIP, "resolve launch failed");
  return 0;
}

int sr_stats(sr_handle* h, uint64_t* segments, uint64_t* traversals, uint64_t* nonfinite) {
  if (!h) return fail(SR_EINVAL, "null handle");
  HIPCHK(hipSetDevice(h->device));
  unsigned long long c[4] = {0, 0, 0, 0};
  uint32_t ovf = 0;
  void* work = h->last_work ? h->last_work : h->work.p;
  uint32_t* ovp = h->last_overflow ? h->last_overflow : (uint32_t*)h->overflow.p;
  HIPCHK(hipMemcpyAsync(c, work_view(work, 1).counters, sizeof(c), hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipMemcpyAsync(&ovf, ovp, 4, hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  if (segments) *segments = c[0];
  if (traversals) *traversals = c[1];
  if (nonfinite) *nonfinite = c[2];
  if (ovf) return fail(SR_EOVERFLOW, "traversal stack overflow in the last render");
  return 0;
}

int sr_last_kernel_ms(sr_handle* h, float* ms) {
  if (!h || !ms) return fail(SR_EINVAL, "null argument");
  if (!h->timed) return fail(SR_ESTATE, "no render recorded");
  HIPCHK(hipEventSynchronize(h->ev1));
  HIPCHK(hipEventElapsedTime(ms, h->ev0, h->ev1));
  return 0;
}

int sr_sync(sr_handle* h) {
  if (!h) return fail(SR_EINVAL, "null handle");
  HIPCHK(hipStreamSynchronize(h->st));
  return 0;
}


// ---------------------------------------------------------------- batches
static void batch_drop_graph(sr_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

int sr_batch_create(sr_batch** out) {
  g_err.clear();
  if (!out) return fail(SR_EINVAL, "out is NULL");
  sr_batch* b = new sr_batch();
  hipError_t e = hipGetDevice(&b->device);
  if (e == hipSuccess) e = hipStreamCreate(&b->st);
  if (e == hipSuccess) e = hipEventCreate(&b->ev0);
  if (e == hipSuccess) e = hipEventCreate(&b->ev1);
  int khz = 0;
  if (e == hipSuccess) e = hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, b->device);
  if (e != hipSuccess) {
    delete b;
    return fail(SR_EHIP, "HIP init failed: %s", hipGetErrorString(e));
  }
  if (khz > 0) b->clock_khz = (double)khz;
  int rc = b->overflow.ensure(64);
  if (rc) {
    delete b;
    return rc;
  }
  (void)hipMemset(b->overflow.p, 0, 64);
  *out = b;
  return 0;
}

int sr_batch_destroy(sr_batch* b) {
  if (!b) return 0;
  (void)hipSetDevice(b->device);
  if (b->st) (void)hipStreamSynchronize(b->st);
  batch_drop_graph(b);
  b->jobs.release();
  b->work.release();
  b->overflow.release();
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->st) (void)hipStreamDestroy(b->st);
  delete b;
  return 0;
}

int sr_batch_submit(sr_batch* b, sr_handle* h, int width, int height, uint64_t seed, uint32_t stream,
                    int first_block, int n_blocks, int max_depth, int flags, float* dev_accum) {
  g_err.clear();
  if (!b) return fail(SR_EINVAL, "null batch");
  if (flags & ~SR_RENDER_FLAGS) return fail(SR_EINVAL, "unknown render flags 0x%x", flags);
  int rc = check_render_args(h, width, height, first_block, n_blocks, max_depth, stream);
  if (rc) return rc;
  if (!dev_accum) return fail(SR_EINVAL, "dev_accum is NULL");
  if (h->device != b->device) return fail(SR_EINVAL, "scene handle is on another device");
  for (const auto& sp : b->specs)
    if (sp.h == h) return fail(SR_EINVAL, "a scene handle may appear once per batch (its block-sum buffer)");
  b->specs.push_back({h, width, height, first_block, n_blocks, max_depth, seed, stream, dev_accum, flags, 0});
  b->dirty = true;
  return 0;
}

int sr_batch_clear(sr_batch* b) {
  if (!b) return fail(SR_EINVAL, "null batch");
  (void)hipStreamSynchronize(b->st);
  b->specs.clear();
  b->dirty = true;
  batch_drop_graph(b);
  return 0;
}

int sr_batch_config(sr_batch* b, int block, int grid_pct, int chunk) {
  if (!b) return fail(SR_EINVAL, "null batch");
  if (block != 64 && block != 128 && block != 256) return fail(SR_EINVAL, "block must be 64, 128 or 256");
  if (grid_pct < 10 || grid_pct > 400) return fail(SR_EINVAL, "grid_pct outside [10, 400]");
  if (chunk < 1 || chunk > (1 << 20)) return fail(SR_EINVAL, "bad chunk");
  b->block = block;
  b->grid_pct = grid_pct;
  b->chunk = (uint32_t)chunk;
  b->dirty = true;  // the captured graph bakes the launch shape in
  return 0;
}

static bool batch_walks(const sr_batch* b) {
  for (const auto& sp : b->specs)
    if (sp.flags & SR_RENDER_SUBSURFACE) return true;
  return false;
}

static int batch_enqueue(sr_batch* b, int grid) {
  const int n = (int)b->specs.size();
  hipStream_t st = b->st;
  if (reset_work(b->work.p, n, st)) return fail(SR_EHIP, "reset launch failed");
  PathArgs a;
  std::memset(&a, 0, sizeof(a));
  WorkView w = work_view(b->work.p, n);
  a.jobs = (const Job*)b->jobs.p;
  a.n_jobs = n;
  a.queues = w.queues;
  a.done_items = w.done;
  a.clocks = w.clocks;
  a.counters = w.counters;
  a.overflow = (uint32_t*)b->overflow.p;
  a.chunk = b->chunk;
  a.tab_bytes = b->tab_bytes;
  a.adj.rough_mat = a.adj.g_mat = a.adj.sig_mat = -2;
  a.adj.rough_idx = a.adj.g_idx = a.adj.sig_off = -1;
  const int lrc = batch_walks(b) ? sr_launch_path_sss(a, grid, b->block, st) : sr_launch_path(a, grid, b->block, st);
  if (lrc) return fail(SR_EHIP, "path kernel launch failed");
  for (int j = 0; j < n; ++j) {
    const auto& sp = b->specs[j];
    const size_t cnt = (size_t)sp.W * sp.H * sp.h->B;
    if (sr_launch_finalize(sp.accum, b->host_jobs[j].sum, cnt, sp.first, sp.nb, st))
      return fail(SR_EHIP, "finalize launch failed");
  }
  return 0;
}

int sr_batch_launch(sr_batch* b, int flags) {
  g_err.clear();
  if (!b) return fail(SR_EINVAL, "null batch");
  if (b->specs.empty()) return fail(SR_ESTATE, "empty batch");
  HIPCHK(hipSetDevice(b->device));
  const int n = (int)b->specs.size();
  // a job table (and a captured graph) holds each scene's device pointers,
  // camera and block-sum buffer: re-fill it when any of them changed
  for (const auto& sp : b->specs)
    if (sp.gen != sp.h->gen) b->dirty = true;
  if (b->dirty) {
    HIPCHK(hipStreamSynchronize(b->st));
    batch_drop_graph(b);
    b->host_jobs.resize(n);
    for (int j = 0; j < n; ++j) {
      auto& sp = b->specs[j];
      int rc = ensure_sum(sp.h, sp.W, sp.H, b->st);
      if (rc) return rc;
      sp.gen = sp.h->gen;
      fill_job(sp.h, b->host_jobs[j], sp.W, sp.H, sp.seed, sp.stream, sp.first, sp.nb, sp.depth);
      b->host_jobs[j].nee_tree = ((sp.flags & SR_RENDER_LIGHT_TREE) && sp.h->n_lnodes > 0) ? 1 : 0;
      b->host_jobs[j].sss = (sp.flags & SR_RENDER_SUBSURFACE) ? 1 : 0;
      b->host_jobs[j].end_dead = (sp.flags & SR_RENDER_END_DEAD_PATHS) ? 1 : 0;
      b->host_jobs[j].tab =
          sr_lds_tables(sp.h->n_mats, sp.h->B, sp.h->B * std::max(sp.h->max_lights, 1), sp.h->n_slots, b->block, false);
    }
    // the jobs' tables share one LDS region per block (the block's waves
    // switch jobs together): every job's must fit, or none goes to LDS
    b->tab_bytes = 0;
    bool all_fit = true;
    for (const Job& jb : b->host_jobs) {
      all_fit = all_fit && jb.tab.bytes > 0;
      b->tab_bytes = std::max(b->tab_bytes, jb.tab.bytes);
    }
    if (!all_fit) {
      b->tab_bytes = 0;
      for (Job& jb : b->host_jobs) std::memset(&jb.tab, 0, sizeof(jb.tab));
    }
    int rc = b->jobs.upload(b->host_jobs.data(), (size_t)n, b->st);
    if (!rc) rc = b->work.ensure(work_bytes(n));
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(b->st));
    b->dirty = false;
  }
  int per = 0;
  if ((batch_walks(b) ? sr_path_sss_grid_bs(b->block, &per) : sr_path_grid_bs(b->block, &per)) != 0)
    return fail(SR_EHIP, "occupancy query failed");
  int grid = std::max(1, per * b->grid_pct / 100);
  {
    uint32_t max_items = 0;
    for (const auto& j : b->host_jobs) max_items = std::max(max_items, j.n_items);
    if (queue_overflows(max_items, grid, b->block, b->chunk))
      return fail(SR_EINVAL, "queue counter would wrap: %u items + %d waves x chunk %u >= 2^32", max_items,
                  grid * b->block / 64, b->chunk);
  }
  HIPCHK(hipEventRecord(b->ev0, b->st));
  if (flags & SR_BATCH_GRAPH) {
    if (!b->exec) {
      HIPCHK(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
      int rc = batch_enqueue(b, grid);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(b->st, &g);
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      if (e != hipSuccess) return fail(SR_EHIP, "graph capture failed: %s", hipGetErrorString(e));
      b->graph = g;
      HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    HIPCHK(hipGraphLaunch(b->exec, b->st));
  } else {
    int rc = batch_enqueue(b, grid);
    if (rc) return rc;
  }
  HIPCHK(hipEventRecord(b->ev1, b->st));
  b->launched = true;
  return 0;
}

int sr_batch_wait(sr_batch* b, float* total_ms, float* job_ms, uint64_t* job_segments, uint64_t* job_traversals) {
  if (!b) return fail(SR_EINVAL, "null batch");
  if (!b->launched) return fail(SR_ESTATE, "nothing launched");
  HIPCHK(hipSetDevice(b->device));
  HIPCHK(hipEventSynchronize(b->ev1));
  const int n = (int)b->specs.size();
  if (total_ms) HIPCHK(hipEventElapsedTime(total_ms, b->ev0, b->ev1));
  std::vector<unsigned long long> clk((size_t)n + 1), cnt((size_t)4 * n);
  WorkView w = work_view(b->work.p, n);
  uint32_t ovf = 0;
  HIPCHK(hipMemcpy(clk.data(), w.clocks, sizeof(unsigned long long) * (n + 1), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(cnt.data(), w.counters, sizeof(unsigned long long) * 4 * n, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(&ovf, b->overflow.p, 4, hipMemcpyDeviceToHost));
  for (int j = 0; j < n; ++j) {
    if (job_ms) job_ms[j] = (float)((double)(clk[1 + j] - clk[0]) / b->clock_khz);
    if (job_segments) job_segments[j] = cnt[4 * j];
    if (job_traversals) job_traversals[j] = cnt[4 * j + 1];
  }
  if (ovf) return fail(SR_EOVERFLOW, "traversal stack overflow in the batch");
  return 0;
}

}  // extern "C"
