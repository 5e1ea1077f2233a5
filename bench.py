"""Benchmark: Mpath-segments/s of the spectral path loop on one MI355X.

Workload (default): the 1,048,588-triangle heightfield box, 31 bands,
1024x1024, depth 12.  One step = one spectral block = one single-band path
per pixel per band (1024*1024*31 = 32.5M paths) through the persistent
path kernel.  Inputs are resident in HBM before timing starts.

Prints one JSON line (see the keys below).  roofline.achieved uses the
algorithmic bytes of the traversals the kernel counted:
    bytes = 128 B * (D4 + 1) * traversals + 8 B * bands * pixels
with D4 = ceil(log4(ceil(N / 4))), divided by the path kernel's device time
(HIP events on the stream it is launched on).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
for p in (str(ROOT / "taichi-spectral-render_amd"), str(ROOT)):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0
# the kernel sources a committed rocprofv3 counter summary was measured on
PROFILED_SOURCES = ("taichi-spectral-render_amd/csrc/sr_path.hip", "taichi-spectral-render_amd/csrc/sr_device.h",
                    "taichi-spectral-render_amd/csrc/sr_layout.h", "taichi-spectral-render_amd/csrc/sr_kernels.h",
                    "taichi-spectral-render_amd/csrc/sr_dual.h", "taichi-spectral-render_amd/Makefile")


def sources_sha256():
    import hashlib

    h = hashlib.sha256()
    for rel in PROFILED_SOURCES:
        h.update((ROOT / rel).read_bytes())
    return h.hexdigest()


def profiled_traffic(name):
    """HBM bytes per launch from the committed rocprofv3 FETCH/WRITE summary,
    only if it was measured on the kernel sources of this tree."""
    pmc = ROOT / "profiles" / f"{name}_pmc.json"
    if not pmc.exists():
        return None, "no counter profile committed"
    try:
        d = json.loads(pmc.read_text())
    except Exception:
        return None, "unreadable counter profile"
    if d.get("sources_sha256") != sources_sha256():
        return None, "counter profile measured on other kernel sources: not reported"
    return d.get("hbm_bytes_per_launch"), None


def d4_levels(n_tris):
    return int(math.ceil(math.log(max(math.ceil(n_tris / 4), 1), 4))) if n_tris > 4 else 1


def algorithmic_bytes(n_tris, traversals, bands, pixels, blocks):
    return 128.0 * (d4_levels(n_tris) + 1) * traversals + 8.0 * bands * pixels * blocks


def cpu_baseline(flat, cfg, target_s=15.0):
    """The CPU restatement (oracle) on a bounded sample of the same workload:
    one block over a band of pixel columns, all host cores."""
    import numpy as np

    from oracle.oracle import OracleScene

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t0 = time.time()
    ora = OracleScene(flat)
    build_s = time.time() - t0
    # calibrate on 4 columns, then size the sample to ~target_s
    cols = 4
    t0 = time.time()
    _, st = ora.render(cfg.width, cfg.height, cfg.seed, 0, 1, cfg.depth, cols=(0, cols))
    dt = max(time.time() - t0, 1e-3)
    ncols = int(min(cfg.width - cols, max(1, cols * target_s / dt)))
    c0 = cfg.width // 2 - ncols // 2
    t0 = time.time()
    _, st = ora.render(cfg.width, cfg.height, cfg.seed, 0, 1, cfg.depth, cols=(c0, c0 + ncols))
    dt = time.time() - t0
    return {"value": st["segments"] / dt / 1e6, "unit": "Mpath-segments/s", "cores": threads, "kind": "port",
            "sample": f"1 block x {ncols} of {cfg.width} pixel columns x {cfg.height} rows x {cfg.bands} bands "
                      f"({st['segments']} segments in {dt:.1f} s; BVH build {build_s:.1f} s not timed); "
                      f"C++ restatement of the reference semantics, not Taichi"}


def adjoint_ratio(reps=2):
    """adjoint : forward device-time ratio on the inverse-render workload
    (volume/fluorescence scene, 256x256, 64 spp, depth 12, 126 parameters):
    path kernel time of the adjoint over that of the differentiable forward
    for the same blocks and seed.  The differentiable forward records each
    path's total (sr_render_record) and the adjoint takes it in place of its
    first pass (sr_render_adjoint_record), as RenderFunction runs them; the
    two-pass adjoint and the plain forward are reported beside them."""
    import torch

    from spectral import configs
    from spectral.autograd import DifferentiableScene, inverse_volume_spec
    from spectral.scene import flatten_scene

    cfg = configs.inverse_volume()
    flat = flatten_scene(cfg.scene, cfg.bands, cfg.wl_min, cfg.wl_max)
    ds = DifferentiableScene(flat, inverse_volume_spec(), cfg.width, cfg.height, cfg.spp, cfg.depth, cfg.seed,
                             device="cuda:0")
    d_acc = torch.rand((cfg.width, cfg.height, cfg.bands), device="cuda:0")
    tot = ds.new_path_totals()
    fwd, adj, plain, adj2, full = [], [], [], [], []
    for _ in range(reps + 1):
        ds.render(path_totals=tot)  # the differentiable forward: dead paths ended, as the adjoint replays them
        fwd.append(ds.r.last_kernel_ms())
        segs = ds.r.stats()["segments"]
        ds.adjoint(d_acc, path_totals=tot)
        adj.append(ds.r.last_kernel_ms())
        ds.render()  # the same forward without the record
        plain.append(ds.r.last_kernel_ms())
        ds.adjoint(d_acc)  # every path replayed twice
        adj2.append(ds.r.last_kernel_ms())
        ds.r.clear_pixel_space()  # the same image with every path run to the depth limit
        ds.r.render(n_blocks=ds.spp, max_depth=ds.depth, seed=ds.seed)
        full.append(ds.r.last_kernel_ms())
    fwd, adj, plain, adj2, full = (min(x[1:]) for x in (fwd, adj, plain, adj2, full))
    fluoresced = torch.isnan(tot).float().mean().item()
    untouched = ((tot == 0) & torch.signbit(tot)).float().mean().item()
    return {"value": round(adj / fwd, 3), "forward_ms": round(fwd, 2), "adjoint_ms": round(adj, 2),
            "two_pass_adjoint_ms": round(adj2, 2), "two_pass_ratio": round(adj2 / plain, 3),
            "forward_without_record_ms": round(plain, 2),
            "adjoint_over_full_forward": round(adj / full, 3), "full_forward_ms": round(full, 2),
            "paths_replayed_twice": round(fluoresced, 4), "paths_not_replayed": round(untouched, 4),
            "workload": f"{cfg.name}: {flat.n_tris} tris, {cfg.bands} bands, {cfg.width}x{cfg.height}, "
                        f"{cfg.spp} spp, depth {cfg.depth}, {ds.n_params} params ({segs} segments per forward; "
                        f"forward and adjoint end paths whose throughput is 0, full_forward runs every path "
                        f"to the depth limit; the forward records each path's total and the adjoint replays "
                        f"once every path whose light stayed in its band, and not at all a path that added "
                        f"nothing)"}


def bench_inverse(iterations=3, width=256, height=256, spp=64, log=print):
    """The inverse-render loop on the volume/fluorescence scene (126
    parameters, depth 12): seconds per iteration (forward + loss + adjoint
    + Adam step) after one untimed iteration, and the two kernels' times."""
    import torch

    from spectral.inverse import InverseRender

    inv = InverseRender(width=width, height=height, spp=spp, depth=12, device="cuda:0")
    inv.step()
    torch.cuda.synchronize()
    t0 = time.time()
    fwd = adj = 0.0
    for _ in range(iterations):
        inv.step()
        adj += inv.ds.r.last_kernel_ms()
    torch.cuda.synchronize()
    dt = (time.time() - t0) / iterations
    inv.ds.render(stream=99)
    fwd = inv.ds.r.last_kernel_ms()
    return {"seconds_per_iteration": round(dt, 4), "forward_kernel_ms": round(fwd, 2),
            "adjoint_kernel_ms": round(adj / iterations, 2), "iterations_timed": iterations,
            "loss_first_last": [round(inv.history.loss[0], 5), round(inv.history.loss[-1], 5)],
            "workload": f"inverse_volume: {width}x{height}, {spp} spp, depth 12, 126 params, Adam"}


def bench_batch(spp_scale=8, graph=True, log=print, build="host"):
    """64-scene batch in one persistent launch (hipGraph) against sequential
    per-scene launches of the same jobs.  spp is divided by spp_scale to
    bound the run time (full spp: ~19G paths).  ``build`` chooses where the
    64 scenes are compiled (host or device; the trees are identical); the
    scene-compile time is reported beside the render figures."""
    import numpy as np
    import torch

    from spectral import configs
    from spectral.batch import Batch, tune
    from spectral.render import Renderer
    from spectral.scene import flatten_scene

    cfgs = configs.batch64()
    flats = [flatten_scene(c.scene, c.bands, c.wl_min, c.wl_max) for c in cfgs]
    torch.cuda.synchronize()
    t0 = time.time()
    rs = []
    for c, f in zip(cfgs, flats):
        rs.append((c, f.n_tris, Renderer(f, c.bands, c.wl_min, c.wl_max, width=c.width, height=c.height,
                                         device="cuda:0", build=build)))
    torch.cuda.synchronize()
    compile_s = time.time() - t0
    log(f"# batch64: {len(rs)} scenes, {sum(n for _, n, _ in rs)} tris, compiled ({build}) and uploaded in "
        f"{compile_s:.2f} s")
    spp = [max(1, c.spp // spp_scale) for c, _, _ in rs]
    # autotune on one block per scene
    best, timings = tune([(r, 1, c.depth, c.seed) for c, _, r in rs], blocks=(128, 256), grid_pcts=(100,),
                         chunks=(64, 256, 1024), graph=graph, reps=1)
    log(f"# autotuner: best (block, grid%, chunk) = {best}; " +
        ", ".join(f"{k}: {v:.1f} ms" for k, v in timings.items()))
    # batched: one launch over all scenes
    for _, _, r in rs:
        r.clear_pixel_space()
    b = Batch()
    b.set_config(*best)
    for (c, _, r), n in zip(rs, spp):
        b.submit(r, n, c.depth, c.seed)
    b.run(graph)  # capture + warm
    for _, _, r in rs:
        r.clear_pixel_space()
    out = b.run(graph)
    segs = int(np.sum(out["segments"]))
    batched = {"ms": round(out["total_ms"], 2), "Mseg_per_s": round(segs / out["total_ms"] / 1e3, 1),
               "p50_latency_ms": round(float(np.percentile(out["latency_ms"], 50)), 2),
               "p99_latency_ms": round(float(np.percentile(out["latency_ms"], 99)), 2)}
    # sequential: one sr_render launch per scene, back to back, same order
    lat = []
    seq_segs = 0
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    start.record()
    evs = []
    for (c, _, r), n in zip(rs, spp):
        r.clear_pixel_space()
        r.render(n_blocks=n, max_depth=c.depth, seed=c.seed)
        e = torch.cuda.Event(enable_timing=True)
        e.record()
        evs.append(e)
    torch.cuda.synchronize()
    lat = [start.elapsed_time(e) for e in evs]
    seq_ms = lat[-1]
    sequential = {"ms": round(seq_ms, 2), "Mseg_per_s": round(segs / seq_ms / 1e3, 1),
                  "p50_latency_ms": round(float(np.percentile(lat, 50)), 2),
                  "p99_latency_ms": round(float(np.percentile(lat, 99)), 2)}
    return {"metric": "batch64 aggregate Mpath-segments/s and per-scene completion latency",
            "workload": f"64 scenes (seed 2), spp/{spp_scale}, 31 bands, depth 12, "
                        f"{sum(c.width * c.height * n * c.bands for (c, _, _), n in zip(rs, spp))} paths",
            "segments": segs, "tuned_config": {"block": best[0], "grid_pct": best[1], "chunk": best[2]},
            "graph": graph, "build": build, "scene_compile_s": round(compile_s, 3),
            "batched": batched,
            "sequential": sequential}


def bench_trace(n=1 << 22, reps=5, log=print):
    """Closest-hit traversal alone (sr_trace_closest, device pointers) on the
    heightfield_1m scene: coherent pixel-centre primary rays and incoherent
    rays (uniform origins in the box, uniform directions)."""
    import numpy as np
    import torch

    from spectral import configs
    from spectral.render import Renderer
    from spectral.scene import flatten_scene

    cfg = configs.heightfield_1m()
    flat = flatten_scene(cfg.scene, cfg.bands, cfg.wl_min, cfg.wl_max)
    r = Renderer(flat, cfg.bands, cfg.wl_min, cfg.wl_max, device="cuda:0")
    rng = np.random.default_rng(0)
    o_p, d_p = r.primary_rays(2048, 2048)
    o_r = rng.uniform(-0.49, 0.49, (n, 3)).astype(np.float32)
    d_r = rng.normal(size=(n, 3)).astype(np.float32)
    d_r /= np.linalg.norm(d_r, axis=1, keepdims=True)
    out = {}
    for name, (o, d) in {"primary": (o_p, d_p), "random": (o_r, d_r)}.items():
        og = torch.from_numpy(np.ascontiguousarray(o)).cuda()
        dg = torch.from_numpy(np.ascontiguousarray(d)).cuda()
        r.trace(og, dg)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            t, p, _ = r.trace(og, dg)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        out[name] = {"rays": int(og.shape[0]), "ms": round(ms, 3), "Grays_per_s": round(og.shape[0] / ms / 1e6, 3),
                     "hit_fraction": round(float((p >= 0).float().mean()), 4)}
    return {"metric": "closest-hit traversal rate, heightfield_1m (1,048,588 tris)", **out}


def bench_build(reps=3, log=print):
    """Scene compile on the host (C++ SAH BVH, sr_upload_scene) against the
    device (sr_upload_scene_gpu, the same arrays built on the GPU) for heightfield_1m,
    and the path kernel's rate on each tree (one block, same seed)."""
    import numpy as np
    import torch

    from spectral import _lib, configs
    from spectral.render import Renderer
    from spectral.scene import flatten_scene

    cfg = configs.heightfield_1m()
    flat = flatten_scene(cfg.scene, cfg.bands, cfg.wl_min, cfg.wl_max)
    out = {"metric": "scene compile time and path rate by BVH builder, heightfield_1m (1,048,588 tris)"}
    rs = {}
    for build in ("host", "device"):
        t0 = time.time()
        r = Renderer(flat, cfg.bands, cfg.wl_min, cfg.wl_max, width=cfg.width, height=cfg.height, device="cuda:0",
                     build=build)
        torch.cuda.synchronize()
        rs[build] = (r, time.time() - t0)
    # device build alone: inputs already resident, best of reps
    dv = torch.from_numpy(np.ascontiguousarray(flat.tris_v, np.float32)).cuda()
    dn = torch.from_numpy(np.ascontiguousarray(flat.tris_n, np.float32)).cuda()
    dm = torch.from_numpy(np.ascontiguousarray(flat.tri_mat, np.int32)).cuda()
    r = rs["device"][0]
    best_dev = 1e9
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.time()
        _lib.call("sr_upload_scene_gpu", r._h, _lib.ptr(dv), _lib.ptr(dn), _lib.ptr(dm), flat.n_tris,
                  _lib.ptr(flat.mat_f32), _lib.ptr(flat.mat_i32), _lib.ptr(flat.mat_eem), flat.n_mats,
                  _lib.ptr(flat.eem), flat.n_eem, _lib.ptr(flat.cam))
        best_dev = min(best_dev, time.time() - t0)
    best_host = 1e9
    rh = rs["host"][0]
    for _ in range(reps):
        t0 = time.time()
        rh._upload(flat)
        best_host = min(best_host, time.time() - t0)
    for build, (rr, first_s) in rs.items():
        rr.render(n_blocks=1, max_depth=cfg.depth, seed=cfg.seed)  # warm
        rr.sync()
        ms = []
        for k in range(reps):
            rr.clear_pixel_space()
            rr.render(n_blocks=1, max_depth=cfg.depth, seed=cfg.seed + 1 + k)
            ms.append(rr.last_kernel_ms())
        st = rr.stats()
        info = rr.scene_info()
        out[build] = {"first_upload_s": round(first_s, 3),
                      "compile_s": round(best_host if build == "host" else best_dev, 4),
                      "bvh_nodes": info["bvh_nodes"], "bvh4_depth": info["bvh4_depth"],
                      "kernel_ms_per_block": round(float(np.median(ms)), 3),
                      "Mseg_per_s": round(st["segments"] / float(np.median(ms)) / 1e3, 1)}
    out["note"] = ("host compile_s = flattened arrays -> normals, lights, SAH BVH and upload (C++); device "
                   "compile_s = sr_upload_scene_gpu with the triangles already in HBM")
    return out


def bench_rows(reps=5, log=print, only=None):
    """Throughput of the rows around the path: the display path on
    heightfield_1m at 1024x1024 (normals view: one closed-hit trace per
    pixel; RGBA8 conversion: 12 B read + 4 B written per pixel; CMF resolve
    of the 31-band accumulator: 124 B read + 12 B written per pixel, and from
    a 512 x 512 accumulator to a 1024 x 1024 display), NEE through
    the light tree against the area CDF (time, pixel variance over 8 blocks and
    efficiency = 1 / (variance x time)) on 128 lamps over 3.5 units, on
    1024 lamps over 32 units and on the box whose icosphere emits in bands
    >= 8 (1282 lights facing every way), and the subsurface walk on the
    closed slab (w = 0.9), all at 1024x1024 so that one block fills the
    persistent grid."""
    import numpy as np
    import torch

    from spectral import configs
    from spectral.render import Renderer
    from spectral.scene import flatten_scene

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        best = 1e9
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            torch.cuda.synchronize()
            best = min(best, e0.elapsed_time(e1))
        return best

    out = {"metric": "display path, light-tree NEE and subsurface walk throughput"}
    if only is not None:
        out["metric"] = f"{only} row of the display / light-tree / subsurface rows"
        return bench_rows_paths(out, reps, only)
    cfg = configs.heightfield_1m()
    flat = flatten_scene(cfg.scene, cfg.bands, cfg.wl_min, cfg.wl_max)
    r = Renderer(flat, cfg.bands, cfg.wl_min, cfg.wl_max, width=1024, height=1024, device="cuda:0")
    box = {}
    ms_n = timed(lambda: box.__setitem__("rgb", r.visualize_normals(1024, 1024)))
    ms_c = timed(lambda: r.to_rgba8(box["rgb"]))
    npx = 1024 * 1024
    out["normals_view"] = {"pixels": npx, "ms": round(ms_n, 3), "Mpixels_per_s": round(npx / ms_n / 1e3, 1)}
    # the reference's default display, twice the render grid: one ray per
    # render pixel, then the gather to 2048 x 2048
    ms_n2 = timed(lambda: r.visualize_normals(2048, 2048))
    out["normals_view_x2"] = {"render_pixels": npx, "display_pixels": 4 * npx, "ms": round(ms_n2, 3),
                              "display_Mpixels_per_s": round(4 * npx / ms_n2 / 1e3, 1)}
    gbs = npx * 16 / ms_c / 1e6
    out["rgba8"] = {"pixels": npx, "ms": round(ms_c, 4), "GB_per_s": round(gbs, 1), "frac_of_8TBps": round(gbs / 8000, 3),
                    "note": "16 MB per frame: launch-bound at this size"}
    # CMF resolve of a full accumulator (color_rays): reads W x H x B f32, writes W x H x 3
    from spectral.data import srgb_colour_system
    cs = srgb_colour_system(cfg.wl_min, cfg.wl_max, cfg.bands)
    r.render(n_blocks=1, max_depth=cfg.depth, seed=cfg.seed)
    ms_r = timed(lambda: [r.color_rays(cs) for _ in range(20)]) / 20  # back to back: amortises the host call
    rb = npx * (cfg.bands + 3) * 4
    out["resolve"] = {"bytes": rb, "ms": round(ms_r, 4), "GB_per_s": round(rb / ms_r / 1e6, 1),
                      "frac_of_8TBps": round(rb / ms_r / 1e6 / 8000, 3)}
    # the reference's default display: the render grid shown at twice its
    # size (RESOLUTION_SCALE = 2): a 512 x 512 accumulator to 1024 x 1024
    r.set_resolution(512, 512)
    r.accum.uniform_()
    ms_u = timed(lambda: [r.color_rays(cs, out_width=1024, out_height=1024) for _ in range(20)]) / 20
    ub = 512 * 512 * cfg.bands * 4 + npx * 12
    out["resolve_x2"] = {"bytes": ub, "ms": round(ms_u, 4), "GB_per_s": round(ub / ms_u / 1e6, 1),
                         "frac_of_8TBps": round(ub / ms_u / 1e6 / 8000, 3)}
    del r
    return bench_rows_paths(out, reps, None)


def bench_rows_paths(out, reps, only):
    """The light-tree and subsurface rows of bench_rows (all, or the one
    named `only`)."""
    import numpy as np

    from spectral import configs
    from spectral.render import Renderer
    from spectral.scene import flatten_scene

    sphere = configs.box_icosphere(width=1024, height=1024)
    sphere.scene["materials"][5] = configs.emitting_sphere_material(sphere.bands, sphere.wl_min, sphere.wl_max)
    rows = (("light_tree", configs.lamps(width=1024, height=1024)),
            ("light_tree_sparse", configs.lamps(width=1024, height=1024, n=32, extent=32.0)),
            ("light_tree_sphere", sphere),
            ("subsurface", configs.subsurface_slab(width=1024, height=1024)))
    for name, c in rows:
        if only is not None and name != only:
            continue
        f = flatten_scene(c.scene, c.bands, c.wl_min, c.wl_max)
        rr = Renderer(f, c.bands, c.wl_min, c.wl_max, width=1024, height=1024, device="cuda:0")
        light = name.startswith("light_tree")
        modes = (("cdf", {}), ("tree", {"light_tree": True})) if light else (("walk", {"subsurface": True}),)
        res = {"workload": f"{c.name}: {f.n_tris} tris, {c.bands} bands, 1024x1024, depth {c.depth}"}
        for mode, flags in modes:
            rr.render(n_blocks=1, max_depth=c.depth, seed=c.seed, **flags)
            rr.sync()
            ms = []
            s1 = s2 = None
            n_var = 8 if light else reps
            for k in range(n_var):
                # block k of one seed: the accumulator holds block / (k + 1)
                rr.accum.zero_()
                rr.blocks_done = k
                rr.render(n_blocks=1, max_depth=c.depth, seed=c.seed, **flags)
                ms.append(rr.last_kernel_ms())
                if light:
                    x = rr.accum.double() * (k + 1)
                    s1 = x if s1 is None else s1 + x
                    s2 = x * x if s2 is None else s2 + x * x
            st = rr.stats()
            m = float(np.median(ms))
            res[mode] = {"kernel_ms_per_block": round(m, 3), "Mseg_per_s": round(st["segments"] / m / 1e3, 1),
                         "traversals_per_segment": round(st["traversals"] / max(st["segments"], 1), 3)}
            if light:
                # per pixel-band sample variance over the blocks, summed over the image
                res[mode]["variance"] = float(((s2 - s1 * s1 / n_var) / (n_var - 1)).sum())
        if light:
            # efficiency = 1 / (variance x time): the tree's gain over the CDF pick
            cdf, tree = res["cdf"], res["tree"]
            res["tree_variance_ratio"] = round(tree["variance"] / cdf["variance"], 4)
            res["tree_efficiency_gain"] = round(cdf["variance"] * cdf["kernel_ms_per_block"]
                                                / (tree["variance"] * tree["kernel_ms_per_block"]), 3)
        out[name] = res
        del rr
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="heightfield_1m",
                    help="heightfield_1m (the headline), trace, build, batch64, rows, or one path row: "
                         "light_tree, light_tree_sparse, light_tree_sphere, subsurface")
    ap.add_argument("--res", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="leave the batch64 key out of the default line")
    ap.add_argument("--no-dead", action="store_true",
                    help="leave out the end_dead_paths key (the same blocks with dead paths ended)")
    ap.add_argument("--no-inverse", action="store_true", help="leave the inverse_render key out")
    ap.add_argument("--no-full", action="store_true",
                    help="leave the full_render key (the configuration's whole render in one launch) out")
    ap.add_argument("--spp-scale", type=int, default=8, help="batch64: divide each scene's spp by this")
    ap.add_argument("--build", choices=("host", "device"), default="host",
                    help="batch64: where the scenes are compiled")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (HBM counter calibration)")
    ap.add_argument("--adjoint-once", action="store_true",
                    help="only run one recording forward and one adjoint on the inverse-render workload "
                         "(counter profiles of the adjoint)")
    args = ap.parse_args()
    if args.adjoint_once:
        import torch

        from spectral import configs
        from spectral.autograd import DifferentiableScene, inverse_volume_spec
        from spectral.scene import flatten_scene

        cfg = configs.inverse_volume()
        flat = flatten_scene(cfg.scene, cfg.bands, cfg.wl_min, cfg.wl_max)
        ds = DifferentiableScene(flat, inverse_volume_spec(), cfg.width, cfg.height, cfg.spp, cfg.depth, cfg.seed,
                                 device="cuda:0")
        d_acc = torch.rand((cfg.width, cfg.height, cfg.bands), device="cuda:0")
        tot = ds.new_path_totals()
        ds.render(path_totals=tot)
        fwd = ds.r.last_kernel_ms()
        ds.adjoint(d_acc, path_totals=tot)
        print(json.dumps({"workload": cfg.name, "forward_ms": round(fwd, 2),
                          "adjoint_ms": round(ds.r.last_kernel_ms(), 2)}), flush=True)
        return
    if args.calibrate_copy:
        import torch

        n = args.calibrate_copy * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_mib": args.calibrate_copy, "bytes_read": n * 4, "bytes_written": n * 4}))
        return
    if args.workload == "trace":
        print(json.dumps(bench_trace()), flush=True)
        return
    if args.workload == "rows":
        print(json.dumps(bench_rows()), flush=True)
        return
    if args.workload in ("light_tree", "light_tree_sparse", "light_tree_sphere", "subsurface"):
        print(json.dumps(bench_rows(only=args.workload)), flush=True)
        return
    if args.workload == "build":
        print(json.dumps(bench_build()), flush=True)
        return
    if args.workload == "batch64":
        print(json.dumps(bench_batch(args.spp_scale, log=lambda m: print(m, file=sys.stderr, flush=True),
                                     build=args.build)),
              flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (multi-GPU sharding is design-only)",
              file=sys.stderr)
        sys.exit(2)

    import torch

    from spectral import configs
    from spectral.render import Renderer
    from spectral.scene import flatten_scene

    t0 = time.time()
    kw = {}
    if args.res:
        kw = {"width": args.res, "height": args.res}
    cfg = configs.WORKLOADS[args.workload](**kw)
    flat = flatten_scene(cfg.scene, cfg.bands, cfg.wl_min, cfg.wl_max)
    print(f"# scene {cfg.name}: {flat.n_tris} tris, {cfg.bands} bands, {cfg.width}x{cfg.height}, depth {cfg.depth} "
          f"(built in {time.time() - t0:.1f} s)", file=sys.stderr, flush=True)
    t0 = time.time()
    r = Renderer(flat, cfg.bands, cfg.wl_min, cfg.wl_max, width=cfg.width, height=cfg.height, device="cuda:0")
    info = r.scene_info()
    print(f"# upload+BVH4 {time.time() - t0:.1f} s: {info}", file=sys.stderr, flush=True)

    # warmup: the same launch shape as the timed one, fewer blocks
    if args.warmup:
        r.render(n_blocks=args.warmup, max_depth=cfg.depth, seed=cfg.seed)
    torch.cuda.synchronize()
    # timed: the K steps (spectral blocks) as ONE persistent launch, the way a
    # multi-block render runs; no host reads inside the timed region
    t0 = time.time()
    r.render(n_blocks=args.steps, max_depth=cfg.depth, seed=cfg.seed)
    torch.cuda.synchronize()
    wall = time.time() - t0
    kms = r.last_kernel_ms()
    st = r.stats()
    segs, travs, nonfinite = st["segments"], st["traversals"], st["nonfinite"]
    ms_step = wall * 1e3 / args.steps
    value = segs / wall / 1e6
    pixels = cfg.width * cfg.height
    abytes = algorithmic_bytes(flat.n_tris, travs, cfg.bands, pixels, args.steps)
    achieved = abytes / (kms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None}
    traffic, note = profiled_traffic(cfg.name)
    if traffic is not None:
        # the profile is per one-block launch; this launch holds args.steps blocks
        roof["traffic"] = int(traffic * args.steps)
        roof["traffic_over_algorithmic"] = round(roof["traffic"] / abytes, 4)
        roof["traffic_GBps_of_8TBps"] = round(roof["traffic"] / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    else:
        roof["traffic_note"] = note
    if roof["frac"] > 1.0:
        # the algorithmic model charges every traversal (D4 + 1) node lines of
        # HBM reads; the caches serve most of them, so the model's bytes per
        # second can exceed the HBM peak while the fabric traffic stays far
        # below it (traffic_GBps_of_8TBps)
        roof["frac_note"] = ("above 1: the algorithmic bytes assume no cache reuse; "
                             "the measured HBM traffic is the binding figure")
    out = {
        "metric": "Mpath-segments/s at 1M tris, 31 bands; adjoint:forward time ratio",
        "value": round(value, 2),
        "unit": "Mpath-segments/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{cfg.name}: {flat.n_tris} tris, {cfg.bands} bands, {cfg.width}x{cfg.height}, "
                               f"depth {cfg.depth}, 1 spectral block per step",
                   "bands": cfg.bands, "width": cfg.width, "height": cfg.height, "depth": cfg.depth,
                   "tris": flat.n_tris},
        "segments_per_step": segs // max(args.steps, 1),
        "traversals_per_segment": round(travs / max(segs, 1), 4),
        "kernel_ms_per_step": round(kms / args.steps, 3),
        "launch": f"{args.steps} blocks in one launch ({args.warmup} warmup blocks in one launch before)",
        "nonfinite": nonfinite,
        "roofline": roof,
    }
    if not args.no_dead:
        # the same blocks with SR_RENDER_END_DEAD_PATHS (a path ends once its
        # throughput is exactly 0): the same image, fewer segments; the
        # headline above counts the reference's segments
        r.clear_pixel_space()
        torch.cuda.synchronize()
        t0 = time.time()
        r.render(n_blocks=args.steps, max_depth=cfg.depth, seed=cfg.seed, end_dead_paths=True)
        torch.cuda.synchronize()
        wall_d = time.time() - t0
        st_d = r.stats()
        out["end_dead_paths"] = {"ms_per_step": round(wall_d * 1e3 / args.steps, 3),
                                 "kernel_ms_per_step": round(r.last_kernel_ms() / args.steps, 3),
                                 "segments_per_step": st_d["segments"] // max(args.steps, 1),
                                 "Mseg_per_s": round(st_d["segments"] / wall_d / 1e6, 2),
                                 "block_time_ratio": round(wall_d / wall, 4),
                                 "nonfinite": st_d["nonfinite"]}
    if not args.no_full and cfg.spp > args.steps:
        # the configuration's whole render -- its spp spectral blocks (64 for
        # the 1M-triangle heightfield) -- as the one launch it runs as
        r.clear_pixel_space()
        torch.cuda.synchronize()
        t0 = time.time()
        r.render(n_blocks=cfg.spp, max_depth=cfg.depth, seed=cfg.seed)
        torch.cuda.synchronize()
        wall_f = time.time() - t0
        st_f = r.stats()
        out["full_render"] = {"blocks": cfg.spp, "seconds": round(wall_f, 3),
                              "kernel_seconds": round(r.last_kernel_ms() / 1e3, 3),
                              "segments": st_f["segments"], "Mseg_per_s": round(st_f["segments"] / wall_f / 1e6, 2),
                              "ms_per_block": round(wall_f * 1e3 / cfg.spp, 3), "nonfinite": st_f["nonfinite"]}
    del r
    if not args.no_adjoint:
        out["adjoint_forward_ratio"] = adjoint_ratio()
    if not args.no_inverse:
        out["inverse_render"] = bench_inverse()
    if not args.no_batch:
        b = bench_batch(8, log=lambda m: print(m, file=sys.stderr, flush=True), build="device")
        out["batch64"] = {"workload": b["workload"], "batched": b["batched"], "sequential": b["sequential"],
                          "speedup": round(b["sequential"]["ms"] / b["batched"]["ms"], 3),
                          "scene_compile_s": b["scene_compile_s"], "tuned_config": b["tuned_config"]}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(flat, cfg, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
