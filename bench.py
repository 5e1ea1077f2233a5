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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="heightfield_1m")
    ap.add_argument("--res", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (HBM counter calibration)")
    args = ap.parse_args()
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

    for _ in range(args.warmup):
        r.render(n_blocks=1, max_depth=cfg.depth, seed=cfg.seed)
    torch.cuda.synchronize()
    segs = travs = nonfinite = 0
    kms = 0.0
    torch.cuda.synchronize()
    t0 = time.time()
    for _ in range(args.steps):
        r.render(n_blocks=1, max_depth=cfg.depth, seed=cfg.seed)
        kms += r.last_kernel_ms()
        st = r.stats()
        segs += st["segments"]
        travs += st["traversals"]
        nonfinite += st["nonfinite"]
    torch.cuda.synchronize()
    wall = time.time() - t0
    ms_step = wall * 1e3 / args.steps
    value = segs / wall / 1e6
    pixels = cfg.width * cfg.height
    abytes = algorithmic_bytes(flat.n_tris, travs, cfg.bands, pixels, args.steps)
    achieved = abytes / (kms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None}
    pmc = ROOT / "profiles" / f"{cfg.name}_pmc.json"
    if pmc.exists():
        try:
            roof["traffic"] = json.loads(pmc.read_text()).get("hbm_bytes_per_launch")
        except Exception:
            pass
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
        "nonfinite": nonfinite,
        "roofline": roof,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(flat, cfg, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
