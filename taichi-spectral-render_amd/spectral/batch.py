"""Batched multi-scene rendering: many scenes, one persistent launch.

The reference renders one scene held in global Taichi fields with one
kernel launch per band (main.py:361-394, 1683-1709, 1831).  Here a batch
holds jobs (scene handle, resolution, spp, depth, seed, accumulator) and runs
them in ONE persistent kernel launch fed from per-job device queues
(sr_batch_* in include/spectral_render.h), optionally captured once as a
hipGraph and replayed.

``tune()`` is the autotuner: it times the batch under candidate launch
shapes -- workgroup size, grid fraction of the resident maximum and queue
chunk (items claimed per atomic) -- on a reduced-spp copy of the batch and
keeps the fastest.
"""
from __future__ import annotations

import ctypes
import itertools
import time

import numpy as np
import torch

from . import _lib
from .render import Renderer, render_flags

SR_BATCH_GRAPH = 1

def _declare():
    return _lib.load()


class Batch:
    """Jobs over several Renderers, launched together."""

    def __init__(self):
        _declare()
        h = ctypes.c_void_p()
        _lib.call("sr_batch_create", ctypes.byref(h))
        self._b = h
        self.jobs = []
        self.config = (256, 100, 64)

    def close(self):
        if getattr(self, "_b", None) is not None and self._b.value:
            _lib.load().sr_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, renderer: Renderer, n_blocks, max_depth, seed=0, stream=0, light_tree=False, subsurface=False,
               end_dead_paths=False):
        """Add a job: `n_blocks` spectral blocks of `renderer`'s scene into its
        accumulator (continuing from renderer.blocks_done), with the render
        modes of Renderer.render."""
        if renderer.accum is None:
            raise RuntimeError("renderer has no resolution")
        flags = render_flags(light_tree, subsurface, end_dead_paths)
        self._submit(renderer, int(n_blocks), int(max_depth), int(seed), int(stream), flags)

    def _submit(self, renderer, n_blocks, max_depth, seed, stream, flags):
        _lib.call("sr_batch_submit", self._b, renderer._h, renderer.width, renderer.height, ctypes.c_uint64(seed),
                  ctypes.c_uint32(stream), renderer.blocks_done, n_blocks, max_depth, flags,
                  _lib.ptr(renderer.accum))
        self.jobs.append((renderer, n_blocks, max_depth, seed, stream, flags, self._state(renderer)))

    @staticmethod
    def _state(r):
        """What a submitted job captured of its renderer: where the blocks
        continue, and the accumulator they fold into."""
        return (r.blocks_done, r.accum.data_ptr(), r.width, r.height)

    def clear(self):
        _lib.call("sr_batch_clear", self._b)
        self.jobs = []

    def set_config(self, block=256, grid_pct=100, chunk=64):
        _lib.call("sr_batch_config", self._b, int(block), int(grid_pct), int(chunk))
        self.config = (int(block), int(grid_pct), int(chunk))

    def launch(self, graph=True):
        torch.cuda.synchronize()  # accumulators are ready before the batch's own stream starts
        _lib.call("sr_batch_launch", self._b, SR_BATCH_GRAPH if graph else 0)

    def wait(self):
        """Blocks; returns dict with total_ms, per-job latency_ms, segments,
        traversals.  Advances each renderer's block counter."""
        n = len(self.jobs)
        total = ctypes.c_float()
        lat = np.zeros(n, np.float32)
        seg = np.zeros(n, np.uint64)
        trav = np.zeros(n, np.uint64)
        _lib.call("sr_batch_wait", self._b, ctypes.byref(total), _lib.ptr(lat), _lib.ptr(seg), _lib.ptr(trav))
        return {"total_ms": total.value, "latency_ms": lat, "segments": seg, "traversals": trav}

    def run(self, graph=True):
        """Launch, wait and advance each renderer's block counter.  A job
        continues from its renderer's blocks_done into its current
        accumulator: when either moved since the jobs were submitted (a
        previous run, clear_pixel_space, set_resolution) the jobs are
        submitted again first (and a graph is captured again)."""
        if any(self._state(r) != st for (r, *_x, st) in self.jobs):
            jobs = self.jobs
            self.clear()
            for (r, nb, depth, seed, stream, flags, _st) in jobs:
                self._submit(r, nb, depth, seed, stream, flags)
        self.launch(graph)
        out = self.wait()
        for (r, nb, *_rest) in self.jobs:
            r.blocks_done += nb
        return out


def tune(jobs_spec, blocks=(128, 256), grid_pcts=(100,), chunks=(64, 256, 1024), graph=True, reps=2, log=None):
    """Autotuner: `jobs_spec` is a list of (renderer, n_blocks, depth, seed).
    Each candidate runs the whole batch `reps` times (accumulators are
    restored afterwards); returns (best_config, timings)."""
    saved = [(r, r.accum.clone(), r.blocks_done) for (r, *_x) in jobs_spec]
    timings = {}
    best = None
    for cfg in itertools.product(blocks, grid_pcts, chunks):
        b = Batch()
        b.set_config(*cfg)
        for (r, nb, depth, seed) in jobs_spec:
            b.submit(r, nb, depth, seed)
        ms = []
        for _ in range(reps):
            ms.append(b.run(graph)["total_ms"])
            for (r, acc, done) in saved:
                r.accum.copy_(acc)
                r.blocks_done = done
        b.close()
        t = min(ms)
        timings[cfg] = t
        if log:
            log(f"tune {cfg}: {t:.2f} ms")
        if best is None or t < timings[best]:
            best = cfg
    return best, timings
