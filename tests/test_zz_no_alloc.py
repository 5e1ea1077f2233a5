"""sr_render_noalloc -- the render entry point that allocates nothing (its
block sums, job table and counters live in a fixed device array) -- against
sr_render through the C ABI: two calls in a row, the second with a smaller
image, each equal to sr_render's result for the same blocks; sizes beyond
the array are refused."""
import ctypes

import numpy as np
import pytest
import torch

from spectral import _lib, configs
from spectral.render import Renderer
from spectral.scene import flatten_scene

pytestmark = pytest.mark.gpu

MAX_ELEMS = 1 << 22  # SR_NOALLOC_MAX_ELEMS


def _noalloc(r, W, H, seed, first, n, depth, acc, flags=0):
    _lib.call("sr_render_noalloc", r._h, W, H, ctypes.c_uint64(seed), ctypes.c_uint32(0), first, n, depth, flags,
              _lib.ptr(acc))


def _reference(flat, c, W, H, n, gpu, flags=0):
    ref = Renderer(flat, c.bands, c.wl_min, c.wl_max, width=W, height=H, device=gpu)
    ref.render(n_blocks=n, max_depth=c.depth, seed=c.seed, light_tree=bool(flags & 1))
    return ref.accum.cpu().numpy(), ref.stats()


def test_noalloc_matches_sr_render_twice_smaller_second(gpu):
    c = configs.box_icosphere()
    flat = flatten_scene(c.scene, c.bands, c.wl_min, c.wl_max)
    r = Renderer(flat, c.bands, c.wl_min, c.wl_max, device=gpu)  # no resolution: no block-sum buffer
    for (W, H, n) in ((96, 80, 6), (40, 24, 3)):
        want, st_want = _reference(flat, c, W, H, n, gpu)
        acc = torch.zeros((W, H, c.bands), dtype=torch.float32, device=gpu)
        _noalloc(r, W, H, c.seed, 0, n, c.depth, acc)
        torch.cuda.synchronize()
        got = acc.cpu().numpy()
        assert np.allclose(got, want, rtol=1e-6, atol=1e-7), (W, H, np.abs(got - want).max())
        st = r.stats()
        assert st["segments"] == st_want["segments"] and st["traversals"] == st_want["traversals"]
    # a running mean continued over two calls equals one call over both ranges
    W, H = 40, 24
    want, _ = _reference(flat, c, W, H, 5, gpu)
    acc = torch.zeros((W, H, c.bands), dtype=torch.float32, device=gpu)
    _noalloc(r, W, H, c.seed, 0, 2, c.depth, acc)
    _noalloc(r, W, H, c.seed, 2, 3, c.depth, acc)
    assert np.allclose(acc.cpu().numpy(), want, rtol=1e-5, atol=1e-7)


def test_noalloc_refuses_sizes_beyond_its_array(gpu):
    c = configs.box_icosphere()
    flat = flatten_scene(c.scene, c.bands, c.wl_min, c.wl_max)
    r = Renderer(flat, c.bands, c.wl_min, c.wl_max, device=gpu)
    W = 512
    H = MAX_ELEMS // (W * c.bands) + 1  # one row too many
    acc = torch.zeros((1,), dtype=torch.float32, device=gpu)
    with pytest.raises(_lib.SpectralRenderError, match="exceeds"):
        _noalloc(r, W, H, c.seed, 0, 1, c.depth, acc)
    # the largest image that fits is accepted (zero blocks: nothing runs)
    _noalloc(r, W, MAX_ELEMS // (W * c.bands), c.seed, 0, 0, c.depth, acc)
