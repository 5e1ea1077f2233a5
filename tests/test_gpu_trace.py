"""Closest-hit parity of the HIP BVH4 traversal (sr_trace_closest) against
brute force over all faces (the oracle), bit-exact on ids and t."""
import numpy as np
import pytest

from oracle.oracle import OracleScene
from spectral import configs
from spectral.scene import flatten_scene, pixel_center_rays

pytestmark = pytest.mark.gpu


def _random_rays(n, seed, lo=-0.49, hi=0.49):
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return o, d.astype(np.float32)


@pytest.fixture(scope="module")
def box(gpu):
    from spectral.render import Renderer

    cfg = configs.box_icosphere()
    flat = flatten_scene(cfg.scene, cfg.bands, cfg.wl_min, cfg.wl_max)
    return flat, Renderer(flat, cfg.bands, cfg.wl_min, cfg.wl_max, device=gpu), OracleScene(flat, brute=True)


def test_primary_rays_bit_exact(box):
    flat, r, ora = box
    o, d = pixel_center_rays(flat, 128, 128)
    t, p, uv = r.trace(o, d)
    t0, p0, uv0 = ora.trace(o, d)
    assert np.array_equal(p, p0)
    assert np.array_equal(t, t0)
    hit = p0 >= 0
    assert np.array_equal(uv[hit], uv0[hit])


def test_random_rays_bit_exact(box):
    flat, r, ora = box
    o, d = _random_rays(20000, 1)
    t, p, _ = r.trace(o, d)
    t0, p0, _ = ora.trace(o, d)
    assert np.array_equal(p, p0)
    assert np.array_equal(t, t0)


def test_device_pointer_path(box):
    import torch

    flat, r, ora = box
    o, d = _random_rays(4096, 2)
    t, p, _ = r.trace(torch.from_numpy(o).cuda(), torch.from_numpy(d).cuda())
    t0, p0, _ = ora.trace(o, d)
    assert np.array_equal(p.cpu().numpy(), p0)
    assert np.array_equal(t.cpu().numpy(), t0)
