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


def test_single_triangle_leaf_is_hit(gpu):
    # One face: its leaf is (first = 0, count = 1), the encoding that once
    # collided with the empty-slot sentinel.
    from spectral.render import Renderer

    B = 4
    scene = {"materials": [configs._mat(B, n=1.0, tv=1), configs._mat(B, n=1.4, ss=1.0)],
             "meshes": [{"vertices": [[-1, -1, 0], [1, -1, 0], [0, 1, 0]], "faces": [[0, 1, 2]], "material_id": 1}],
             "camera": {"origin": [0, 0, 2], "lookat": [0, 0, 0], "up": [0, 1, 0]}}
    flat = flatten_scene(scene, B, 400.0, 700.0)
    r = Renderer(flat, B, 400.0, 700.0, device=gpu)
    t, p, _ = r.trace(np.array([[0, 0, 2], [5, 5, 2]], np.float32), np.array([[0, 0, -1], [0, 0, -1]], np.float32))
    assert list(p) == [0, -1]
    assert t[0] == np.float32(2.0)
