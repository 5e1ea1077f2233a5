"""Scene compile on the device (sr_upload_scene_gpu): the linear BVH built on
the GPU gives the same closest hits as brute force (bit-exact ids, t, u, v)
and as the host-built SAH BVH, on the 1292-face box scene and the
1,048,588-face heightfield; the face and light records match the host
build, so matched-seed renders of the two builds agree to float summation
order; degenerate scenes and bad material ids behave as in the host build."""
import numpy as np
import pytest

from oracle.oracle import OracleScene
from spectral import _lib, configs
from spectral.render import Renderer
from spectral.scene import flatten_scene, pixel_center_rays

pytestmark = pytest.mark.gpu


def _random_rays(n, seed, lo=-0.49, hi=0.49):
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return o, d.astype(np.float32)


def _pair(cfg, gpu, **kw):
    flat = flatten_scene(cfg.scene, cfg.bands, cfg.wl_min, cfg.wl_max)
    host = Renderer(flat, cfg.bands, cfg.wl_min, cfg.wl_max, device=gpu, **kw)
    dev = Renderer(flat, cfg.bands, cfg.wl_min, cfg.wl_max, device=gpu, build="device", **kw)
    return flat, host, dev


@pytest.fixture(scope="module")
def box(gpu):
    return _pair(configs.box_icosphere(), gpu)


def test_box_hits_bit_exact_vs_brute_force(box):
    flat, _, dev = box
    ora = OracleScene(flat, brute=True)
    for o, d in (pixel_center_rays(flat, 96, 96), _random_rays(20000, 3)):
        t, p, uv = dev.trace(o, d)
        t0, p0, uv0 = ora.trace(o, d)
        assert np.array_equal(p, p0)
        assert np.array_equal(t, t0)
        hit = p0 >= 0
        assert np.array_equal(uv[hit], uv0[hit])


def test_box_tree_within_stack_bound(box):
    _, host, dev = box
    info = dev.scene_info()
    assert 1 <= info["bvh4_depth"] <= 13
    assert info["max_lights_per_band"] == host.scene_info()["max_lights_per_band"]


def test_box_render_matches_host_build(box):
    """Same hits, same records, same random numbers: the two builds render
    the same paths; only the order of the float atomics differs."""
    _, host, dev = box
    cfg = configs.box_icosphere()
    imgs = []
    for r in (host, dev):
        r.set_resolution(24, 24)
        r.clear_pixel_space()
        r.render(n_blocks=4, max_depth=cfg.depth, seed=5)
        imgs.append(r.accum.cpu().numpy().astype(np.float64))
    a, b = imgs
    assert np.allclose(a, b, rtol=1e-5, atol=1e-7)
    assert host.stats() == dev.stats()


def test_heightfield_1m_hits_match_host_build(gpu):
    cfg = configs.heightfield_1m()
    flat, host, dev = _pair(cfg, gpu)
    info = dev.scene_info()
    assert info["bvh4_depth"] <= 13
    for o, d in (pixel_center_rays(flat, 256, 256), _random_rays(65536, 4)):
        t0, p0, uv0 = host.trace(o, d)
        t, p, uv = dev.trace(o, d)
        assert np.array_equal(p, p0)
        assert np.array_equal(t, t0)
        hit = p0 >= 0
        assert np.array_equal(uv[hit], uv0[hit])


def test_single_triangle_and_empty_scene(gpu):
    cfg = configs.box_icosphere()
    flat = flatten_scene(cfg.scene, cfg.bands, cfg.wl_min, cfg.wl_max)
    one = flat.__class__(**{**flat.__dict__, "tris_v": flat.tris_v[:1].copy(), "tris_n": flat.tris_n[:1].copy(),
                            "tri_mat": flat.tri_mat[:1].copy()})
    r = Renderer(one, cfg.bands, cfg.wl_min, cfg.wl_max, device=gpu, build="device")
    v = one.tris_v[0].astype(np.float64)
    c = v.mean(0)
    nrm = np.cross(v[1] - v[0], v[2] - v[0])
    nrm /= np.linalg.norm(nrm)
    o = (c + 0.3 * nrm)[None].astype(np.float32)
    d = (-nrm)[None].astype(np.float32)
    t, p, _ = r.trace(o, d)
    assert p[0] == 0 and t[0] == pytest.approx(0.3, rel=1e-4)
    empty = flat.__class__(**{**flat.__dict__, "tris_v": flat.tris_v[:0].copy(), "tris_n": flat.tris_n[:0].copy(),
                              "tri_mat": flat.tri_mat[:0].copy()})
    e = Renderer(empty, cfg.bands, cfg.wl_min, cfg.wl_max, width=8, height=8, device=gpu, build="device")
    e.render(n_blocks=1, max_depth=cfg.depth, seed=0)
    assert float(e.accum.abs().sum()) == 0.0


def test_bad_material_id_rejected(gpu):
    cfg = configs.box_icosphere()
    flat = flatten_scene(cfg.scene, cfg.bands, cfg.wl_min, cfg.wl_max)
    bad = flat.tri_mat.copy()
    bad[7] = flat.n_mats + 3
    broken = flat.__class__(**{**flat.__dict__, "tri_mat": bad})
    with pytest.raises(_lib.SpectralRenderError, match="material id"):
        Renderer(broken, cfg.bands, cfg.wl_min, cfg.wl_max, device=gpu, build="device")
