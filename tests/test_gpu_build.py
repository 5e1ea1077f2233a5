"""Scene compile on the device (sr_upload_scene_gpu) against the host build.

The device builder reproduces the host builder's tree, so the test is
byte identity of every compiled array (nodes, triangle records, shading
records, light tables) read back through sr_scene_array, on the 1292-face
box scene, the 1,048,588- and 4,194,316-face heightfields, 16 of the 64
batch scenes, and
scenes made to reach the builder's other paths: coincident centroids (every
split an object median with ties, in both the chunked and the one-wave
kernels), faces at x = -2^-k (SAH splits that peel off a few faces until
the depth budget forces median splits, on large ranges too), and
leaf-sized and empty scenes.  The host tree itself is pinned against brute force
(bit-exact hits) here and in the trace tests."""
import numpy as np
import pytest

from oracle.oracle import OracleScene
from spectral import _lib, configs
from spectral.render import Renderer
from spectral.scene import flatten_scene, pixel_center_rays

pytestmark = pytest.mark.gpu

ARRAYS = ("nodes", "tris", "shade", "lights", "lverts", "light_count", "light_area", "light_tree", "light_power",
          "light_keys")


def _random_rays(n, seed, lo=-0.49, hi=0.49):
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return o, d.astype(np.float32)


def _pair(flat, gpu, bands, wl_min, wl_max, **kw):
    host = Renderer(flat, bands, wl_min, wl_max, device=gpu, **kw)
    dev = Renderer(flat, bands, wl_min, wl_max, device=gpu, build="device", **kw)
    return host, dev


def _assert_same_scene(host, dev):
    assert host.scene_info() == dev.scene_info()
    for name in ARRAYS:
        a, b = host.scene_array(name), dev.scene_array(name)
        assert a.shape == b.shape, name
        if not np.array_equal(a, b):
            bad = np.flatnonzero(a != b)
            pytest.fail(f"{name}: {bad.size} bytes differ, first at byte {bad[0]} of {a.size}")


def _with_tris(flat, v, mat=0):
    v = np.ascontiguousarray(v, np.float32)
    e1, e2 = v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]
    nrm = np.cross(e1, e2)
    nrm /= np.maximum(np.linalg.norm(nrm, axis=1, keepdims=True), 1e-20)
    n = np.repeat(nrm[:, None, :], 3, axis=1).astype(np.float32)
    return flat.__class__(**{**flat.__dict__, "tris_v": v, "tris_n": n,
                             "tri_mat": np.full(v.shape[0], mat, np.int32)})


@pytest.fixture(scope="module")
def box_flat():
    cfg = configs.box_icosphere()
    return cfg, flatten_scene(cfg.scene, cfg.bands, cfg.wl_min, cfg.wl_max)


@pytest.fixture(scope="module")
def box(gpu, box_flat):
    cfg, flat = box_flat
    return (flat,) + _pair(flat, gpu, cfg.bands, cfg.wl_min, cfg.wl_max)


def test_box_scene_identical_to_host_build(box):
    _, host, dev = box
    _assert_same_scene(host, dev)


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


def test_box_render_matches_host_build(box):
    """Same records, same random numbers: the two builds render the same
    paths; only the order of the float atomics differs."""
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


def test_heightfield_1m_identical_to_host_build(gpu):
    cfg = configs.heightfield_1m()
    flat = flatten_scene(cfg.scene, cfg.bands, cfg.wl_min, cfg.wl_max)
    host, dev = _pair(flat, gpu, cfg.bands, cfg.wl_min, cfg.wl_max)
    assert dev.scene_info()["bvh4_depth"] <= 13
    _assert_same_scene(host, dev)


def test_heightfield_4m_identical_to_host_build(gpu):
    """4,194,316 faces (2048 x 1024 cells): the level kernels at four times
    the headline size, a tree at the 4-wide depth bound."""
    cfg = configs.heightfield_1m(nx=2048, nz=1024, width=64, height=64)
    flat = flatten_scene(cfg.scene, cfg.bands, cfg.wl_min, cfg.wl_max)
    host, dev = _pair(flat, gpu, cfg.bands, cfg.wl_min, cfg.wl_max)
    assert dev.scene_info()["bvh4_depth"] <= 13
    _assert_same_scene(host, dev)


def test_batch_scenes_identical_to_host_build(gpu):
    cfgs = configs.batch64()
    for i in range(0, 64, 4):
        c = cfgs[i]
        flat = flatten_scene(c.scene, c.bands, c.wl_min, c.wl_max)
        host, dev = _pair(flat, gpu, c.bands, c.wl_min, c.wl_max)
        print(f"batch scene {i}: {flat.n_tris} faces", flush=True)
        _assert_same_scene(host, dev)
        del host, dev


def test_coincident_centroids_median_splits(gpu, box_flat):
    """Every face has its centroid exactly at the origin (vertices p, q,
    -(p + q) on a 2^-9 grid, so the f32 sums are exact): no SAH split
    exists, every split is an object median whose ties keep the current
    order."""
    cfg, flat = box_flat
    rng = np.random.default_rng(11)
    for n in (3000, 50):  # chunked level kernels, one-wave subtree
        p = rng.integers(-200, 201, (n, 3)).astype(np.float32) / np.float32(512)
        q = rng.integers(-200, 201, (n, 3)).astype(np.float32) / np.float32(512)
        v = np.stack([p, q, -(p + q)], axis=1)
        c = (v[:, 0] + v[:, 1] + v[:, 2]) * np.float32(1.0 / 3.0)
        assert not c.any()
        f = _with_tris(flat, v)
        host, dev = _pair(f, gpu, cfg.bands, cfg.wl_min, cfg.wl_max)
        _assert_same_scene(host, dev)
        o, d = _random_rays(4000, n, -0.6, 0.6)
        t0, p0, _ = OracleScene(f, brute=True).trace(o, d)
        t, pp, _ = dev.trace(o, d)
        assert np.array_equal(pp, p0) and np.array_equal(t, t0)


def test_depth_budget_forces_median_splits(gpu, box_flat):
    """16 faces at each x = -2^-k, k < 125: the 16 bins separate about four
    positions per split, so the SAH tree runs into the binary depth budget
    and median splits are forced, on ranges of more than 64 faces too."""
    cfg, flat = box_flat
    k = np.repeat(np.arange(125), 16)
    n = k.size
    x = -(np.float32(2.0) ** -k.astype(np.float32))
    s = np.float32(1e-6) * (1 + (np.arange(n) % 16)).astype(np.float32)
    v = np.zeros((n, 3, 3), np.float32)
    v[:, :, 0] = x[:, None]
    v[:, 1, 0] += s
    v[:, 2, 0] -= s
    v[:, 1, 1] = s
    v[:, 2, 2] = s
    f = _with_tris(flat, v)
    host, dev = _pair(f, gpu, cfg.bands, cfg.wl_min, cfg.wl_max)
    assert dev.scene_info()["bvh4_depth"] <= 13
    _assert_same_scene(host, dev)


def test_single_triangle_small_and_empty_scenes(gpu, box_flat):
    cfg, flat = box_flat
    for n in (1, 3, 4, 5, 63, 64, 65, 200):
        k = min(n, flat.n_tris)
        one = flat.__class__(**{**flat.__dict__, "tris_v": flat.tris_v[:k].copy(),
                                "tris_n": flat.tris_n[:k].copy(), "tri_mat": flat.tri_mat[:k].copy()})
        host, dev = _pair(one, gpu, cfg.bands, cfg.wl_min, cfg.wl_max)
        _assert_same_scene(host, dev)
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
    host, dev = _pair(empty, gpu, cfg.bands, cfg.wl_min, cfg.wl_max, width=8, height=8)
    _assert_same_scene(host, dev)
    dev.render(n_blocks=1, max_depth=cfg.depth, seed=0)
    assert float(dev.accum.abs().sum()) == 0.0


def test_bad_material_id_rejected(gpu, box_flat):
    cfg, flat = box_flat
    bad = flat.tri_mat.copy()
    bad[7] = flat.n_mats + 3
    broken = flat.__class__(**{**flat.__dict__, "tri_mat": bad})
    with pytest.raises(_lib.SpectralRenderError, match="material id"):
        Renderer(broken, cfg.bands, cfg.wl_min, cfg.wl_max, device=gpu, build="device")


def test_scene_array_errors(box):
    """sr_scene_array: a size query needs no buffer; a buffer smaller than
    the array, an unknown array id and a handle without a scene are errors."""
    import ctypes

    _, host, _ = box
    size = ctypes.c_size_t()
    _lib.call("sr_scene_array", host._h, 0, None, 0, ctypes.byref(size))
    assert size.value == host.scene_info()["bvh_nodes"] * 128
    small = np.empty(size.value - 16, np.uint8)
    with pytest.raises(_lib.SpectralRenderError, match="capacity"):
        _lib.call("sr_scene_array", host._h, 0, small.ctypes.data_as(ctypes.c_void_p), small.nbytes,
                  ctypes.byref(size))
    with pytest.raises(_lib.SpectralRenderError, match="unknown scene array"):
        _lib.call("sr_scene_array", host._h, 10, None, 0, ctypes.byref(size))
    h = ctypes.c_void_p()
    _lib.call("sr_create", 8, 400.0, 700.0, ctypes.byref(h))
    try:
        with pytest.raises(_lib.SpectralRenderError, match="ESTATE"):
            _lib.call("sr_scene_array", h, 0, None, 0, ctypes.byref(size))
    finally:
        _lib.call("sr_destroy", h)
