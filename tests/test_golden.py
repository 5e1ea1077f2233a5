"""The oracle against the committed golden fixtures (tests/golden/, written by
make_golden.py): brute-force closest hits of a frozen ray batch, the BSDF
closed-form tables, and the frozen 4096-block C0 image."""
import os

import numpy as np
import pytest

import oracle.oracle as O
from oracle.oracle import OracleScene
from spectral import configs
from spectral.scene import flatten_scene

G = os.path.join(os.path.dirname(__file__), "golden")


def _load(name):
    return np.load(os.path.join(G, name), allow_pickle=False)


def _digest(flat):
    import hashlib

    h = hashlib.sha256()
    for a in (flat.tris_v, flat.tris_n, flat.tri_mat, flat.mat_f32, flat.mat_i32, flat.mat_eem, flat.eem, flat.cam):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


@pytest.fixture(scope="module")
def box():
    cfg = configs.box_icosphere()
    flat = flatten_scene(cfg.scene, cfg.bands, cfg.wl_min, cfg.wl_max)
    return cfg, flat


def test_golden_scene_is_current(box):
    _, flat = box
    for name in ("c0_rays.npz", "c0_image.npz"):
        assert str(_load(name)["scene_sha256"]) == _digest(flat), f"{name} was made from another box scene"


def test_rays_brute_force(box):
    _, flat = box
    g = _load("c0_rays.npz")
    t, p, uv = OracleScene(flat, brute=True).trace(g["orig"], g["dirs"])
    assert np.array_equal(p, g["prim"])
    assert np.array_equal(t, g["t"])
    hit = g["prim"] >= 0
    assert np.array_equal(uv[hit], g["uv"][hit])
    # the box is open towards the camera: some rays leave it
    assert hit[: int(g["n_pixel_rays"])].mean() > 0.5 and hit.mean() > 0.75


def test_rays_reference_shaped_bvh_same_t(box):
    """The oracle's reference-shaped BVH finds the same closest distance
    (face ids may differ only on exact ties in t)."""
    _, flat = box
    g = _load("c0_rays.npz")
    t, p, _ = OracleScene(flat).trace(g["orig"], g["dirs"])
    assert np.array_equal(t, g["t"])
    assert (p == g["prim"]).mean() > 0.99


def test_bsdf_tables():
    g = _load("bsdf_tables.npz")
    for (ni, ki, nt, kt), row in zip(g["fresnel_media"], g["fresnel"]):
        got = np.array([O.fresnel(float(c), ni, ki, nt, kt) for c in g["fresnel_cos"]])
        assert np.allclose(got, row, rtol=1e-4, atol=1e-6), (ni, ki, nt, kt)
    for r, row in zip(g["ggx_roughness"], g["ggx_D"]):
        got = np.array([O.ggx_D(float(c), float(r)) for c in g["ggx_nh"]])
        # f32 cancellation in nh^2 (a2 - 1) + 1: relative error ~ 2 ulp / den, squared
        a2 = float(r) ** 4
        den = g["ggx_nh"] ** 2 * (a2 - 1.0) + 1.0
        tol = 2e-5 + 2 * 2.4e-7 / den
        assert (np.abs(got - row) <= tol * row + 1e-7).all(), r
    for gg, row in zip(g["hg_g"], g["hg"]):
        got = np.array([O.hg_eval(float(c), float(gg)) for c in g["hg_cos"]])
        # f32 cancellation in 1 + g^2 - 2 g c, raised to the power 3/2
        den = 1.0 + gg * gg - 2.0 * gg * g["hg_cos"]
        tol = 2e-5 + 1.5 * 3 * 1.2e-7 * (1.0 + gg * gg + 2.0 * abs(gg)) / den
        assert (np.abs(got - row) <= tol * row + 1e-7).all(), gg


def test_c0_image_independent_stream(box):
    """64 blocks from an independent Philox stream against the frozen
    4096-block image: z scores with the per-block spread of the fixture."""
    cfg, flat = box
    g = _load("c0_image.npz")
    W, H = g["mean"].shape[:2]
    n = 64
    a, _ = OracleScene(flat).render(W, H, int(g["seed"]), 0, n, int(g["depth"]), stream=1)
    sem = g["sem"].astype(np.float64)
    s = np.sqrt(sem ** 2 + sem ** 2 * (int(g["chunks"]) * int(g["per_chunk"]) / n))
    m = s > 0
    assert np.array_equal(a[~m], g["mean"][~m])
    z = (a.astype(np.float64) - g["mean"])[m] / s[m]
    # Monte Carlo radiance is heavy-tailed (rare bright paths): robust checks.
    # A bias of a few percent on lit pixels moves the median |z| far above
    # the normal 0.674.
    q50, q90 = np.quantile(np.abs(z), [0.5, 0.9])
    assert 0.55 < q50 < 0.8 and 1.4 < q90 < 2.0, (q50, q90)
    assert (np.abs(z) > 4).mean() < 0.015
    zs = np.sort(z)
    k = len(zs) // 100
    assert abs(zs[k:-k].mean()) < 0.1


# ---------------------------------------------------------------- heightfield_1m (1,048,588 faces)

@pytest.fixture(scope="module")
def hf():
    cfg = configs.heightfield_1m()
    return cfg, flatten_scene(cfg.scene, cfg.bands, cfg.wl_min, cfg.wl_max)


def c1_all_rays(flat, g):
    from spectral.scene import pixel_center_rays

    n = int(g["pixel_grid"])
    po, pd = pixel_center_rays(flat, n, n)
    return np.concatenate([po, g["orig"]]).astype(np.float32), np.concatenate([pd, g["dirs"]]).astype(np.float32)


def test_c1_fixtures_are_current(hf):
    import hashlib

    _, flat = hf
    g = _load("c1_rays.npz")
    for name in ("c1_rays.npz", "c1_image.npz"):
        assert str(_load(name)["scene_sha256"]) == _digest(flat), f"{name} was made from another scene"
    o, d = c1_all_rays(flat, g)
    n = int(g["pixel_grid"]) ** 2
    h = hashlib.sha256()
    h.update(o[:n].tobytes())
    h.update(d[:n].tobytes())
    assert h.hexdigest() == str(g["pixel_rays_sha256"])
    assert g["t"].shape == (o.shape[0],)


def test_c1_rays_brute_force_sample(hf):
    """Brute force over all 1,048,588 faces on a sample of the batch (the
    whole batch took minutes to freeze): bit-exact."""
    _, flat = hf
    g = _load("c1_rays.npz")
    o, d = c1_all_rays(flat, g)
    idx = np.random.default_rng(5).choice(o.shape[0], 384, replace=False)
    t, p, uv = OracleScene(flat, brute=True).trace(o[idx], d[idx])
    assert np.array_equal(p, g["prim"][idx])
    assert np.array_equal(t, g["t"][idx])
    hit = p >= 0
    assert np.array_equal(uv[hit], g["uv"][idx][hit])


def test_c1_rays_reference_shaped_bvh(hf):
    """The oracle's reference-shaped BVH on the whole batch: the same t
    everywhere, the same face except on exact ties in t; the batch hits the
    heightfield (material 5) a lot."""
    _, flat = hf
    g = _load("c1_rays.npz")
    o, d = c1_all_rays(flat, g)
    t, p, _ = OracleScene(flat).trace(o, d)
    assert np.array_equal(t, g["t"])
    # a ray through a shared edge of two grid triangles hits both at the
    # same t: traversal order then picks the face (brute force: the lowest id)
    assert (p == g["prim"]).mean() > 0.995
    assert np.array_equal(p >= 0, g["prim"] >= 0)
    hit = g["prim"] >= 0
    assert hit.mean() > 0.7
    assert (flat.tri_mat[g["prim"][hit]] == 5).mean() > 0.4


def test_c1_image_independent_stream(hf):
    """16 blocks from an independent Philox stream against the frozen
    4096-block heightfield image (robust z statistics, as for the box)."""
    cfg, flat = hf
    g = _load("c1_image.npz")
    W, H = g["mean"].shape[:2]
    n = 16
    a, _ = OracleScene(flat).render(W, H, int(g["seed"]), 0, n, int(g["depth"]), stream=1)
    sem = g["sem"].astype(np.float64)
    s = np.sqrt(sem ** 2 + sem ** 2 * (int(g["chunks"]) * int(g["per_chunk"]) / n))
    m = s > 0
    assert np.array_equal(a[~m], g["mean"][~m])
    z = (a.astype(np.float64) - g["mean"])[m] / s[m]
    q50 = np.quantile(np.abs(z), 0.5)
    assert 0.4 < q50 < 0.85, q50
    assert (np.abs(z) > 5).mean() < 0.02
    lit = g["mean"] > 1e-3
    assert abs(a[lit].sum() / g["mean"][lit].sum() - 1.0) < 0.03
