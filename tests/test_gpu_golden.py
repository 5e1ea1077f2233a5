"""The HIP path against the committed golden fixtures (tests/golden/), for
the 1,292-face box scene and the 1,048,588-face heightfield scene:
sr_trace_closest on the frozen ray batches (bit-exact t, face id, u, v,
host- and device-compiled scene), and sr_render at 32x32 x 4096 blocks
against the frozen oracle images: with the oracle's seed, at least 99.7% of
pixel-bands within 3 standard errors and mean relative error below 1%; with
an independent Philox stream, the z-score distribution of the difference."""
import os

import numpy as np
import pytest

from spectral import configs
from spectral.render import Renderer
from spectral.scene import flatten_scene

pytestmark = pytest.mark.gpu

G = os.path.join(os.path.dirname(__file__), "golden")


def _load(name):
    return np.load(os.path.join(G, name), allow_pickle=False)


@pytest.fixture(scope="module")
def box():
    cfg = configs.box_icosphere()
    return cfg, flatten_scene(cfg.scene, cfg.bands, cfg.wl_min, cfg.wl_max)


@pytest.mark.parametrize("build", ["host", "device"])
def test_golden_rays(box, gpu, build):
    cfg, flat = box
    g = _load("c0_rays.npz")
    r = Renderer(flat, cfg.bands, cfg.wl_min, cfg.wl_max, device=gpu, build=build)
    t, p, uv = r.trace(g["orig"], g["dirs"])
    assert np.array_equal(p, g["prim"])
    assert np.array_equal(t, g["t"])
    hit = g["prim"] >= 0
    assert np.array_equal(uv[hit], g["uv"][hit])


def _render(flat, cfg, gpu, g, stream):
    W, H = g["mean"].shape[:2]
    r = Renderer(flat, cfg.bands, cfg.wl_min, cfg.wl_max, width=W, height=H, device=gpu)
    r.render(n_blocks=int(g["chunks"]) * int(g["per_chunk"]), max_depth=int(g["depth"]), seed=int(g["seed"]),
             stream=stream)
    return r.accum.cpu().numpy().astype(np.float64)


def test_golden_image_matched_seed_4096_blocks(box, gpu):
    cfg, flat = box
    g = _load("c0_image.npz")
    a = _render(flat, cfg, gpu, g, 0)
    mean, sem = g["mean"].astype(np.float64), g["sem"].astype(np.float64)
    diff = np.abs(a - mean)
    ok = diff <= 3 * sem + 1e-6 * np.abs(mean) + 1e-7
    assert ok.mean() >= 0.997, f"{(~ok).sum()} of {ok.size} pixel-bands beyond 3 sigma"
    lit = mean > 1e-3
    mre = (diff[lit] / mean[lit]).mean()
    assert mre < 0.01, f"mean relative error {mre:.4f}"


def test_golden_image_independent_stream(box, gpu):
    cfg, flat = box
    g = _load("c0_image.npz")
    a = _render(flat, cfg, gpu, g, 1)
    sem = g["sem"].astype(np.float64)
    s = np.sqrt(2.0) * sem  # both sides are 4096-block means of the same estimator
    m = s > 0
    assert np.array_equal(a[~m], g["mean"][~m])
    z = (a - g["mean"])[m] / s[m]
    # heavy-tailed Monte Carlo radiance: robust statistics of z
    q50, q90 = np.quantile(np.abs(z), [0.5, 0.9])
    assert 0.5 < q50 < 0.85 and 1.3 < q90 < 2.1, (q50, q90)
    assert (np.abs(z) > 4).mean() < 0.015
    zs = np.sort(z)
    k = len(zs) // 100
    assert abs(zs[k:-k].mean()) < 0.1
    lit = g["mean"] > 1e-3
    assert abs(a[lit].sum() / g["mean"][lit].sum() - 1.0) < 0.005


# ---------------------------------------------------------------- heightfield_1m


@pytest.fixture(scope="module")
def hf():
    cfg = configs.heightfield_1m()
    return cfg, flatten_scene(cfg.scene, cfg.bands, cfg.wl_min, cfg.wl_max)


@pytest.mark.parametrize("build", ["host", "device"])
def test_c1_golden_rays(hf, gpu, build):
    """256x256 pixel-centre rays and 65,536 seeded rays against brute force
    over all 1,048,588 faces (trace, main.py:831-864)."""
    from spectral.scene import pixel_center_rays

    cfg, flat = hf
    g = _load("c1_rays.npz")
    n = int(g["pixel_grid"])
    po, pd = pixel_center_rays(flat, n, n)
    o = np.concatenate([po, g["orig"]]).astype(np.float32)
    d = np.concatenate([pd, g["dirs"]]).astype(np.float32)
    r = Renderer(flat, cfg.bands, cfg.wl_min, cfg.wl_max, device=gpu, build=build)
    t, p, uv = r.trace(o, d)
    assert np.array_equal(p, g["prim"]), f"{(p != g['prim']).sum()} face ids differ"
    assert np.array_equal(t, g["t"])
    hit = g["prim"] >= 0
    assert np.array_equal(uv[hit], g["uv"][hit])


def test_c1_golden_image_matched_seed_4096_blocks(hf, gpu):
    cfg, flat = hf
    g = _load("c1_image.npz")
    a = _render(flat, cfg, gpu, g, 0)
    mean, sem = g["mean"].astype(np.float64), g["sem"].astype(np.float64)
    diff = np.abs(a - mean)
    ok = diff <= 3 * sem + 1e-6 * np.abs(mean) + 1e-7
    lit = mean > 1e-3
    mre = (diff[lit] / mean[lit]).mean()
    print(f"heightfield 32x32 x 4096 blocks: {ok.mean():.5f} of pixel-bands within 3 sigma, MRE {mre:.2e}")
    assert ok.mean() >= 0.997, f"{(~ok).sum()} of {ok.size} pixel-bands beyond 3 sigma"
    assert mre < 0.01, f"mean relative error {mre:.4f}"


def test_c1_golden_image_independent_stream(hf, gpu):
    cfg, flat = hf
    g = _load("c1_image.npz")
    a = _render(flat, cfg, gpu, g, 1)
    s = np.sqrt(2.0) * g["sem"].astype(np.float64)
    m = s > 0
    assert np.array_equal(a[~m], g["mean"][~m])
    z = (a - g["mean"])[m] / s[m]
    q50, q90 = np.quantile(np.abs(z), [0.5, 0.9])
    print(f"heightfield independent stream: |z| median {q50:.3f}, 90% {q90:.3f}")
    assert 0.5 < q50 < 0.85 and 1.3 < q90 < 2.1, (q50, q90)
    assert (np.abs(z) > 4).mean() < 0.015
    lit = g["mean"] > 1e-3
    assert abs(a[lit].sum() / g["mean"][lit].sum() - 1.0) < 0.005
