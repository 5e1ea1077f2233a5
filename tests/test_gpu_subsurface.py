"""Random-walk subsurface scattering (sr_render flag SR_RENDER_SUBSURFACE).

The reference leaves it as a TODO (main.py:1066), so there is no reference
image to match; the walk is pinned by radiative-transfer closed forms:

* a semi-infinite, index-matched medium with isotropic scattering of single
  albedo w under uniform sky radiance L reflects L times its spherical
  albedo r_s(w) = 1 - 2 sqrt(1 - w) int_0^1 H(mu) mu dmu (Chandrasekhar's
  H-function; a cosine-distributed entry and a Lambertian exit make the
  outgoing radiance uniform), checked at w = 0.5 and 0.9 on a closed slab
  many mean free paths thick;
* a closed sphere with w = 1 under the same sky returns exactly L (white
  furnace: every walk leaves, nothing is absorbed);
* with no opaque scattering material the flag changes nothing;
* with the light tree as well (both sr_render flags), the image's energy is
  the one of the walk with the area-CDF pick.
The H-function solver itself is checked on the CPU against its identities
and Chandrasekhar's tabulated H(1) = 2.9078 for w = 1."""
import math

import numpy as np
import pytest

from spectral import configs
from spectral.render import Renderer
from spectral.scene import flatten_scene

B = 8


def chandrasekhar_h(w, n=96, iters=20000):
    """H(mu) for isotropic scattering by iterating
    H(mu) = 1 / (1 - (w/2) mu int_0^1 H(mu') / (mu + mu') dmu')
    on Gauss-Legendre nodes (slow but convergent for w = 1 too)."""
    x, wt = np.polynomial.legendre.leggauss(n)
    mu, wt = 0.5 * (x + 1.0), 0.5 * wt
    h = np.ones(n)
    for _ in range(iters):
        integ = ((wt * h)[None, :] / (mu[:, None] + mu[None, :])).sum(1)
        hn = 1.0 / (1.0 - 0.5 * w * mu * integ)
        if np.abs(hn - h).max() < 1e-13:
            h = hn
            break
        h = hn
    return mu, wt, h


def spherical_albedo(w):
    mu, wt, h = chandrasekhar_h(w)
    return 1.0 - 2.0 * math.sqrt(1.0 - w) * float((wt * mu * h).sum())


def test_h_function_solver():
    mu, wt, h = chandrasekhar_h(1.0, n=64, iters=20000)
    # conservative case: int_0^1 H dmu = 2 and H(1) = 2.9078 (tabulated)
    assert float((wt * h).sum()) == pytest.approx(2.0, abs=2e-3)
    # H at mu = 1, extrapolated from the last nodes
    h_at_1 = h[-1] + (h[-1] - h[-2]) / (mu[-1] - mu[-2]) * (1.0 - mu[-1])
    assert h_at_1 == pytest.approx(2.9078, abs=2e-3)
    assert spherical_albedo(1.0) == pytest.approx(1.0, abs=1e-6)
    assert spherical_albedo(1e-6) == pytest.approx(0.0, abs=1e-5)
    for w in (0.2, 0.5, 0.9, 0.99):
        mu, wt, h = chandrasekhar_h(w)
        # zeroth moment identity: int_0^1 H dmu = (2/w) (1 - sqrt(1 - w))
        assert float((wt * h).sum()) == pytest.approx(2.0 / w * (1.0 - math.sqrt(1.0 - w)), rel=1e-9)


def _scene(body, w, sigma_s):
    cfg = configs.subsurface_slab(w=w, sigma_s=sigma_s, bands=B, body=body)
    return flatten_scene(cfg.scene, cfg.bands, cfg.wl_min, cfg.wl_max)


def _box(lo, hi, mat, outward=True):
    return configs.closed_box(lo, hi, mat, outward)


def _mean_and_se(r, n_chunks=16, per=16, depth=5):
    vals = []
    for c in range(n_chunks):
        r.accum.zero_()
        r.blocks_done = c * per
        r.render(n_blocks=per, max_depth=depth, seed=4, subsurface=True)
        vals.append(float(r.accum.double().mean()) * (c + 1))
    vals = np.array(vals)
    return vals.mean(), vals.std(ddof=1) / math.sqrt(n_chunks)


@pytest.mark.gpu
@pytest.mark.parametrize("w", [0.5, 0.9])
def test_semi_infinite_slab_spherical_albedo(gpu, w):
    # sigma_t <= 10: the slab is >= 25 mean free paths deep, and scattering
    # within the trace's 1e-4 minimum hit distance of the boundary (where the
    # walk cannot see it) stays below 1e-3 of the events
    flat = _scene(_box([-10, -5, -10], [10, 0, 10], 1), w, sigma_s=5.0)
    r = Renderer(flat, B, 400.0, 700.0, width=16, height=16, device=gpu)
    m, se = _mean_and_se(r)
    rs = spherical_albedo(w)
    print(f"w={w}: rendered {m:.5f} +- {se:.5f}, r_s = {rs:.5f}")
    assert abs(m - rs) < 4 * se + 1e-3


@pytest.mark.gpu
def test_white_furnace_sphere(gpu):
    body = configs.icosphere([0.0, 0.0, 0.0], 0.5, 1, subdivisions=3)
    flat = _scene(body, 1.0, sigma_s=5.0)
    r = Renderer(flat, B, 400.0, 700.0, width=16, height=16, device=gpu)
    m, se = _mean_and_se(r)
    assert abs(m - 1.0) < 4 * se + 2e-3, (m, se)


@pytest.mark.gpu
def test_flag_is_a_no_op_without_scattering_opaque_materials(gpu):
    cfg = configs.box_icosphere()
    for m in (1, 2, 3):  # walls: no scattering coefficient
        cfg.scene["materials"][m]["scattering_coefficient"] = [0.0] * cfg.bands
    flat = flatten_scene(cfg.scene, cfg.bands, cfg.wl_min, cfg.wl_max)
    r = Renderer(flat, cfg.bands, cfg.wl_min, cfg.wl_max, width=16, height=16, device=gpu)
    imgs = []
    for sss in (False, True):
        r.clear_pixel_space()
        r.render(n_blocks=8, max_depth=cfg.depth, seed=2, subsurface=sss)
        imgs.append(r.accum.cpu().numpy().astype(np.float64))
    assert np.allclose(imgs[0], imgs[1], rtol=1e-5, atol=1e-7)


@pytest.mark.gpu
def test_walk_with_light_tree_same_expectation(gpu):
    # the sky box's 12 faces are the lights; the walk's exit NEE picks them
    # through the tree in one mode and by the area CDF in the other
    flat = _scene(_box([-10, -5, -10], [10, 0, 10], 1), 0.9, sigma_s=5.0)
    r = Renderer(flat, B, 400.0, 700.0, width=16, height=16, device=gpu)
    tot = {}
    for tree in (False, True):
        vals = []
        for c in range(32):
            r.accum.zero_()
            r.blocks_done = c * 16
            r.render(n_blocks=16, max_depth=5, seed=4, stream=1 if tree else 0, subsurface=True, light_tree=tree)
            vals.append(float(r.accum.double().sum()) * (c + 1))
        tot[tree] = np.array(vals)
    se = math.sqrt(tot[False].var(ddof=1) / 32 + tot[True].var(ddof=1) / 32)
    m0, m1 = tot[False].mean(), tot[True].mean()
    print(f"tree / CDF with the walk: {m1 / m0:.5f} +- {se / m0:.5f}")
    assert abs(m1 - m0) < 4 * se + 5e-4 * m0


@pytest.mark.gpu
def test_walks_end_at_the_step_limit(gpu):
    # w = 1 and sigma_s = 1000 (mean free path 1e-3) in a 5-unit slab: a
    # walk enters 0.1 mean free paths deep, so most leave within a few
    # steps, but the long tail passes 256 steps and is ended there; the
    # render finishes, stays finite, and returns visibly less than the sky
    # radiance 1 that unbounded walks would (measured 0.868)
    flat = _scene(_box([-10, -5, -10], [10, 0, 10], 1), 1.0, sigma_s=1000.0)
    r = Renderer(flat, B, 400.0, 700.0, width=16, height=16, device=gpu)
    r.clear_pixel_space()
    r.render(n_blocks=4, max_depth=5, seed=4, subsurface=True)
    a = r.accum.cpu().numpy()
    assert np.isfinite(a).all() and (a >= 0).all()
    st = r.stats()
    assert st["nonfinite"] == 0
    assert 0.6 < float(a.mean()) < 0.95
