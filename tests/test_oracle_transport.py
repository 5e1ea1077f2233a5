"""Pins the CPU restatement's transport functions (scatter_volume,
shadow_transmittance, sample_direct_light, cast_rays) with hand-computed
cases and closed forms on small custom scenes."""
import math

import numpy as np
import pytest
from scipy import stats

from oracle import oracle as O
from spectral import configs
from spectral.scene import flatten_scene

B = 4
WL = (400.0, 700.0)


def mat(**kw):
    return configs._mat(B, **kw)


def quad(v, m):
    return {"vertices": v, "faces": [[0, 2, 1], [0, 3, 2]], "material_id": m}


def flat_of(meshes, materials, cam=None, eems=None):
    scene = {"meshes": meshes, "materials": materials,
             "camera": cam or {"origin": [0, 0, 0], "lookat": [0, 1, 0], "up": [0, 0, 1]}}
    if eems is not None:
        scene["eems"] = eems
    return flatten_scene(scene, B, *WL)


def lam0_m():
    return WL[0] * 1e-9


# ---------------------------------------------------------------- shadow walk
def _slab_scene(sigma_s=0.1, sigma_a=0.3):
    k = sigma_a * lam0_m() / (4 * math.pi)
    mats = [mat(n=1.0, tv=1), mat(n=1.0, k=k, tv=1, ss=sigma_s), mat(se=1, em=5.0), mat(n=1.4, tv=0, ss=1.0)]
    sl = configs.slab([-1, 1.0, -1], [1, 1.5, 1], 1, 1)
    light = quad([[-0.5, 3, -0.3], [0.5, 3, -0.3], [0.5, 3, 0.7], [-0.5, 3, 0.7]], 2)
    return mats, sl, light


def test_shadow_unoccluded_air_is_one():
    mats, sl, light = _slab_scene()
    flat = flat_of([light], mats)
    ora = O.OracleScene(flat)
    assert ora.shadow([0, 0, 0], [0, 1, 0], 3.0, 0, 0, 0) == 1.0


def test_shadow_inside_medium_double_attenuation():
    # main.py:1463 then 1468: the last span is attenuated twice.
    mats, sl, light = _slab_scene()
    flat = flat_of([light], mats)
    ora = O.OracleScene(flat)
    got = ora.shadow([0, 0, 0], [0, 1, 0], 3.0, 0, 1, 0)
    assert got == pytest.approx(math.exp(-0.4 * 3.0) ** 2, rel=1e-5)


def test_shadow_walk_restarts_from_origin():
    # An index-matched slab y in [1, 1.5] between the origin and the light at
    # y = 3.  Each step restarts from the ORIGINAL point plus the last hit
    # distance (main.py:1451, 1496), so the walk re-enters the slab:
    # spans in the medium: 0.4999, 0.4999, 0.5, then 0.4998 twice-counted end.
    mats, sl, light = _slab_scene()
    flat = flat_of([light, sl], mats)
    ora = O.OracleScene(flat)
    got = ora.shadow([0, 0, 0], [0, 1, 0], 3.0, 0, 0, 0)
    e = 1e-4
    spans = (0.5 - e) + (0.5 - e) + 0.5 + (0.5 - 2 * e)
    assert got == pytest.approx(math.exp(-0.4 * spans), rel=2e-4)


def test_shadow_opaque_blocks():
    mats, sl, light = _slab_scene()
    wall = quad([[-1, 2, -1], [-1, 2, 1], [1, 2, 1], [1, 2, -1]], 3)
    flat = flat_of([light, wall], mats)
    ora = O.OracleScene(flat)
    assert ora.shadow([0, 0, 0], [0, 1, 0], 3.0, 0, 0, 0) == 0.0


# ---------------------------------------------------------------- NEE
def _corner_ff(X, Y):
    """Form factor from a differential area to a parallel rectangle with one
    corner above it (Howell's catalogue, case B-3)."""
    a = X / math.sqrt(1 + X * X) * math.atan(Y / math.sqrt(1 + X * X))
    b = Y / math.sqrt(1 + Y * Y) * math.atan(X / math.sqrt(1 + Y * Y))
    return (a + b) / (2 * math.pi)


def test_nee_parallel_rectangle_irradiance():
    # Lambertian point (n = 1, k = 0: F = 0; roughness 0: Oren-Nayar A = 1,
    # B = 0) below the centre of a square light: E[factor] = albedo * FF.
    a, h = 0.3, 0.5
    mats = [mat(n=1.0, tv=1), mat(n=1.0, r=0.0, tv=0, ss=1.0), mat(se=1, em=5.0)]
    # ceiling-light winding (main.py:192): normal -y, facing the point
    light = quad([[-a / 2, h, -a / 2], [-a / 2, h, a / 2], [a / 2, h, a / 2], [a / 2, h, -a / 2]], 2)
    flat = flat_of([light], mats)
    ora = O.OracleScene(flat)
    out = ora.nee([0, 0, 0], [0, -1, 0], [0, 1, 0], 1, 0, 0, seed=5, n=400000)
    albedo = 1.0 / (1.0 + 1e-6)
    ff = 4 * _corner_ff(a / (2 * h), a / (2 * h))
    assert out[2] == pytest.approx(albedo * ff, rel=5e-3)
    assert out[1] == pytest.approx(5.0)


# ---------------------------------------------------------------- volume events
def _medium_scene(sigma_s, sigma_a, g=0.0, eem=None):
    k = sigma_a * lam0_m() / (4 * math.pi)
    mats = [mat(n=1.0, tv=1), mat(n=1.0, k=k, tv=1, ss=sigma_s, g=g, eem_id=1 if eem is not None else 0)]
    sphere = configs.icosphere([0, 0, 0], 5.0, 1, subdivisions=2)
    eems = {1: eem} if eem is not None else None
    return flat_of([sphere], mats, eems=eems)


def test_free_flight_exponential_and_event_split():
    ss, sa = 1.2, 0.8
    flat = _medium_scene(ss, sa)
    ora = O.OracleScene(flat)
    tr, ac, es, nb = ora.volume_events([0, 0, 0], [0.3, 0.4, math.sqrt(0.75)], 1, 0, seed=9, n=200000)
    st = ss + sa + 1e-6
    vol = ac != 1
    assert vol.mean() > 0.999
    # free-flight distances: exponential with rate sigma_t (main.py:1204)
    assert stats.kstest(tr[vol], "expon", args=(0, 1 / st)).pvalue > 1e-3
    # absorption-or-null branch with probability sigma_a / sigma_t (main.py:1245-1246)
    p_abs = (ac[vol] == 0).mean()
    assert p_abs == pytest.approx(sa / st, abs=4 * math.sqrt(0.25 / vol.sum()))
    # estimators: T on scatter, T / (1 + 1e-6) on a null event (main.py:1282-1290)
    T = np.exp(-st * tr.astype(np.float64))
    sc = ac == 2
    assert np.allclose(es[sc], T[sc], rtol=1e-5)
    nu = ac == 0
    assert np.allclose(es[nu], T[nu] / (1 + 1e-6), rtol=1e-5)


def test_volume_fluorescence_band_pmf():
    # Volume fluorescence (main.py:1248-1278): QY = row max is the
    # probability of re-emission after an absorption event; the new band
    # follows the row-normalised EEM.
    eem = np.zeros((B, B))
    eem[0] = [0.0, 0.2, 0.5, 0.3]
    flat = _medium_scene(0.0, 2.0, eem=eem)
    ora = O.OracleScene(flat)
    tr, ac, es, nb = ora.volume_events([0, 0, 0], [0, 0, 1], 1, 0, seed=10, n=200000)
    vol = ac != 1
    fl = ac == 3
    assert fl.sum() / vol.sum() == pytest.approx(0.5, abs=0.005)
    counts = np.bincount(nb[fl], minlength=B)
    pmf = eem[0] / eem[0].sum()
    assert stats.chisquare(counts[1:], pmf[1:] * fl.sum()).pvalue > 1e-3
    # estimator (1 - T) + T * emission / (sigma_a + 1e-6) with zero emission
    T = np.exp(-(2.0 + 1e-6) * tr[fl].astype(np.float64))
    assert np.allclose(es[fl], 1 - T, rtol=1e-4, atol=1e-6)


# ---------------------------------------------------------------- cast_rays
def _light_only_scene():
    mats = [mat(n=1.0, tv=1), mat(se=1, em=10.0)]
    light = quad([[-0.3, 1, -0.3], [-0.3, 1, 0.3], [0.3, 1, 0.3], [0.3, 1, -0.3]], 1)
    return flat_of([light], mats, cam={"origin": [0, 0, 0], "lookat": [0, 1, 0], "up": [0, 0, 1]})


def test_light_visible_pixels_closed_form():
    # A camera ray that hits an emitter: throughput 1 / (1 + 1e-6) (main.py:1236)
    # and MIS weight sp^2 / (0 + sp^2 + 1e-6) with pdf_light * min(1, 0) = 0.
    flat = _light_only_scene()
    ora = O.OracleScene(flat)
    W = H = 16
    acc, st = ora.render(W, H, seed=1, first_block=0, n_blocks=3, depth=4)
    expect = np.float32(np.float32(np.float32(1 / (1 + 1e-6)) * 10.0) * np.float32(1 / (1 + 1e-6)))
    # the light spans |x|,|z| <= 0.3 at distance 1: the central 4x4 pixels are
    # fully covered for every jitter (image plane half-width 0.5 at distance 1)
    c = acc[6:10, 6:10]
    assert np.allclose(c, expect, rtol=2e-6)
    # pixels fully outside the light see only escaped paths: exactly 0
    assert (acc[0:2] == 0).all() and (acc[:, 0:2] == 0).all()
    assert st["nonfinite"] == 0


def test_zero_light_scene_is_black():
    cfg = configs.box_icosphere()
    scene = dict(cfg.scene)
    scene["materials"] = [dict(m) for m in scene["materials"]]
    scene["materials"][4]["surface_emission"] = [0] * cfg.bands
    scene["materials"][4]["emission"] = [0.0] * cfg.bands
    flat = flatten_scene(scene, cfg.bands, cfg.wl_min, cfg.wl_max)
    ora = O.OracleScene(flat)
    acc, st = ora.render(12, 12, seed=0, first_block=0, n_blocks=2, depth=6)
    assert (acc == 0).all()
    assert st["segments"] > 0


def test_running_mean_over_blocks():
    # average_and_reset_block applied per block (main.py:1603) equals the
    # mean of the block sums.
    cfg = configs.box_icosphere()
    flat = flatten_scene(cfg.scene, cfg.bands, cfg.wl_min, cfg.wl_max)
    ora = O.OracleScene(flat)
    one, _ = ora.render(8, 8, seed=4, first_block=0, n_blocks=4, depth=5)
    acc = np.zeros_like(one)
    per = []
    for b in range(4):
        blk, _ = ora.render(8, 8, seed=4, first_block=b, n_blocks=1, depth=5, accum=np.zeros_like(one))
        per.append(blk * (b + 1))  # alone, block b is folded in as blk / (b + 1)
        ora.render(8, 8, seed=4, first_block=b, n_blocks=1, depth=5, accum=acc)
    # blocks rendered alone reproduce the same paths (RNG keyed by block)
    assert np.allclose(acc, one, rtol=1e-5, atol=1e-7)
    assert np.allclose(np.mean(per, axis=0), one, rtol=1e-5, atol=1e-7)


# ---------------------------------------------------------- non-finite terms
def _nf_stats(cfg, w, n_blocks, roughness=None, mat_id=None, cols=None):
    scene = dict(cfg.scene)
    if roughness is not None:
        scene["materials"] = [dict(m) for m in scene["materials"]]
        scene["materials"][mat_id]["roughness"] = [roughness] * cfg.bands
    flat = flatten_scene(scene, cfg.bands, cfg.wl_min, cfg.wl_max)
    _, st = O.OracleScene(flat).render(w, w, cfg.seed, 0, n_blocks, cfg.depth, cols=cols)
    return st


# (config, render size, blocks, rough dielectric volume, non-finite events per
# path segment measured on the oracle at this size, +-10%)
NF_CASES = [("box_icosphere", 64, 4, 5, 0.0652), ("volume_fluor", 48, 2, 6, 0.0245)]


@pytest.mark.parametrize("name,w,nb,mat_id,rate", NF_CASES)
def test_nonfinite_source_is_rough_btdf_leaving_the_medium(name, w, nb, mat_id, rate):
    # Leaving a rough (roughness >= 0.0018) true-volume dielectric, the
    # reference takes n_i from the ray's volume and n_t from the hit surface,
    # the same material: refract() returns the incident direction, L.H =
    # -V.H, the denominator n_i V.H + n_t L.H is 0 and the BTDF estimator of
    # main.py:1176-1185 is 0/0.  The path's throughput becomes NaN; it is
    # counted (and the path ends) at the next segment, so an event on the
    # last segment is never counted.  Every other source stays at zero.
    cfg = getattr(configs, name)()
    st = _nf_stats(cfg, w, nb)
    assert st["nf_other_estimator"] == 0 and st["nf_direct"] == 0 and st["nf_indirect"] == 0
    assert st["nf_btdf_eta1"] > 0
    assert st["nonfinite"] == st["nf_btdf_eta1"] - st["nf_btdf_eta1_last"]
    got = st["nonfinite"] / st["segments"]
    assert abs(got - rate) <= 0.1 * rate, f"{got:.4f} non-finite events per segment"


@pytest.mark.parametrize("name,w,nb,mat_id,rate", NF_CASES)
def test_nonfinite_vanishes_for_a_smooth_dielectric(name, w, nb, mat_id, rate):
    # Below roughness 0.0018 the transmission estimator is 1 - F
    # (main.py:1167-1169): finite, so the same scene has no non-finite term;
    # at 0.0018 the rough branch (and its 0/0) is taken again.  So near the
    # threshold a second source appears: with alpha = roughness^2 the GGX
    # D = alpha^2 / (pi denom^2) of main.py:900-907 and the terms built on
    # it leave float range on some half vectors and the estimator is
    # inf/inf; by roughness 0.02 it no longer fires.
    cfg = getattr(configs, name)()
    smooth = _nf_stats(cfg, w, nb, roughness=0.0017, mat_id=mat_id)
    assert smooth["nonfinite"] == 0 and smooth["nf_btdf_eta1"] == 0
    edge = _nf_stats(cfg, w, nb, roughness=0.0018, mat_id=mat_id)
    assert edge["nf_btdf_eta1"] > 0 and edge["nf_other_estimator"] > 0
    assert edge["nonfinite"] <= edge["nf_btdf_eta1"] + edge["nf_other_estimator"]
    mid = _nf_stats(cfg, w, nb, roughness=0.02, mat_id=mat_id)
    assert mid["nf_other_estimator"] == 0 and mid["nonfinite"] == mid["nf_btdf_eta1"] - mid["nf_btdf_eta1_last"]


def test_heightfield_has_no_nonfinite_term():
    # The heightfield is an open sheet: no path leaves a closed rough volume.
    st = _nf_stats(configs.heightfield_1m(), 1024, 1, cols=(300, 316))
    assert st["segments"] > 100000
    assert st["nonfinite"] == 0 and st["nf_btdf_eta1"] == 0
