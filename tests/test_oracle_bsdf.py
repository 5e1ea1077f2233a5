"""Pins the oracle's surface shading -- the Oren-Nayar factor, the estimators
of scatter_surface (GGX reflection, reflection on total internal reflection,
rough transmission, the Oren-Nayar lobe) and evaluate_bsdf -- against what
the mathematics fixes, independently of how main.py writes them:

* Oren-Nayar (main.py:1071-1088, 1396-1413): the qualitative model of Oren &
  Nayar (1994) evaluated in float64 from the polar and azimuthal angles of a
  tangent frame, sigma = roughness * pi / 2;
* GGX reflection (main.py:1039-1053, 1142-1156): with bsdf = D G F / (4 NL NV)
  and pdf = D NH / (4 V.H) the estimator bsdf cos / pdf reduces to
  G F (V.H) cos / (NL NV NH) for every sample -- D and both 4s cancel --, and
  its expectation is the integral of that over the sampler's half-vector
  density D_{alpha=r}(h) (n.h) (a float64 quadrature);
* rough transmission (main.py:1166-1186): bsdf NL / pdf reduces to
  G |V.H| / (NV NH);
* the Oren-Nayar lobe: estimator (1 - F) albedo ON, whose expectation under
  cosine sampling is the float64 quadrature of the model (at most 1 for
  albedo 1 and F = 0);
* evaluate_bsdf (main.py:1360-1444): its GGX lobe equals Walter et al.
  (2007)'s D (alpha = roughness^2) times G F / (4 NL NV) pointwise, its pdf
  D NH F / (4 V.H) integrates to 1 over all light directions at normal view
  (the Jacobian of the half-vector map is 1 / (4 V.H)), and its Oren-Nayar
  lobe is (1 - F) (albedo / pi) ON with pdf (1 - F) NL / pi.

Each check is also run on the oracle built with one planted mistake (sin and
tan swapped, a 4 dropped, a cosine exchanged, ...): it must fail there."""
import math
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O

ROOT = Path(__file__).resolve().parents[1]
WL = 550.0
N = np.array([0.0, 0.0, 1.0])


# ---------------------------------------------------------------- closed forms (float64)
def fresnel_bw(c, n1, n2):
    """Unpolarised Fresnel reflectance for complex indices (Born & Wolf)."""
    c = np.asarray(c, np.float64)
    s2 = 1.0 - c * c
    a = n1 * c
    b = np.sqrt(n2 * n2 - n1 * n1 * s2 + 0j)
    rs = (a - b) / (a + b)
    rp = (n2 * n2 * a - n1 * n1 * b) / (n2 * n2 * a + n1 * n1 * b)
    return 0.5 * (np.abs(rs) ** 2 + np.abs(rp) ** 2)


def ggx_D_walter(cos_h, alpha):
    """Walter et al. 2007, eq. 33: alpha^2 / (pi cos^4 (alpha^2 + tan^2)^2)."""
    c2 = np.asarray(cos_h, np.float64) ** 2
    tan2 = (1.0 - c2) / c2
    return alpha ** 2 / (np.pi * c2 * c2 * (alpha ** 2 + tan2) ** 2)


def smith_schlick_G(nv, nl, r):
    """The separable Smith-Schlick G of main.py:909-923: k = alpha^2 / 2 with
    alpha = max(r^2, 1e-4) (pinned on its own in test_oracle_units.py)."""
    a = max(r * r, 1e-4)
    k = a * a / 2.0
    nv, nl = np.maximum(nv, 0.0), np.maximum(nl, 0.0)
    return nv / (nv * (1 - k) + k + 1e-7) * nl / (nl * (1 - k) + k + 1e-7)


def frame(n):
    n = np.asarray(n, np.float64)
    a = np.array([1.0, 0.0, 0.0]) if abs(n[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    t = np.cross(a, n)
    t /= np.linalg.norm(t)
    return t, np.cross(n, t)


def oren_nayar_model(r, l, v, n):
    """Oren & Nayar (1994), qualitative model: A + B max(0, cos(phi_i - phi_o))
    sin(alpha) tan(beta), A = 1 - 0.5 s^2 / (s^2 + 0.33), B = 0.45 s^2 /
    (s^2 + 0.09), alpha = max(theta_i, theta_o), beta = min(...)."""
    s2 = (r * np.pi / 2) ** 2
    A = 1.0 - 0.5 * s2 / (s2 + 0.33)
    B = 0.45 * s2 / (s2 + 0.09)
    t, b = frame(n)
    ci, co = float(np.dot(n, l)), float(np.dot(n, v))
    ti, to = math.acos(max(1e-4, ci)), math.acos(max(1e-4, co))
    cphi = 0.0
    if ci > 1e-4 and co > 1e-4:
        pi_ = math.atan2(float(np.dot(l, b)), float(np.dot(l, t)))
        po = math.atan2(float(np.dot(v, b)), float(np.dot(v, t)))
        cphi = max(0.0, math.cos(pi_ - po))
    al, be = max(ti, to), min(ti, to)
    return A + B * cphi * math.sin(al) * math.tan(be)


def hemisphere_dirs(rng, n, k, cmin=0.05):
    """k random unit vectors with cos to n in (cmin, 0.999)."""
    t, b = frame(n)
    c = rng.uniform(cmin, 0.999, k)
    ph = rng.uniform(0, 2 * np.pi, k)
    s = np.sqrt(1 - c * c)
    return (np.outer(s * np.cos(ph), t) + np.outer(s * np.sin(ph), b) + np.outer(c, n))


def view(cosv, phi=0.0):
    s = math.sqrt(1 - cosv * cosv)
    return np.array([s * math.cos(phi), s * math.sin(phi), cosv])


# ---------------------------------------------------------------- the checks
def check_oren_nayar():
    rng = np.random.default_rng(1)
    for r in (0.1, 0.3, 0.6):
        for _ in range(6):
            n = rng.normal(size=3)
            n /= np.linalg.norm(n)
            ls = hemisphere_dirs(rng, n, 40)
            vs = hemisphere_dirs(rng, n, 40)
            for l, v in zip(ls, vs):
                got = O.oren_nayar(r, max(1e-4, l @ n), max(1e-4, v @ n), l, v, n)
                want = oren_nayar_model(r, l, v, n)
                assert abs(got - want) <= 2e-4 * abs(want) + 1e-6, (r, got, want)
    # closed-form reductions: r = 0 -> 1 exactly; l = v -> A + B sin(t) tan(t)
    assert O.oren_nayar(0.0, 0.6, 0.8, [0.8, 0, 0.6], [0, 0.6, 0.8], N) == 1.0
    s2 = (0.5 * np.pi / 2) ** 2
    th = math.acos(0.6)
    want = 1 - 0.5 * s2 / (s2 + 0.33) + 0.45 * s2 / (s2 + 0.09) * math.sin(th) * math.tan(th)
    assert O.oren_nayar(0.5, 0.6, 0.6, [0.8, 0, 0.6], [0.8, 0, 0.6], N) == pytest.approx(want, rel=1e-5)


def ggx_reflection_expectation(cosv, r, ni, nt_c, nth=1500, nph=360):
    """E[estimator of the reflected samples] = integral over the sampler's
    half-vector density D_{alpha=r}(h) (n.h) of P(reflect | h) times
    G F (V.h) cos / (NL NV NH).  A sample reflects with probability F (1
    once F > 0.99: the uniform is clamped to 0.99, main.py:1027) and, when
    it penetrates a volume material, again on total internal reflection
    (eta^2 (1 - (d.h)^2) > 1 for eta = n_i / n_t, main.py:1129-1156), with
    the same estimator (G vanishes below the horizon in both branches)."""
    V = view(cosv)
    s = (np.arange(nth) + 0.5) / nth
    th = (np.pi / 2) * s * s  # dense near the normal
    dth = (np.pi / 2) * 2 * s / nth
    ph = (np.arange(nph) + 0.5) / nph * 2 * np.pi
    T, P = np.meshgrid(th, ph, indexing="ij")
    h = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1)
    dw = (np.sin(T) * dth[:, None]) * (2 * np.pi / nph)
    dens = ggx_D_walter(np.cos(T), r) * np.cos(T)
    vh = h @ V
    ia = np.maximum(1e-4, vh)
    R = 2 * vh[..., None] * h - V  # reflect(-V, h)
    nr = R[..., 2]
    F = fresnel_bw(ia, complex(ni), nt_c)
    pref = np.where(F > 0.99, 1.0, F)
    eta = ni / nt_c.real
    tir = eta * eta * (1.0 - vh * vh) > 1.0
    pref = pref + (1.0 - pref) * tir
    nl = np.maximum(1e-4, nr)
    nh = np.maximum(1e-4, np.cos(T))
    nv = max(1e-4, cosv)
    G = smith_schlick_G(cosv, nr, r)
    est = G * F * ia * np.maximum(0.0, nr) / (nl * nv * nh)
    return float(np.sum(dens * pref * est * dw))


def check_ggx_reflection():
    cases = [((1.0, 0.0, 0.2, 3.6, 0.2, 0.0, 1), 0.9), ((1.0, 0.0, 0.2, 3.6, 0.5, 0.0, 1), 0.5),
             ((1.0, 0.0, 1.5, 1.8, 0.3, 0.0, 1), 0.3), ((1.0, 0.0, 0.47, 2.4, 0.2, 0.0, 1), 0.1)]
    for mat, cosv in cases:
        ni, nt_c, r = mat[0], complex(mat[2], mat[3]), mat[4]
        V = view(cosv)
        res = O.scatter_surface(mat, WL, -V, N, 11, 100_000)
        # conductor (is_true_volume = 1): the penetrating branch refracts with
        # G = 0, so only reflections carry an estimator
        m = res["reflect"]
        R, H = res["dir"][m].astype(np.float64), res["h"][m].astype(np.float64)
        ia = np.maximum(1e-4, H @ V)
        nr = R[:, 2]
        F = fresnel_bw(ia, complex(ni), nt_c)
        G = smith_schlick_G(cosv, nr, r)
        want = G * F * ia * np.maximum(0.0, nr) / (np.maximum(1e-4, nr) * max(1e-4, cosv) * np.maximum(1e-4, H[:, 2]))
        got = res["est"][m].astype(np.float64)
        ok = np.abs(got - want) <= 2e-3 * np.abs(want) + 1e-6
        assert ok.mean() > 0.999, (mat, cosv, ok.mean())
        # the expectation over the sampler (every sample, 0 off this branch)
        est = np.where(res["reflect"], res["est"], 0.0).astype(np.float64)
        mean, se = est.mean(), est.std() / math.sqrt(len(est))
        q = ggx_reflection_expectation(cosv, r, ni, nt_c)
        assert abs(mean - q) <= 4 * se + 2e-3 * q, (mat, cosv, mean, se, q)


def check_ggx_energy_at_normal_view():
    """At normal view the sampler and G agree well enough that the GGX
    reflection of a near-perfect conductor (n = 1, k = 1e4: F within 1e-7 of
    1) stays below 1 for every roughness.  (At grazing views the
    reference's estimator exceeds 1: its half vectors are drawn with
    alpha = r while G uses alpha = r^2, main.py:970 vs 911.)"""
    for r in (0.05, 0.2, 0.5, 0.8):
        res = O.scatter_surface((1.0, 0.0, 1.0, 1e4, r, 0.0, 1), WL, -N, N, 12, 50_000)
        est = np.where(res["reflect"], res["est"], 0.0).astype(np.float64)
        se = est.std() / math.sqrt(len(est))
        assert est.mean() <= 1.0 + 3 * se, (r, est.mean())
        q = ggx_reflection_expectation(0.99999, r, 1.0, complex(1.0, 1e4))
        assert abs(est.mean() - q) <= 4 * se + 2e-3 * q, (r, est.mean(), q)


def check_rough_transmission():
    """Glass to air at oblique views (the only case where the refracted ray
    can leave on the viewer's side, N.L > 0, so that G > 0)."""
    for cosv, seed in ((0.3, 21), (0.1, 22), (0.6, 23)):
        mat = (1.5, 0.0, 1.0, 0.0, 0.3, 0.0, 1)
        V = view(cosv)
        res = O.scatter_surface(mat, WL, -V, N, seed, 50_000)
        m = ~res["reflect"]
        L, H = res["dir"][m].astype(np.float64), res["h"][m].astype(np.float64)
        vh = H @ V
        G = smith_schlick_G(cosv, L[:, 2], 0.3)
        want = G * np.abs(vh) / (max(1e-4, cosv) * np.maximum(1e-4, H[:, 2]))
        got = res["est"][m].astype(np.float64)
        fin = np.isfinite(got)
        assert fin.mean() > 0.999
        ok = np.abs(got[fin] - want[fin]) <= 2e-3 * np.abs(want[fin]) + 1e-6
        assert ok.mean() > 0.999, (cosv, ok.mean())
        if cosv <= 0.3:
            assert (want > 0).sum() > 20  # the identity was exercised off zero
    # air to glass: every refracted ray goes below the surface, G = 0
    res = O.scatter_surface((1.0, 0.0, 1.5, 0.0, 0.3, 0.0, 1), WL, -view(0.7), N, 24, 5000)
    assert np.all(res["est"][~res["reflect"]] == 0.0)


def on_lobe_expectation(cosv, r, nth=400, nph=360):
    """Integral of ON(l, v) cos / pi over the hemisphere (float64 model)."""
    V = view(cosv)
    th = (np.arange(nth) + 0.5) / nth * np.pi / 2
    ph = (np.arange(nph) + 0.5) / nph * 2 * np.pi
    s2 = (r * np.pi / 2) ** 2
    A = 1 - 0.5 * s2 / (s2 + 0.33)
    B = 0.45 * s2 / (s2 + 0.09)
    T, P = np.meshgrid(th, ph, indexing="ij")
    to = math.acos(cosv)
    al, be = np.maximum(T, to), np.minimum(T, to)
    cphi = np.maximum(0.0, np.cos(P))  # V at azimuth 0
    on = A + B * cphi * np.sin(al) * np.tan(be)
    dw = np.sin(T) * (np.pi / 2 / nth) * (2 * np.pi / nph)
    return float(np.sum(on * np.cos(T) / np.pi * dw))


def back_facing_half_vectors(cosv, r, nth=1500, nph=360):
    """Probability that the sampler's half vector (density D_{alpha=r}(h)
    (n.h)) faces away from the viewer, V.h <= 0."""
    V = view(cosv)
    s = (np.arange(nth) + 0.5) / nth
    th = (np.pi / 2) * s * s
    dth = (np.pi / 2) * 2 * s / nth
    ph = (np.arange(nph) + 0.5) / nph * 2 * np.pi
    T, P = np.meshgrid(th, ph, indexing="ij")
    h = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1)
    dw = (np.sin(T) * dth[:, None]) * (2 * np.pi / nph)
    dens = ggx_D_walter(np.cos(T), r) * np.cos(T)
    return float(np.sum(dens * (h @ V <= 0.0) * dw))


def check_oren_nayar_lobe():
    """Opaque wall with n = 1, sigma_s = 1, k = 0 (albedo 1 / (1 + 1e-6)).
    The lobe's direction comes from its own uniforms, so given that the
    sample took the lobe, E[estimator] = albedo x the cosine integral of the
    model (and at most 1).  Which samples take it: F at the half vector's
    max(1e-4, V.h) is 0 for matched indices -- except at the 1e-4 clamp,
    where the reference's f32 arithmetic returns F = 1 (sin^2 rounds to 1),
    so exactly the half vectors facing away from the viewer reflect
    (main.py:1020-1028)."""
    assert O.fresnel(1e-4, 1.0, 0.0, 1.0, 0.0) == 1.0 and O.fresnel(1e-3, 1.0, 0.0, 1.0, 0.0) < 1e-3
    for cosv, r, seed in ((1.0, 0.2, 31), (0.7, 0.6, 32), (0.3, 1.0, 33), (0.1, 0.6, 34)):
        res = O.scatter_surface((1.0, 0.0, 1.0, 0.0, r, 1.0, 0), WL, -view(cosv), N, seed, 100_000)
        S = res["dir"].astype(np.float64)
        lobe = np.abs(res["pdf"] - 0.318309886184 * np.maximum(1e-4, S[:, 2])) <= 1e-5 * res["pdf"] + 1e-9
        back = back_facing_half_vectors(min(cosv, 0.99999), r)
        n = len(lobe)
        assert abs((1.0 - lobe.mean()) - back) <= 4 * math.sqrt(back * (1 - back) / n) + 1e-4, (cosv, r, back)
        est = res["est"][lobe].astype(np.float64)
        mean, se = est.mean(), est.std() / math.sqrt(len(est))
        q = on_lobe_expectation(min(cosv, 0.9999), r) / (1 + 1e-6)
        assert abs(mean - q) <= 4 * se + 1e-3 * q, (cosv, r, mean, se, q)
        assert mean <= 1.0


def check_eval_bsdf():
    rng = np.random.default_rng(3)
    # GGX lobe alone (is_true_volume = 1 removes the diffuse lobe)
    for r, nt_c in ((0.3, complex(0.2, 3.6)), (0.6, complex(1.5, 0.0)), (0.15, complex(0.47, 2.4))):
        mat = (1.0, 0.0, nt_c.real, nt_c.imag, r, 0.0, 1)
        Vs = hemisphere_dirs(rng, N, 300)
        Ls = hemisphere_dirs(rng, N, 300)
        b, p = O.eval_bsdf(mat, WL, Vs, Ls, N)
        H = Vs + Ls
        H /= np.linalg.norm(H, axis=1, keepdims=True)
        vh = np.maximum(1e-4, np.sum(Vs * H, 1))
        nv, nl, nh = Vs[:, 2], Ls[:, 2], H[:, 2]
        F = fresnel_bw(vh, complex(1.0), nt_c)
        D = ggx_D_walter(nh, r * r)
        G = smith_schlick_G(nv, nl, r)
        assert np.allclose(b, D * G * F / (4 * nl * nv), rtol=2e-4, atol=1e-7), r
        assert np.allclose(p, D * nh * F / (4 * vh), rtol=2e-4, atol=1e-7), r
    # normalisation of the pdf at normal view: F = 1 (n = 1, k = 1e4)
    for r in (0.4, 0.7):
        nth, nph = 3000, 90
        s = (np.arange(nth) + 0.5) / nth
        th = np.pi * s * s  # all light directions, dense near the normal
        dth = np.pi * 2 * s / nth
        ph = (np.arange(nph) + 0.5) / nph * 2 * np.pi
        T, P = np.meshgrid(th, ph, indexing="ij")
        L = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
        dw = (np.sin(T) * dth[:, None] * (2 * np.pi / nph)).reshape(-1)
        V = np.broadcast_to(N, L.shape)
        _, p = O.eval_bsdf((1.0, 0.0, 1.0, 1e4, r, 0.0, 1), WL, V, L, N)
        assert float(np.sum(p * dw)) == pytest.approx(1.0, abs=5e-3), r
    # Oren-Nayar lobe: (1 - F) albedo / pi ON, pdf (1 - F) NL / pi
    for r in (0.2, 0.7):
        mat = (1.0, 0.0, 1.4, 2e-8, r, 1.0, 0)
        Vs = hemisphere_dirs(rng, N, 200)
        Ls = hemisphere_dirs(rng, N, 200)
        b, p = O.eval_bsdf(mat, WL, Vs, Ls, N)
        H = Vs + Ls
        H /= np.linalg.norm(H, axis=1, keepdims=True)
        vh = np.maximum(1e-4, np.sum(Vs * H, 1))
        F = fresnel_bw(vh, complex(1.0), complex(1.4, 2e-8))
        albedo = 1.0 / (1.0 + 4 * np.pi * 2e-8 / (WL * 1e-9) + 1e-6)
        on = np.array([oren_nayar_model(r, l, v, N) for l, v in zip(Ls, Vs)])
        nl, nv, nh = Ls[:, 2], Vs[:, 2], H[:, 2]
        spec = ggx_D_walter(nh, r * r) * smith_schlick_G(nv, nl, r) * F / (4 * nl * nv)
        assert np.allclose(b, spec + (1 - F) * albedo / np.pi * on, rtol=5e-4, atol=1e-7), r
        assert np.allclose(p, ggx_D_walter(nh, r * r) * nh * F / (4 * vh) + (1 - F) * nl / np.pi, rtol=5e-4,
                           atol=1e-7), r


CHECKS = {"oren_nayar": check_oren_nayar, "ggx_reflection": check_ggx_reflection,
          "ggx_energy": check_ggx_energy_at_normal_view, "rough_transmission": check_rough_transmission,
          "oren_nayar_lobe": check_oren_nayar_lobe, "eval_bsdf": check_eval_bsdf}


@pytest.mark.parametrize("name", list(CHECKS))
def test_oracle_bsdf(name):
    CHECKS[name]()


# planted mistake (OR_PLANT in oracle.cpp) -> the checks that must catch it
PLANTS = {1: ("oren_nayar", "oren_nayar_lobe", "eval_bsdf"),   # sin / tan swapped
          2: ("ggx_reflection", "ggx_energy"),                   # GGX sampling pdf without its 4
          3: ("rough_transmission",),                            # rough BTDF pdf without |L.H|
          4: ("eval_bsdf",),                                     # evaluate_bsdf pdf without its 4
          5: ("eval_bsdf",),                                     # evaluate_bsdf with N.H for N.L
          6: ("oren_nayar", "oren_nayar_lobe", "eval_bsdf")}     # sigma without pi / 2


@pytest.fixture(scope="module")
def plants():
    subprocess.run(["make", "-s", "-j4", "-C", str(ROOT / "oracle"), "plants"], check=True)


@pytest.mark.parametrize("plant,name", [(k, n) for k, names in PLANTS.items() for n in names])
def test_checks_catch_planted_mistake(plants, plant, name):
    with O.planted(plant):
        with pytest.raises(AssertionError):
            CHECKS[name]()
