"""Pins the CPU restatement's unit functions against what mathematics fixes:
published known-answer vectors, closed forms, normalisation integrals and
chi^2 tests of each sampler against the density it must draw from."""
import math

import numpy as np
import pytest
from scipy import integrate, stats

from oracle import oracle as O


# ---------------------------------------------------------------- RNG
def test_philox_known_answers():
    # Random123 kat_vectors, philox4x32_10
    assert list(O.philox([0, 0, 0, 0], [0, 0])) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert list(O.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2)) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert list(O.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0])) == [
        0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_uniform_stream_layout():
    seed = 0x0123456789ABCDEF
    u = O.uniforms(seed, 7, 3, 5, 10)
    key = [seed & 0xFFFFFFFF, seed >> 32]
    w0 = O.philox([7, 3, 5, 0], key)
    w1 = O.philox([7, 3, 5, 1], key)
    words = list(w0) + list(w1)
    expect = np.array([(w >> 8) / 16777216.0 for w in words[:8]], np.float32)
    assert np.array_equal(u[:8], expect)
    assert (u >= 0).all() and (u < 1).all()


# ---------------------------------------------------------------- Fresnel
@pytest.mark.parametrize("ni,ki,nt,kt", [(1.0, 0.0, 1.5, 0.0), (1.0, 0.0, 0.2, 3.6), (1.5, 0.0, 1.0, 0.0),
                                         (1.0, 0.0, 1.4, 1e-3), (1.33, 0.1, 0.8, 2.0)])
def test_fresnel_normal_incidence_closed_form(ni, ki, nt, kt):
    expect = ((ni - nt) ** 2 + (ki - kt) ** 2) / ((ni + nt) ** 2 + (ki + kt) ** 2)
    assert O.fresnel(1.0, ni, ki, nt, kt) == pytest.approx(expect, rel=2e-5, abs=1e-7)


def test_fresnel_index_matched_is_zero():
    for c in (1.0, 0.7, 0.3, 0.05):
        assert O.fresnel(c, 1.0, 0.0, 1.0, 0.0) == pytest.approx(0.0, abs=1e-6)


def _fresnel_textbook(c, n1c, n2c):
    """Unpolarised Fresnel reflectance for complex indices (Born & Wolf):
    n2 cos(t) = sqrt(n2^2 - n1^2 sin^2(i)), principal branch."""
    s2 = 1.0 - c * c
    a = n1c * c
    b = np.sqrt(n2c * n2c - n1c * n1c * s2 + 0j)
    rs = (a - b) / (a + b)
    rp = (n2c * n2c * a - n1c * n1c * b) / (n2c * n2c * a + n1c * n1c * b)
    return 0.5 * (abs(rs) ** 2 + abs(rp) ** 2)


@pytest.mark.parametrize("ni,ki,nt,kt", [(1.0, 0.0, 1.5, 0.0), (1.0, 0.0, 0.47, 2.4), (1.0, 0.0, 1.4, 0.0),
                                         (1.5, 0.0, 1.0, 0.0)])
def test_fresnel_oblique_against_textbook(ni, ki, nt, kt):
    for c in np.linspace(0.02, 1.0, 25):
        got = O.fresnel(float(c), ni, ki, nt, kt)
        exp = _fresnel_textbook(float(c), complex(ni, ki), complex(nt, kt))
        assert got == pytest.approx(exp, rel=1e-4, abs=1e-6), c
    # total internal reflection beyond the critical angle
    if ni > nt and ki == 0 and kt == 0:
        crit = math.sqrt(1 - (nt / ni) ** 2)
        assert O.fresnel(crit * 0.5, ni, ki, nt, kt) == pytest.approx(1.0, abs=1e-5)


# ---------------------------------------------------------------- microfacet
@pytest.mark.parametrize("r", [0.3, 0.45, 0.6, 0.9])
def test_ggx_D_normalised(r):
    # D uses alpha = r^2 (main.py:902-904): integral of D(h) (n.h) over the
    # hemisphere is 1; with x = cos^2: pi * int_0^1 D(sqrt(x)) dx.
    val, _ = integrate.quad(lambda x: O.ggx_D(math.sqrt(x), r), 0.0, 1.0, points=[1 - r ** 4, 1.0], limit=200)
    assert math.pi * val == pytest.approx(1.0, rel=2e-4)


def test_ggx_G_limits():
    n = [0, 0, 1]
    assert O.ggx_G(n, n, n, 0.5) == pytest.approx(1.0, abs=1e-5)  # both at normal incidence
    grazing = [1, 0, 0]
    assert O.ggx_G(n, n, grazing, 0.5) == pytest.approx(0.0, abs=1e-6)
    # k = max(r^2, 1e-4)^2 / 2 (main.py:912-915): G1 at NdotV = 0.5
    r = 0.7
    k = (r * r) ** 2 / 2
    v = [math.sqrt(0.75), 0, 0.5]
    g1 = 0.5 / (0.5 * (1 - k) + k + 1e-7)
    assert O.ggx_G(n, v, n, r) == pytest.approx(g1 * (1.0 / (1.0 + 1e-7)), rel=1e-5)


def _chi2_cos(cos_samples, cdf, nbins=40):
    edges = np.linspace(cos_samples.min() - 1e-9, 1.0, nbins + 1)
    edges[0] = min(edges[0], -1.0 if cos_samples.min() < 0 else 0.0)
    obs, _ = np.histogram(cos_samples, bins=edges)
    probs = np.diff([cdf(e) for e in edges])
    exp = probs * len(cos_samples)
    keep = exp > 20
    chi2 = ((obs[keep] - exp[keep]) ** 2 / exp[keep]).sum()
    dof = keep.sum() - 1
    return stats.chi2.sf(chi2, dof)


def test_cosine_hemisphere_sampler():
    axis = np.array([0.3, -0.5, 0.8], np.float32)
    axis /= np.linalg.norm(axis)
    d = O.sample(0, 11, 200000, axis)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=2e-6)
    c = d @ axis.astype(np.float64)
    assert c.min() > -1e-6
    assert _chi2_cos(np.clip(c, 0, 1), lambda x: min(max(x, 0.0), 1.0) ** 2) > 1e-4


@pytest.mark.parametrize("g", [0.0, 0.5, 0.8, -0.3])
def test_hg_normalised_and_sampler(g):
    val, _ = integrate.quad(lambda m: O.hg_eval(m, g), -1.0, 1.0, limit=200)
    assert 2 * math.pi * val == pytest.approx(1.0, rel=1e-3)
    axis = np.array([0.0, 1.0, 0.0], np.float32)  # exercises the (1,0,0) up-vector switch
    d = O.sample(1, 12, 200000, axis, g)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=2e-6)
    c = d @ axis.astype(np.float64)

    def cdf(x):
        x = min(max(x, -1.0), 1.0)
        v, _ = integrate.quad(lambda m: 2 * math.pi * O.hg_eval(m, g), -1.0, x, limit=200)
        return v / (2 * math.pi * val)

    assert _chi2_cos(c, cdf) > 1e-4


@pytest.mark.parametrize("r", [0.2, 0.5])
def test_ggx_half_vector_sampler_density(r):
    # importance_sample_ggx is called with alpha = roughness, so a2 = r^2
    # (main.py:970, 1019): the half vectors follow D_{a2=r^2}(h) (n.h).
    a2 = r * r

    def D_s(m):
        den = m * m * (a2 - 1) + 1
        return a2 / (math.pi * den * den)

    axis = np.array([0.0, 0.0, 1.0], np.float32)
    h = O.sample(2, 13, 200000, axis, r)
    c = h @ axis.astype(np.float64)
    tot, _ = integrate.quad(lambda m: 2 * math.pi * D_s(m) * m, 0.0, 1.0, limit=200)
    assert tot == pytest.approx(1.0, rel=1e-4)

    def cdf(x):
        x = min(max(x, 0.0), 1.0)
        v, _ = integrate.quad(lambda m: 2 * math.pi * D_s(m) * m, 0.0, x, limit=200)
        return v

    assert _chi2_cos(np.clip(c, 0, 1), cdf) > 1e-4


def test_unit_sphere_sampler():
    d = O.sample(3, 14, 200000, [0, 0, 1])
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=2e-6)
    assert _chi2_cos(d[:, 2].astype(np.float64), lambda x: (min(max(x, -1.0), 1.0) + 1) / 2) > 1e-4
    phi = np.arctan2(d[:, 1], d[:, 0])
    obs, _ = np.histogram(phi, bins=24, range=(-math.pi, math.pi))
    assert stats.chisquare(obs).pvalue > 1e-4
