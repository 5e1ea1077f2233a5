"""Spectral authoring formats beyond the reference's loaders: every
refractiveindex.info YAML variant (RefractiveIndex(formulas=True)) and the
curve editors' export formats.

The dispersion formulas are pinned by catalogue values (Schott N-BK7 n_d =
1.5168 from its Sellmeier coefficients, fused silica n_d = 1.4585 from
Malitson 1965, standard air n - 1 = 2.77e-4 at 587.6 nm) and by the
reductions between formulas that the mathematics fixes (formula 4 with two
poles and C3 = C5 = C7 = C9 = 2 is Sellmeier; a Lorentz-Lorenz constant
(formula 8) has a closed-form n; one Exotic pole (formula 9) is a
Sellmeier-2 term; Herzberger without its poles is a Cauchy series in
lambda^2), not by restating each formula."""
import math

import numpy as np
import pytest

from spectral.data import (ColourSystem, EEMMatrix, RefractiveIndex, SpecularVec, _dispersion, eem_table_text,
                           spec_table_text)

ND = 0.5875618  # helium d line, micrometres
BK7_S2 = [0, 1.03961212, 0.00600069867, 0.231792344, 0.0200179144, 1.01046945, 103.560653]
FS_S1 = [0, 0.6961663, 0.0684043, 0.4079426, 0.1162414, 0.8974794, 9.896161]


def _yaml(tmp_path, body, name="m.yml"):
    p = tmp_path / name
    p.write_text(body)
    return str(p)


def test_catalogue_values():
    assert _dispersion(2, BK7_S2, ND) == pytest.approx(1.5168, abs=5e-5)
    assert _dispersion(1, FS_S1, ND) == pytest.approx(1.4585, abs=5e-5)
    air = [0, 0.05792105, 238.0185, 0.00167917, 57.362]
    assert _dispersion(6, air, ND) - 1 == pytest.approx(2.772e-4, abs=2e-7)


def test_formula_reductions():
    lam = np.linspace(0.4, 0.8, 7)
    # Sellmeier with squared poles == Sellmeier-2 with the squares
    s2 = [FS_S1[0]] + [v for a, b in zip(FS_S1[1::2], FS_S1[2::2]) for v in (a, b * b)]
    assert np.allclose(_dispersion(1, FS_S1, lam), _dispersion(2, s2, lam), rtol=1e-14)
    # formula 4: two poles with exponents 2 == Sellmeier, and n^2 = 1 + C1 + ...
    c = [1.0, FS_S1[1], 2, FS_S1[2], 2, FS_S1[3], 2, FS_S1[4], 2]
    ref = np.sqrt(1 + FS_S1[1] * lam ** 2 / (lam ** 2 - FS_S1[2] ** 2) + FS_S1[3] * lam ** 2 / (lam ** 2 - FS_S1[4] ** 2))
    assert np.allclose(_dispersion(4, c, lam), ref, rtol=1e-14)
    # Lorentz-Lorenz with a constant: n^2 = (1 + 2 r) / (1 - r)
    r = 0.3
    assert np.allclose(_dispersion(8, [r, 0, 0, 0], lam), math.sqrt((1 + 2 * r) / (1 - r)))
    # one Exotic pole is a Sellmeier-2 term: 1 + a + b L^2/(L^2 - p) = (1 + a + b) + b p / (L^2 - p)
    a, b, p = 0.2, 1.1, 0.01
    assert np.allclose(_dispersion(9, [1 + a + b, b * p, p, 0, 0, 1], lam), _dispersion(2, [a, b, p], lam),
                       rtol=1e-13)
    # Herzberger without poles == Cauchy series in L^2; polynomial n^2 with one term
    h = _dispersion(7, [1.5, 0, 0, 0.01, 0.002, 0.0003], lam)
    assert np.allclose(h, _dispersion(5, [1.5, 0.01, 2, 0.002, 4, 0.0003, 6], lam), rtol=1e-14)
    assert np.allclose(_dispersion(3, [2.25], lam), 1.5)


def test_yaml_variants(tmp_path):
    wl = "0.40 1.50 0.10\n0.60 1.48 0.05\n0.80 1.47 0.01"
    nk = RefractiveIndex(_yaml(tmp_path, f"DATA:\n  - type: tabulated nk\n    data: |\n" +
                               "".join(f"      {l}\n" for l in wl.split("\n"))), 400, 800, 9)
    sep = RefractiveIndex(_yaml(tmp_path, "DATA:\n  - type: tabulated n\n    data: |\n" +
                                "".join(f"      {l.split()[0]} {l.split()[1]}\n" for l in wl.split("\n")) +
                                "  - type: tabulated k\n    data: |\n" +
                                "".join(f"      {l.split()[0]} {l.split()[2]}\n" for l in wl.split("\n")), "s.yml"),
                          400, 800, 9, formulas=True)
    assert np.allclose(nk.n, sep.n) and np.allclose(nk.k, sep.k)
    # the same file through both loaders gives the reference's numbers
    both = RefractiveIndex(nk.file_path, 400, 800, 9, formulas=True)
    assert both.n == nk.n and both.k == nk.k
    # formula n + tabulated k; wavelengths outside the range are clamped
    body = ("DATA:\n  - type: formula 2\n    wavelength_range: 0.45 0.75\n"
            "    coefficients: " + " ".join(map(str, BK7_S2)) + "\n"
            "  - type: tabulated k\n    data: |\n      0.4 1e-8\n      0.8 3e-8\n")
    g = RefractiveIndex(_yaml(tmp_path, body, "g.yml"), 400, 800, 5, formulas=True)
    lam = np.clip(np.linspace(0.4, 0.8, 5), 0.45, 0.75)
    assert np.allclose(g.n, _dispersion(2, BK7_S2, lam))
    assert np.allclose(g.k, np.interp(np.linspace(400, 800, 5), [400, 800], [1e-8, 3e-8]))


def test_reference_mode_rejects_formulas(tmp_path, capsys):
    body = "DATA:\n  - type: formula 2\n    coefficients: " + " ".join(map(str, BK7_S2)) + "\n"
    r = RefractiveIndex(_yaml(tmp_path, body), 400, 800, 5)
    assert r.n is None and r.k is None
    assert "tabulated nk" in capsys.readouterr().out


def test_spectrum_editor_export(tmp_path):
    wl = np.linspace(380.0, 780.0, 41)
    curves = np.stack([np.exp(-((wl - c) / 40.0) ** 2) for c in (600, 550, 450)])
    text = spec_table_text(wl, curves)
    first = text.splitlines()[0].split()
    assert first[0] == "380" and first[1] == f"{curves[0, 0]:.5f}"
    p = tmp_path / "curves.txt"
    p.write_text(text)
    for col in (1, 2, 3):
        v = SpecularVec(str(p), 2.0, 400, 700, 13, column=col).vec
        assert np.allclose(v, 2.0 * np.interp(np.linspace(400, 700, 13), wl, curves[col - 1]), atol=2e-5)
    cs = ColourSystem(str(p), (0.64, 0.33), (0.30, 0.60), (0.15, 0.06), (0.3127, 0.3290), 400, 700, 13)
    assert cs.cmf.shape == (13, 3)


def test_eem_editor_export(tmp_path):
    n = 16
    t = np.linspace(0, 1, n)
    ex = np.exp(-((t - 0.3) / 0.1) ** 2)
    em = np.exp(-((t - 0.6) / 0.15) ** 2)
    text = eem_table_text(400.0, 700.0, ex, em, 0.8)
    rows = [l.split() for l in text.splitlines()]
    assert rows[0][0] == "#" and rows[1][0] == "0" and len(rows) == n + 2
    z = np.array([[float(v) for v in r[1:]] for r in rows[2:]])
    assert np.allclose(np.triu(z, 1), np.triu(np.outer(ex, em) * 0.8, 1), atol=1e-6)
    assert (np.tril(z) == 0).all()
    p = tmp_path / "eem.txt"
    p.write_text(text)
    m = EEMMatrix(400, 700, n, file_path=str(p)).matrix
    # on the source grid the interpolating spline reproduces the table;
    # the Stokes mask zeroes emission <= excitation
    assert np.allclose(np.triu(m, 1), np.triu(z, 1), atol=1e-5)
    assert (np.tril(m) == 0).all()
