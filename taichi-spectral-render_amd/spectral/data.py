"""Spectral table loaders with the reference's file formats and semantics.

Mirrors Xyzonox/taichi-spectral-render data.py:

* ``RefractiveIndex``  -- data.py:6-83: refractiveindex.info YAML, the first
  ``tabulated nk`` block, micrometres -> nanometres, ``np.interp`` onto
  ``np.linspace(wl_min, wl_max, B)``.  Errors print and leave ``n``/``k``
  as ``None`` (data.py:37-42, 51-53).
* ``ColourSystem``     -- data.py:86-128: xy primaries -> RGB->XYZ matrix,
  its inverse scaled so white maps to (1,1,1) (``T``), CMFs interpolated
  onto the band centres, ``normalization_factor = sum(cmf_y) * step``.
* ``SpecularVec``      -- data.py:130-143: 1-D table times a scale.
* ``EEMMatrix``        -- data.py:145-182: 2-D ``RectBivariateSpline``
  (kx = ky = 3, s = 0) of an excitation-emission table, Stokes mask
  (emission <= excitation -> 0), clamp >= 0, ids from a class counter
  starting at 1 (id 0 means "no EEM").

Beyond the reference, every loader also accepts an in-memory table
(``numpy`` array) where the reference takes a file path, so synthetic
tables need no files.  The file formats are unchanged.
"""
from __future__ import annotations

import io
import os

import numpy as np
import yaml
from scipy.interpolate import RectBivariateSpline


def band_wavelengths(wavelength_min, wavelength_max, spectral_bands):
    """Band centres in nm, as every loader computes them (data.py:19)."""
    return np.linspace(wavelength_min, wavelength_max, spectral_bands)


def _load_table(src):
    if isinstance(src, np.ndarray):
        return np.asarray(src, dtype=np.float64)
    if isinstance(src, (str, os.PathLike)):
        return np.loadtxt(src, comments="#")
    if isinstance(src, io.IOBase):
        return np.loadtxt(src, comments="#")
    return np.asarray(src, dtype=np.float64)


def _dispersion(kind, c, lam_um):
    """refractiveindex.info dispersion formula ``kind`` (1-9) with
    coefficients c = [C1, C2, ...] at wavelengths lam_um (micrometres), in
    float64; returns n.  The formulas, as the database documents them:

    1 Sellmeier     n^2 - 1 = C1 + sum C(2i) L^2 / (L^2 - C(2i+1)^2)
    2 Sellmeier-2   n^2 - 1 = C1 + sum C(2i) L^2 / (L^2 - C(2i+1))
    3 Polynomial    n^2 = C1 + sum C(2i) L^C(2i+1)
    4 RI.INFO       n^2 = C1 + C2 L^C3 / (L^2 - C4^C5) + C6 L^C7 / (L^2 - C8^C9) + sum_{i>=10} C(i) L^C(i+1)
    5 Cauchy        n = C1 + sum C(2i) L^C(2i+1)
    6 Gases         n - 1 = C1 + sum C(2i) / (C(2i+1) - L^-2)
    7 Herzberger    n = C1 + C2 / (L^2 - 0.028) + C3 / (L^2 - 0.028)^2 + C4 L^2 + C5 L^4 + C6 L^6
    8 Retro         (n^2 - 1) / (n^2 + 2) = C1 + C2 L^2 / (L^2 - C3) + C4 L^2
    9 Exotic        n^2 = C1 + C2 / (L^2 - C3) + C4 (L - C5) / ((L - C5)^2 + C6)
    """
    c = [float(x) for x in c]
    m = len(c)
    C = c + [0.0] * 18
    L = np.asarray(lam_um, dtype=np.float64)
    L2 = L * L
    pairs = [(C[i], C[i + 1]) for i in range(1, m - 1, 2)]
    if kind == 1:
        return np.sqrt(1.0 + C[0] + sum(a * L2 / (L2 - b * b) for a, b in pairs))
    if kind == 2:
        return np.sqrt(1.0 + C[0] + sum(a * L2 / (L2 - b) for a, b in pairs))
    if kind == 3:
        return np.sqrt(C[0] + sum(a * L ** b for a, b in pairs))
    if kind == 4:
        n2 = C[0] + C[1] * L ** C[2] / (L2 - C[3] ** C[4]) + C[5] * L ** C[6] / (L2 - C[7] ** C[8])
        n2 = n2 + sum(C[i] * L ** C[i + 1] for i in range(9, max(m - 1, 9), 2))
        return np.sqrt(n2)
    if kind == 5:
        return C[0] + sum(a * L ** b for a, b in pairs)
    if kind == 6:
        return 1.0 + C[0] + sum(a / (b - 1.0 / L2) for a, b in pairs)
    if kind == 7:
        q = 1.0 / (L2 - 0.028)
        return C[0] + C[1] * q + C[2] * q * q + C[3] * L2 + C[4] * L2 ** 2 + C[5] * L2 ** 3
    if kind == 8:
        r = C[0] + C[1] * L2 / (L2 - C[2]) + C[3] * L2
        return np.sqrt((1.0 + 2.0 * r) / (1.0 - r))
    if kind == 9:
        return np.sqrt(C[0] + C[1] / (L2 - C[2]) + C[3] * (L - C[4]) / ((L - C[4]) ** 2 + C[5]))
    raise ValueError(f"unknown dispersion formula {kind}")


def _parse_columns(raw, ncol):
    rows = []
    for line in str(raw).strip().split("\n"):
        parts = line.split()
        if len(parts) == ncol:
            try:
                rows.append([float(x) for x in parts])
            except ValueError:
                print(f"Warning: Could not parse line in YAML data: '{line}'")
    return np.array(rows, dtype=np.float64).reshape(-1, ncol)


class RefractiveIndex:
    """n(lambda), k(lambda) from a refractiveindex.info YAML file (data.py:6-83).

    With ``formulas=False`` (the default) this is the reference loader: the
    first ``tabulated nk`` block only, anything else prints an error and
    leaves ``n``/``k`` as ``None``.  ``formulas=True`` reads every variant of
    the database format: ``tabulated nk``, ``tabulated n``, ``tabulated k``
    and ``formula 1`` .. ``formula 9`` (see ``_dispersion``); n comes from
    the first entry that gives n, k from the first that gives k (0 if none).
    Tables are interpolated with ``np.interp`` as in the reference; formulas
    are evaluated at the band wavelengths clamped to the entry's
    ``wavelength_range`` (the same constant continuation ``np.interp`` gives
    a table)."""

    def __init__(self, file_path, wavelength_min, wavelength_max, spectral_bands, formulas=False):
        self.file_path = file_path
        self.formulas = bool(formulas)
        self.target_wavelengths_nm = band_wavelengths(wavelength_min, wavelength_max, spectral_bands)
        self.n, self.k = self._load_and_interpolate_data()

    def _interp(self, wl_um, vals):
        wl_nm = np.asarray(wl_um) * 1000.0
        order = np.argsort(wl_nm)
        return np.interp(self.target_wavelengths_nm, wl_nm[order], np.asarray(vals)[order])

    def _all_variants(self, data):
        n = k = None
        for item in (data or {}).get("DATA", []):
            kind = str(item.get("type", ""))
            if kind.startswith("formula"):
                if n is not None:
                    continue
                f = int(kind.split()[1])
                coef = [float(x) for x in str(item.get("coefficients", "")).split()]
                rng = [float(x) for x in str(item.get("wavelength_range", "0 1e9")).split()]
                lam = np.clip(self.target_wavelengths_nm / 1000.0, rng[0], rng[1])
                n = _dispersion(f, coef, lam)
            elif kind == "tabulated nk":
                t = _parse_columns(item.get("data", ""), 3)
                if len(t) == 0:
                    continue
                if n is None:
                    n = self._interp(t[:, 0], t[:, 1])
                if k is None:
                    k = self._interp(t[:, 0], t[:, 2])
            elif kind in ("tabulated n", "tabulated k"):
                t = _parse_columns(item.get("data", ""), 2)
                if len(t) == 0:
                    continue
                if kind == "tabulated n" and n is None:
                    n = self._interp(t[:, 0], t[:, 1])
                if kind == "tabulated k" and k is None:
                    k = self._interp(t[:, 0], t[:, 1])
        if n is None:
            print(f"Error: Could not find refractive index data in {self.file_path}.")
            return None, None
        if k is None:
            k = np.zeros_like(n)
        return np.asarray(n).tolist(), np.asarray(k).tolist()

    def _load_and_interpolate_data(self):
        try:
            if isinstance(self.file_path, (str, os.PathLike)):
                with open(self.file_path, "r") as f:
                    data = yaml.safe_load(f)
            else:
                data = yaml.safe_load(self.file_path)
        except FileNotFoundError:
            print(f"Error: The file {self.file_path} was not found.")
            return None, None
        except yaml.YAMLError as e:
            print(f"Error parsing YAML file: {e}")
            return None, None
        if self.formulas:
            return self._all_variants(data)

        raw = ""
        for item in (data or {}).get("DATA", []):
            if item.get("type") == "tabulated nk":
                raw = item.get("data")
                break
        if not raw:
            print(f"Error: Could not find 'tabulated nk' data in {self.file_path}.")
            return None, None

        wl_um, n_src, k_src = [], [], []
        for line in raw.strip().split("\n"):
            parts = line.split()
            if len(parts) == 3:
                try:
                    wl_um.append(float(parts[0]))
                    n_src.append(float(parts[1]))
                    k_src.append(float(parts[2]))
                except ValueError:
                    print(f"Warning: Could not parse line in YAML data: '{line}'")
                    continue
        wl_nm = np.array([w * 1000 for w in wl_um])
        order = np.argsort(wl_nm)
        wl_nm = wl_nm[order]
        n_src = np.array(n_src)[order]
        k_src = np.array(k_src)[order]
        n = np.interp(self.target_wavelengths_nm, wl_nm, n_src)
        k = np.interp(self.target_wavelengths_nm, wl_nm, k_src)
        return n.tolist(), k.tolist()


class ColourSystem:
    """CMF table + display primaries (data.py:86-128)."""

    def __init__(self, file_path, red, green, blue, white, wavelength_min, wavelength_max, spectral_bands):
        self.red = self.xyz_from_xy(red)
        self.green = self.xyz_from_xy(green)
        self.blue = self.xyz_from_xy(blue)
        self.white = self.xyz_from_xy(white)
        self.M = np.vstack((self.red, self.green, self.blue)).T
        self.MI = np.linalg.inv(self.M)
        self.wscale = self.MI.dot(self.white)
        self.T = self.MI / self.wscale[:, np.newaxis]
        self.cmf_raw = _load_table(file_path)
        self.target_wavelengths = band_wavelengths(wavelength_min, wavelength_max, spectral_bands)
        self.cmf = self._interpolate_cmf()
        self.wavelength_step = (wavelength_max - wavelength_min) / (spectral_bands - 1)
        self.normalization_factor = np.sum(self.cmf[:, 1]) * self.wavelength_step

    @staticmethod
    def xyz_from_xy(vec):
        x, y = vec[0], vec[1]
        return np.array((x, y, 1 - x - y))

    def _interpolate_cmf(self):
        src_wl = self.cmf_raw[:, 0]
        src = self.cmf_raw[:, 1:]
        ix = np.interp(self.target_wavelengths, src_wl, src[:, 0])
        iy = np.interp(self.target_wavelengths, src_wl, src[:, 1])
        iz = np.interp(self.target_wavelengths, src_wl, src[:, 2])
        return np.vstack((ix, iy, iz)).T


class SpecularVec:
    """A spectrum table times ``scale`` (data.py:130-143).  ``column``
    picks the curve of a multi-curve table (the spectrum editor exports one
    column per curve); 1, the reference's only column, by default."""

    def __init__(self, file_path, scale, wavelength_min, wavelength_max, spectral_bands, column=1):
        self.raw = _load_table(file_path)
        self.column = int(column)
        self.target_wavelengths = band_wavelengths(wavelength_min, wavelength_max, spectral_bands)
        self.vec = self._interpolate(scale)

    def _interpolate(self, scale):
        return (np.interp(self.target_wavelengths, self.raw[:, 0], self.raw[:, self.column]) * scale).tolist()


class EEMMatrix:
    """Excitation-emission matrix on the band grid (data.py:145-182)."""

    _next_eem_id = 1

    def __init__(self, wavelength_min, wavelength_max, spectral_bands, file_path=None, matrix=None):
        self.target_wavelengths = band_wavelengths(wavelength_min, wavelength_max, spectral_bands)
        if file_path is not None:
            self.matrix = self._load_and_interpolate(file_path)
        elif matrix is not None:
            # data.py:151-152: a given matrix is used as is (no Stokes mask)
            self.matrix = np.asarray(matrix, dtype=np.float64)
        else:
            raise ValueError("EEMMatrix Missing file_path and Matrix, one of either must be provided")
        self.id = EEMMatrix._next_eem_id
        EEMMatrix._next_eem_id += 1

    def _load_and_interpolate(self, file_path):
        raw = _load_table(file_path)
        src_em = raw[0, 1:]
        src_ex = raw[1:, 0]
        z = raw[1:, 1:]
        interp = RectBivariateSpline(src_ex, src_em, z)
        m = interp(self.target_wavelengths, self.target_wavelengths)
        wl = self.target_wavelengths
        m[wl[None, :] <= wl[:, None]] = 0.0
        return np.maximum(0, m)


# ----------------------------------------------------------------------------
# The curve editors' export formats.

def _js_round(x):  # Math.round: halves round up
    return int(np.floor(x + 0.5))


def spec_table_text(wavelengths, curves):
    """The spectrum editor's export (SpecEdit.html:433-467): one line per
    sample, the wavelength rounded to an integer, then one value per curve
    with 5 decimals.  Readable by ``SpecularVec`` (``column`` = curve + 1) and,
    with three curves, by ``ColourSystem`` as a CMF table."""
    wl = np.asarray(wavelengths, dtype=np.float64)
    cv = np.atleast_2d(np.asarray(curves, dtype=np.float64))
    if cv.shape[1] != wl.shape[0]:
        cv = cv.T
    lines = []
    for i, w in enumerate(wl):
        lines.append(" ".join([str(_js_round(w))] + [f"{v:.5f}" for v in cv[:, i]]))
    return "\n".join(lines) + "\n"


def eem_table_text(start, end, excitation, emission, amplitude):
    """The EEM editor's export (EEMEdit.html:464-505) for n samples between
    ``start`` and ``end`` nm: a comment line, a header row of 0 and the
    emission wavelengths, then per excitation wavelength the values
    excitation[i] * emission[j] * amplitude for j > i and 0 otherwise (6
    decimals).  Readable by ``EEMMatrix(file_path=...)``."""
    ex = np.asarray(excitation, dtype=np.float64)
    em = np.asarray(emission, dtype=np.float64)
    n = ex.shape[0]
    if n < 2 or em.shape[0] != n:
        raise ValueError("excitation and emission need the same length >= 2")
    step = (end - start) / (n - 1)
    out = ["# EEM Data: Rows = Excitation, Cols = Emission", " ".join(["0"] + [f"{start + j * step:.2f}"
                                                                            for j in range(n)])]
    for i in range(n):
        row = [f"{start + i * step:.2f}"]
        for j in range(n):
            row.append(f"{(ex[i] * em[j] * amplitude if j > i else 0.0):.6f}")
        out.append(" ".join(row))
    return "\n".join(out) + "\n"


# ----------------------------------------------------------------------------
# Synthetic tables used by the benchmark scenes (no files ship with them).

def wyman_cmf(wavelengths_nm):
    """CIE 1931 2-degree CMFs from the multi-lobe piecewise-Gaussian fit of
    Wyman, Sloan & Shirley, "Simple Analytic Approximations to the CIE XYZ
    Color Matching Functions", JCGT 2(2), 2013 (their equation 2)."""
    lam = np.asarray(wavelengths_nm, dtype=np.float64)

    def g(mu, s1, s2):
        s = np.where(lam < mu, s1, s2)
        return np.exp(-0.5 * ((lam - mu) / s) ** 2)

    x = 1.056 * g(599.8, 37.9, 31.0) + 0.362 * g(442.0, 16.0, 26.7) - 0.065 * g(501.1, 20.4, 26.2)
    y = 0.821 * g(568.8, 46.9, 40.5) + 0.286 * g(530.9, 16.3, 31.1)
    z = 1.217 * g(437.0, 11.8, 36.0) + 0.681 * g(459.0, 26.0, 13.8)
    return np.stack([x, y, z], axis=1)


def wyman_cmf_table(lo=360.0, hi=830.0, step=1.0):
    """A CMF table in the reference's ``lambda x y z`` layout."""
    lam = np.arange(lo, hi + 0.5 * step, step)
    return np.column_stack([lam, wyman_cmf(lam)])


SRGB_PRIMARIES = dict(red=[0.64, 0.33], green=[0.30, 0.60], blue=[0.15, 0.06], white=[0.33242, 0.34743])


def srgb_colour_system(wavelength_min, wavelength_max, spectral_bands):
    """ColourSystem with the reference's primaries (main.py:55-64) over the
    analytic CMF table."""
    return ColourSystem(wyman_cmf_table(), wavelength_min=wavelength_min, wavelength_max=wavelength_max,
                        spectral_bands=spectral_bands, **SRGB_PRIMARIES)
