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


class RefractiveIndex:
    """n(lambda), k(lambda) from a refractiveindex.info YAML file (data.py:6-83)."""

    def __init__(self, file_path, wavelength_min, wavelength_max, spectral_bands):
        self.file_path = file_path
        self.target_wavelengths_nm = band_wavelengths(wavelength_min, wavelength_max, spectral_bands)
        self.n, self.k = self._load_and_interpolate_data()

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
    """A 1-D spectrum table times ``scale`` (data.py:130-143)."""

    def __init__(self, file_path, scale, wavelength_min, wavelength_max, spectral_bands):
        self.raw = _load_table(file_path)
        self.target_wavelengths = band_wavelengths(wavelength_min, wavelength_max, spectral_bands)
        self.vec = self._interpolate(scale)

    def _interpolate(self, scale):
        return (np.interp(self.target_wavelengths, self.raw[:, 0], self.raw[:, 1]) * scale).tolist()


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
