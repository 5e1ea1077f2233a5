"""Host-side scene compile and spectral loaders: the setup_scene and
data.py semantics, pinned by closed forms and numpy/scipy library calls on
synthetic fixture files."""
import io
import math
import re
from pathlib import Path

import numpy as np
import pytest
from scipy.interpolate import RectBivariateSpline

from spectral import configs, data
from spectral.scene import camera_basis, face_normals_f32, flatten_scene, scene_dimensions

ROOT = Path(__file__).resolve().parents[1]


def test_icosphere_counts_and_normals():
    for sub, nf, nv in [(0, 20, 12), (1, 80, 42), (3, 1280, 642)]:
        m = configs.icosphere([0.1, 0.2, 0.3], 0.25, 5, subdivisions=sub)
        assert len(m["faces"]) == nf and len(m["vertices"]) == nv
        v = np.asarray(m["vertices"], np.float64) - [0.1, 0.2, 0.3]
        assert np.allclose(np.linalg.norm(v, axis=1), 0.25, atol=1e-6)
        n = np.asarray(m["vertex_normals"])
        assert np.allclose(n, v / 0.25, atol=1e-6)


def test_workload_triangle_counts():
    assert flatten_scene(configs.box_icosphere().scene, 25, 400, 700).n_tris == 1292
    c1 = configs.heightfield_1m(nx=64, nz=32)
    assert flatten_scene(c1.scene, 31, 400, 700).n_tris == 12 + 64 * 32 * 2
    c2 = configs.volume_fluor()
    f2 = flatten_scene(c2.scene, 31, 400, 700)
    assert f2.n_tris == 12 + 20480 + 10008
    assert f2.eem.shape == (2, 31, 31)


def test_face_normals_area_closed_form():
    v0 = np.array([[0, 0, 0], [1, 1, 1]], np.float32)
    v1 = np.array([[2, 0, 0], [1, 1, 1]], np.float32)
    v2 = np.array([[0, 3, 0], [1, 1, 1]], np.float32)
    n, a = face_normals_f32(v0, v1, v2)
    assert a[0] == 3.0 and np.array_equal(n[0], [0, 0, 1])
    # degenerate face: area 0, normal left unnormalised (main.py:536)
    assert a[1] == 0.0 and np.array_equal(n[1], [0, 0, 0])


def test_flatten_defaults_and_errors():
    B = 3
    scene = {"materials": [{"surface_emission": [0] * B}, {"surface_emission": [1, 0, 1], "emission": [2.0] * B}],
             "meshes": [{"vertices": [[0, 0, 0], [1, 0, 0], [0, 1, 0]], "faces": [[0, 1, 2]], "material_id": 1}],
             "camera": {"origin": [0, 0, 1], "lookat": [0, 0, 0], "up": [0, 1, 0]}}
    f = flatten_scene(scene, B, 400, 700)
    assert np.array_equal(f.mat_f32[0], np.zeros((6, B)))  # missing keys -> 0 (main.py:548-561)
    assert np.array_equal(f.mat_i32[1, 1], [1, 0, 1])
    assert np.array_equal(f.tris_n[0], np.tile([0, 0, 1], (3, 1)))  # face normal when no vertex normals
    dims = scene_dimensions(scene, B)
    assert dims == {"num_materials": 2, "num_triangles": 1, "num_bvh_nodes": 1, "max_num_lights_per_band": 1}
    bad = dict(scene, materials=[{}, {}])
    with pytest.raises(KeyError):
        flatten_scene(bad, B, 400, 700)
    bad = dict(scene, meshes=[dict(scene["meshes"][0], material_id=5)])
    with pytest.raises(ValueError):
        flatten_scene(bad, B, 400, 700)


def test_camera_basis_matches_update_camera():
    o, u, v, w = camera_basis(np.array([0, 0, 1.75, 0, 0, -0.5, 0, 1, 0], np.float32))
    assert np.allclose(w, [0, 0, 1]) and np.allclose(u, [1, 0, 0]) and np.allclose(v, [0, 1, 0])


YML = """# synthetic refractiveindex.info-style file
DATA:
  - type: formula 2
    coefficients: 0 1 2
  - type: tabulated nk
    data: |
        0.70 1.5 3.5
        0.40 1.2 2.0
        0.55 1.3 2.5
"""


def test_refractive_index_loader(tmp_path):
    p = tmp_path / "m.yml"
    p.write_text(YML)
    ri = data.RefractiveIndex(str(p), 400, 700, 7)
    wl = np.linspace(400, 700, 7)
    assert np.allclose(ri.n, np.interp(wl, [400, 550, 700], [1.2, 1.3, 1.5]))
    assert np.allclose(ri.k, np.interp(wl, [400, 550, 700], [2.0, 2.5, 3.5]))
    missing = data.RefractiveIndex(str(tmp_path / "nope.yml"), 400, 700, 7)
    assert missing.n is None and missing.k is None


def test_colour_system_white_maps_to_unit_rgb():
    cs = data.srgb_colour_system(400, 700, 31)
    white = cs.white / cs.white[1]
    rgb = cs.T @ white
    assert np.allclose(rgb / rgb[0], 1.0, atol=1e-9)
    assert cs.normalization_factor == pytest.approx(np.sum(cs.cmf[:, 1]) * 10.0)
    # analytic CMF fit: y-bar peaks near 555 nm at ~1
    lam = np.arange(380, 781)
    y = data.wyman_cmf(lam)[:, 1]
    assert abs(lam[np.argmax(y)] - 556) <= 3 and 0.95 < y.max() < 1.05


def test_colour_system_from_file_format(tmp_path):
    tab = data.wyman_cmf_table(380, 780, 5)
    p = tmp_path / "cmf.txt"
    np.savetxt(p, tab, fmt="%.6f")
    a = data.ColourSystem(str(p), wavelength_min=400, wavelength_max=700, spectral_bands=25, **data.SRGB_PRIMARIES)
    b = data.ColourSystem(tab, wavelength_min=400, wavelength_max=700, spectral_bands=25, **data.SRGB_PRIMARIES)
    assert np.allclose(a.cmf, b.cmf, atol=1e-6)


def test_eem_loader_stokes_mask_and_spline(tmp_path):
    ex = np.linspace(300, 800, 11)
    em = np.linspace(300, 800, 11)
    z = np.exp(-((ex[:, None] - 450) / 60) ** 2 - ((em[None, :] - 550) / 50) ** 2)
    raw = np.zeros((12, 12))
    raw[0, 1:] = em
    raw[1:, 0] = ex
    raw[1:, 1:] = z
    p = tmp_path / "eem.txt"
    np.savetxt(p, raw)
    e = data.EEMMatrix(400, 700, 7, file_path=str(p))
    wl = np.linspace(400, 700, 7)
    ref = np.maximum(0, RectBivariateSpline(ex, em, z)(wl, wl))
    ref[wl[None, :] <= wl[:, None]] = 0
    assert np.allclose(e.matrix, ref)
    assert np.all(np.tril(e.matrix) == 0)
    e2 = data.EEMMatrix(400, 700, 7, matrix=np.ones((7, 7)))
    assert e2.id == e.id + 1 and np.all(e2.matrix == 1)  # matrix path skips the mask
    with pytest.raises(ValueError):
        data.EEMMatrix(400, 700, 7)


def test_fluorescence_eem_definition():
    wl = np.linspace(400, 700, 31)
    m = configs.fluorescence_eem(wl)
    assert m.max() <= 0.6 + 1e-12
    assert np.all(np.tril(m) == 0)
    i450 = int(np.argmin(abs(wl - 450)))
    assert m[i450].max() == pytest.approx(0.6, rel=1e-6)


def test_albedo_to_k_roundtrip():
    wl = np.linspace(400, 700, 25)
    k = np.asarray(configs.albedo_to_k(np.full(25, 0.7), wl))
    absorb = (4 * math.pi * k) / (wl * 1e-9)
    assert np.allclose(1.0 / (1.0 + absorb + 1e-6), 0.7, rtol=1e-7)
