"""Frozen synthetic scenes for the five benchmark workloads.

Every scene is a reference-style ``create_scene()`` dict (materials, meshes,
camera; Xyzonox/taichi-spectral-render main.py:89-247).  The geometry is the
reference's Cornell box (main.py:179-192) scaled by 1/2 to a unit box, with
the reference camera (main.py:242-246) scaled the same way.  All tables are
synthetic, so no data files are needed.

Workloads (render = width x height, spp = spectral blocks, depth = path loop
bound):

* ``box_icosphere``   1,292 tris, 25 bands 400-700 nm, 256x256, 16 spp,
  depth 8: walls (Oren-Nayar, 2 step spectra), ceiling quad light with flat
  emission 10, GGX dielectric icosphere (r = 0.25, 3 subdivisions, Cauchy n).
* ``heightfield_1m``  1,048,588 tris, 31 bands, 1024x1024, 64 spp, depth 12:
  the box with a 1024x512-cell fBm heightfield conductor instead of the
  sphere.
* ``volume_fluor``    31 bands, 512x512, 128 spp, depth 16: the box with a
  Henyey-Greenstein medium sphere (5 subdivisions) and a fluorescent slab
  carrying a 31x31 EEM.
* ``batch64``         64 random scenes (displaced icospheres + box).
* ``inverse_volume``  ``volume_fluor`` at 256x256, 64 spp, depth 12, with the
  126 differentiable parameters named.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .data import EEMMatrix, band_wavelengths

BOX_SCALE = 0.5


@dataclass
class RenderConfig:
    name: str
    scene: dict
    bands: int
    wl_min: float
    wl_max: float
    width: int
    height: int
    spp: int
    depth: int
    seed: int
    extra: dict = field(default_factory=dict)


# ---------------------------------------------------------------- materials

def albedo_to_k(albedo, wavelengths_nm, sigma_s=1.0):
    """Extinction coefficient k giving Oren-Nayar diffuse albedo ``albedo``
    under the reference's mapping albedo = s / (s + 4 pi k / lambda[m] + 1e-6)
    (main.py:1057, 1070)."""
    a = np.asarray(albedo, np.float64)
    lam_m = np.asarray(wavelengths_nm, np.float64) * 1e-9
    return ((sigma_s * (1.0 - a) / a - 1e-6) * lam_m / (4.0 * math.pi)).tolist()


def sigma_to_k(sigma_a, wavelengths_nm):
    """k giving absorption coefficient sigma_a = 4 pi k / lambda[m]."""
    lam_m = np.asarray(wavelengths_nm, np.float64) * 1e-9
    return (np.asarray(sigma_a, np.float64) * lam_m / (4.0 * math.pi)).tolist()


def _mat(B, n=1.0, k=0.0, r=0.0, tv=0, ss=0.0, g=0.0, se=0, em=0.0, eem_id=0):
    def vec(x):
        return list(np.broadcast_to(np.asarray(x, np.float64), (B,)).tolist())

    return {
        "refractive_index": vec(n),
        "extinction_coefficient": vec(k),
        "roughness": vec(r),
        "is_true_volume": [int(v) for v in np.broadcast_to(np.asarray(tv), (B,))],
        "scattering_coefficient": vec(ss),
        "anistropy_factor": vec(g),
        "surface_emission": [int(v) for v in np.broadcast_to(np.asarray(se), (B,))],
        "emission": vec(em),
        "eem_id": int(eem_id),
    }


WALL_ROUGHNESS = 20.0 / 90.0  # Oren-Nayar sigma = 20 deg with sigma = r * pi / 2 (main.py:1071)


def box_materials(B, wl):
    """0 air, 1 white (0.7), 2 red step, 3 green step, 4 light (emission 10)."""
    red = np.where(wl >= 600.0, 0.8, 0.1)
    green = np.where((wl >= 500.0) & (wl <= 570.0), 0.8, 0.1)
    return [
        _mat(B, n=1.0, k=0.0, r=0.0, tv=1, ss=0.0, g=0.8),  # air, as main.py:91-101
        _mat(B, n=1.4, k=albedo_to_k(np.full(B, 0.7), wl), r=WALL_ROUGHNESS, tv=0, ss=1.0),
        _mat(B, n=1.4, k=albedo_to_k(red, wl), r=WALL_ROUGHNESS, tv=0, ss=1.0),
        _mat(B, n=1.4, k=albedo_to_k(green, wl), r=WALL_ROUGHNESS, tv=0, ss=1.0),
        _mat(B, n=1.0, k=0.0, r=0.0, tv=0, ss=0.0, se=1, em=10.0),
    ]


# ---------------------------------------------------------------- geometry

def _quad(verts, mat):
    s = BOX_SCALE
    return {"vertices": [[c * s for c in v] for v in verts], "faces": [[0, 2, 1], [0, 3, 2]], "material_id": mat}


def box_meshes(white=1, left=2, right=3, light=4):
    """Walls and light of main.py:179-192 (same winding), scaled to a unit box."""
    return [
        _quad([[-1, -1, -1], [1, -1, -1], [1, -1, 1], [-1, -1, 1]], white),  # floor
        _quad([[-1, 1, -1], [-1, 1, 1], [1, 1, 1], [1, 1, -1]], white),  # ceiling
        _quad([[-1, -1, -1], [-1, 1, -1], [1, 1, -1], [1, -1, -1]], white),  # back
        _quad([[1, -1, -1], [1, 1, -1], [1, 1, 1], [1, -1, 1]], right),  # right
        _quad([[-1, -1, -1], [-1, -1, 1], [-1, 1, 1], [-1, 1, -1]], left),  # left
        _quad([[-0.3, 0.99, -0.3], [-0.3, 0.99, 0.3], [0.3, 0.99, 0.3], [0.3, 0.99, -0.3]], light),
    ]


def box_camera():
    s = BOX_SCALE
    return {"origin": [0.0, 0.0, 3.5 * s], "lookat": [0.0, 0.0, -1.0 * s], "up": [0.0, 1.0, 0.0]}


_ICO_T = (1.0 + 5.0 ** 0.5) / 2.0
_ICO_V = [[-1, _ICO_T, 0], [1, _ICO_T, 0], [-1, -_ICO_T, 0], [1, -_ICO_T, 0], [0, -1, _ICO_T], [0, 1, _ICO_T],
          [0, -1, -_ICO_T], [0, 1, -_ICO_T], [_ICO_T, 0, -1], [_ICO_T, 0, 1], [-_ICO_T, 0, -1], [-_ICO_T, 0, 1]]
_ICO_F = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6],
          [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10],
          [8, 6, 7], [9, 8, 1]]


_UNIT_ICO = {}


def _unit_icosphere(subdivisions):
    """Unit-sphere vertices and faces of create_sphere's subdivision
    (main.py:209-233), memoised per level."""
    if subdivisions in _UNIT_ICO:
        return _UNIT_ICO[subdivisions]
    verts = [np.array(v, np.float32) for v in _ICO_V]
    verts = [v / np.float32(np.linalg.norm(v)) for v in verts]
    faces = [list(f) for f in _ICO_F]
    for _ in range(subdivisions):
        new_faces = []
        mid = {}
        for f in faces:
            v1, v2, v3 = f
            m12 = tuple(sorted((v1, v2)))
            m23 = tuple(sorted((v2, v3)))
            m31 = tuple(sorted((v3, v1)))
            for m in (m12, m23, m31):
                if m not in mid:
                    mp = (verts[m[0]] + verts[m[1]]) / np.float32(2.0)
                    mid[m] = len(verts)
                    verts.append((mp / np.float32(np.linalg.norm(mp))).astype(np.float32))
            new_faces.append([v1, mid[m12], mid[m31]])
            new_faces.append([v2, mid[m23], mid[m12]])
            new_faces.append([v3, mid[m31], mid[m23]])
            new_faces.append([mid[m12], mid[m23], mid[m31]])
        faces = new_faces
    out = (np.stack(verts).astype(np.float32), np.asarray(faces, np.int64))
    _UNIT_ICO[subdivisions] = out
    return out


def icosphere(center, radius, material_id, subdivisions=2, displace=None):
    """create_sphere, main.py:195-237: f32 icosahedron, midpoint subdivision
    with a shared-edge midpoint cache, vertex normals = unit positions.
    ``displace`` (optional) maps unit vertex directions [V,3] to radial
    scale factors (used by the batch scenes)."""
    V, F = _unit_icosphere(subdivisions)
    normals = V / np.linalg.norm(V, axis=1, keepdims=True)
    scale = np.float32(radius)
    if displace is not None:
        scale = (np.float32(radius) * np.asarray(displace(V), np.float32))[:, None]
    final = V * scale + np.asarray(center, np.float32)
    out = {"vertices": final, "faces": F.copy(), "material_id": material_id}
    if displace is None:
        out["vertex_normals"] = normals.astype(np.float32)
    return out


# fBm heightfield: hashed lattice value noise, quintic fade, 6 octaves,
# lacunarity 2, gain 0.5, base frequency 4 per unit, normalised to [0, 1).
_M32 = np.uint64(0xFFFFFFFF)


def _hash01(ix, iz, seed):
    """lowbias32 integer hash (C. Wellons) of a lattice point -> [0, 1)."""
    ix = ix.astype(np.int64).astype(np.uint64) & _M32
    iz = iz.astype(np.int64).astype(np.uint64) & _M32
    h = (ix * np.uint64(0x8DA6B343)) ^ (iz * np.uint64(0xD8163841)) ^ np.uint64((seed * 0xCB1AB31F) & 0xFFFFFFFF)
    h &= _M32
    h ^= h >> np.uint64(16)
    h = (h * np.uint64(0x7FEB352D)) & _M32
    h ^= h >> np.uint64(15)
    h = (h * np.uint64(0x846CA68B)) & _M32
    h ^= h >> np.uint64(16)
    return h.astype(np.float64) / 4294967296.0


def value_noise(x, z, seed):
    ix = np.floor(x)
    iz = np.floor(z)
    fx = x - ix
    fz = z - iz
    ux = fx * fx * fx * (fx * (fx * 6 - 15) + 10)
    uz = fz * fz * fz * (fz * (fz * 6 - 15) + 10)
    a = _hash01(ix, iz, seed)
    b = _hash01(ix + 1, iz, seed)
    c = _hash01(ix, iz + 1, seed)
    d = _hash01(ix + 1, iz + 1, seed)
    return (a * (1 - ux) + b * ux) * (1 - uz) + (c * (1 - ux) + d * ux) * uz


def fbm(x, z, seed, octaves=6, lacunarity=2.0, gain=0.5, base_freq=4.0):
    total = np.zeros_like(x, dtype=np.float64)
    norm = 0.0
    amp = 1.0
    f = base_freq
    for o in range(octaves):
        total += amp * value_noise(x * f, z * f, seed * 1000 + o)
        norm += amp
        amp *= gain
        f *= lacunarity
    return total / norm


def heightfield(nx, nz, amplitude, seed, material_id, base_y=-0.5 + 0.002):
    """nx x nz cells over the unit box floor, two triangles per cell, face
    normals pointing +y (no vertex normals: flat shading)."""
    xs = np.linspace(-0.5, 0.5, nx + 1)
    zs = np.linspace(-0.5, 0.5, nz + 1)
    X, Z = np.meshgrid(xs, zs, indexing="ij")
    Y = base_y + amplitude * fbm(X, Z, seed)
    V = np.stack([X, Y, Z], -1).reshape(-1, 3).astype(np.float32)
    a = np.arange(nx)[:, None]
    c = np.arange(nz)[None, :]
    p00 = (a * (nz + 1) + c).ravel()
    p01 = (a * (nz + 1) + c + 1).ravel()
    p10 = ((a + 1) * (nz + 1) + c).ravel()
    p11 = ((a + 1) * (nz + 1) + c + 1).ravel()
    f1 = np.stack([p00, p01, p11], 1)
    f2 = np.stack([p00, p11, p10], 1)
    F = np.stack([f1, f2], 1).reshape(-1, 3)
    return {"vertices": V, "faces": F, "material_id": material_id}


def slab(lo, hi, n, material_id):
    """Closed axis-aligned box with n x n gridded top and bottom faces and
    one quad per side; outward winding."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    xs = np.linspace(x0, x1, n + 1)
    zs = np.linspace(z0, z1, n + 1)
    X, Z = np.meshgrid(xs, zs, indexing="ij")
    verts, faces = [], []

    def grid(y, up):
        base = sum(len(v) for v in verts)
        V = np.stack([X, np.full_like(X, y), Z], -1).reshape(-1, 3)
        a = np.arange(n)[:, None]
        c = np.arange(n)[None, :]
        p00 = (a * (n + 1) + c).ravel() + base
        p01 = (a * (n + 1) + c + 1).ravel() + base
        p10 = ((a + 1) * (n + 1) + c).ravel() + base
        p11 = ((a + 1) * (n + 1) + c + 1).ravel() + base
        if up:
            F = np.stack([np.stack([p00, p01, p11], 1), np.stack([p00, p11, p10], 1)], 1).reshape(-1, 3)
        else:
            F = np.stack([np.stack([p00, p11, p01], 1), np.stack([p00, p10, p11], 1)], 1).reshape(-1, 3)
        verts.append(V)
        faces.append(F)

    grid(y1, True)
    grid(y0, False)
    # sides: quads (a, b, c, d) counter-clockwise seen from outside
    sides = [
        [[x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]],  # +z
        [[x1, y0, z0], [x0, y0, z0], [x0, y1, z0], [x1, y1, z0]],  # -z
        [[x1, y0, z1], [x1, y0, z0], [x1, y1, z0], [x1, y1, z1]],  # +x
        [[x0, y0, z0], [x0, y0, z1], [x0, y1, z1], [x0, y1, z0]],  # -x
    ]
    for q in sides:
        base = sum(len(v) for v in verts)
        verts.append(np.asarray(q, np.float64))
        faces.append(np.array([[base, base + 1, base + 2], [base, base + 2, base + 3]]))
    return {"vertices": np.concatenate(verts).astype(np.float32), "faces": np.concatenate(faces).astype(np.int64),
            "material_id": material_id}


# ---------------------------------------------------------------- workloads

def box_icosphere(bands=25, wl_min=400.0, wl_max=700.0, width=256, height=256, spp=16, depth=8, seed=0):
    wl = band_wavelengths(wl_min, wl_max, bands)
    mats = box_materials(bands, wl)
    cauchy = 1.5 + 0.0042 / (wl * 1e-3) ** 2
    mats.append(_mat(bands, n=cauchy, k=0.0, r=0.05, tv=1, ss=0.0))  # 5: GGX dielectric
    meshes = box_meshes()
    meshes.append(icosphere([0.0, -0.245, -0.05], 0.25, 5, subdivisions=3))
    scene = {"materials": mats, "meshes": meshes, "camera": box_camera()}
    return RenderConfig("box_icosphere", scene, bands, wl_min, wl_max, width, height, spp, depth, seed)


def heightfield_1m(bands=31, wl_min=400.0, wl_max=700.0, width=1024, height=1024, spp=64, depth=12, seed=1,
                   nx=1024, nz=512):
    wl = band_wavelengths(wl_min, wl_max, bands)
    t = (wl - 400.0) / 300.0
    s = t * t * (3.0 - 2.0 * t)
    mats = box_materials(bands, wl)
    mats.append(_mat(bands, n=0.2 + 1.3 * s, k=1.8 + 1.8 * s, r=0.2, tv=1, ss=0.0))  # 5: conductor
    meshes = box_meshes()
    meshes.append(heightfield(nx, nz, 0.1, seed, 5))
    scene = {"materials": mats, "meshes": meshes, "camera": box_camera()}
    return RenderConfig("heightfield_1m", scene, bands, wl_min, wl_max, width, height, spp, depth, seed)


MEDIUM_SIGMA_A_PEAK = 1.5
MEDIUM_SIGMA_A_WIDTH = 40.0  # nm, Gaussian width of the medium's absorption bump
SLAB_ROUGHNESS = 0.1


def gaussian(wl, mu, sd):
    return np.exp(-0.5 * ((np.asarray(wl, np.float64) - mu) / sd) ** 2)


def fluorescence_eem(wl):
    """EEM[i, j] = 0.6 G_ex(l_i; 450, 25) G_em(l_j; 520, 20) / max G_em,
    zero unless l_j > l_i (Stokes shift, as data.py:176-180 masks)."""
    gex = gaussian(wl, 450.0, 25.0)
    gem = gaussian(wl, 520.0, 20.0)
    m = 0.6 * gex[:, None] * gem[None, :] / gem.max()
    m[wl[None, :] <= wl[:, None]] = 0.0
    return m


def volume_fluor(bands=31, wl_min=400.0, wl_max=700.0, width=512, height=512, spp=128, depth=16, seed=2,
                 params=None):
    """``params`` (optional) overrides the differentiable quantities: dict
    with 'albedo' [3,B], 'roughness', 'g', 'sigma_a' [B]."""
    wl = band_wavelengths(wl_min, wl_max, bands)
    p = default_inverse_params(bands, wl_min, wl_max) if params is None else params
    mats = box_materials(bands, wl)
    for i in range(3):
        mats[1 + i]["extinction_coefficient"] = albedo_to_k(np.asarray(p["albedo"][i]), wl)
    # 5: HG medium, index-matched boundary
    mats.append(_mat(bands, n=1.0, k=sigma_to_k(p["sigma_a"], wl), r=0.0, tv=1, ss=2.0, g=p["g"]))
    # 6: fluorescent slab (volume fluorescence, EEM id 1)
    slab_sigma = 5.0 + 30.0 * gaussian(wl, 450.0, 25.0)
    mats.append(_mat(bands, n=1.5, k=sigma_to_k(slab_sigma, wl), r=p["roughness"], tv=1, ss=0.0, eem_id=1))
    meshes = box_meshes()
    meshes.append(icosphere([-0.2, -0.2, -0.1], 0.22, 5, subdivisions=5))
    meshes.append(slab([0.02, -0.5 + 0.001, -0.35], [0.42, -0.5 + 0.08, 0.15], 50, 6))
    scene = {"materials": mats, "meshes": meshes, "camera": box_camera(), "eems": {1: fluorescence_eem(wl)}}
    return RenderConfig("volume_fluor", scene, bands, wl_min, wl_max, width, height, spp, depth, seed)


def default_inverse_params(bands=31, wl_min=400.0, wl_max=700.0):
    wl = band_wavelengths(wl_min, wl_max, bands)
    red = np.where(wl >= 600.0, 0.8, 0.1)
    green = np.where((wl >= 500.0) & (wl <= 570.0), 0.8, 0.1)
    return {
        "albedo": np.stack([np.full(bands, 0.7), red, green]),
        "roughness": SLAB_ROUGHNESS,
        "g": 0.8,
        "sigma_a": MEDIUM_SIGMA_A_PEAK * gaussian(wl, 450.0, MEDIUM_SIGMA_A_WIDTH),
    }


def inverse_volume(bands=31, width=256, height=256, spp=64, depth=12, seed=3, params=None):
    cfg = volume_fluor(bands=bands, width=width, height=height, spp=spp, depth=depth, seed=seed, params=params)
    cfg.name = "inverse_volume"
    cfg.extra["param_materials"] = {"albedo": [1, 2, 3], "roughness": 6, "g": 5, "sigma_a": 5}
    return cfg


def batch64(n_scenes=64, seed=2, bands=31, wl_min=400.0, wl_max=700.0, tri_lo=1_000, tri_hi=500_000):
    """64 scenes: a displaced icosphere (tris log-uniform in [1k, 500k],
    realised with the closest subdivision level and a ring of copies) in the
    box; material uniform in {Oren-Nayar, GGX dielectric, GGX conductor};
    resolution uniform in {128, 256, 512, 1024}^2; spp uniform in
    {8, 16, 32, 64}."""
    rng = np.random.default_rng(seed)
    wl = band_wavelengths(wl_min, wl_max, bands)
    out = []
    for s in range(n_scenes):
        target = float(np.exp(rng.uniform(np.log(tri_lo), np.log(tri_hi))))
        kind = int(rng.integers(0, 3))
        res = int(rng.choice([128, 256, 512, 1024]))
        spp = int(rng.choice([8, 16, 32, 64]))
        sub = int(np.clip(np.round(np.log(max(target - 12, 20) / 20.0) / np.log(4.0)), 0, 7))
        sub = min(sub, 6)
        per = 20 * 4 ** sub
        copies = max(1, int(round(target / per)))
        mats = box_materials(bands, wl)
        if kind == 0:
            mats.append(_mat(bands, n=1.4, k=albedo_to_k(np.full(bands, 0.6), wl), r=0.4, tv=0, ss=1.0))
        elif kind == 1:
            mats.append(_mat(bands, n=1.5 + 0.0042 / (wl * 1e-3) ** 2, k=0.0, r=0.08, tv=1, ss=0.0))
        else:
            t = (wl - 400.0) / 300.0
            mats.append(_mat(bands, n=0.3 + t, k=2.0 + t, r=0.25, tv=1, ss=0.0))
        meshes = box_meshes()
        amp = float(rng.uniform(0.05, 0.25))
        freq = float(rng.uniform(2.0, 6.0))
        phase = rng.uniform(0, 2 * np.pi, 3)

        def disp(V, amp=amp, freq=freq, phase=phase):
            return 1.0 + amp * np.sin(freq * V[:, 0] + phase[0]) * np.sin(freq * V[:, 1] + phase[1]) * np.sin(
                freq * V[:, 2] + phase[2])

        radius = 0.18 if copies == 1 else 0.3 / math.sqrt(copies)
        for c in range(copies):
            ang = 2 * np.pi * c / copies
            rr = 0.0 if copies == 1 else 0.25
            center = [rr * math.cos(ang), -0.5 + radius * 1.3 + 0.02, rr * math.sin(ang) - 0.05]
            meshes.append(icosphere(center, radius, 5, subdivisions=sub, displace=disp))
        scene = {"materials": mats, "meshes": meshes, "camera": box_camera()}
        out.append(RenderConfig(f"batch{s:02d}", scene, bands, wl_min, wl_max, res, res, spp, 12, seed * 1000 + s,
                                extra={"kind": kind, "target_tris": target}))
    return out


def closed_box(lo, hi, material_id, outward=True):
    """A closed axis-aligned box mesh, every triangle's normal pointing
    outward (or inward, for an enclosing emitter)."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    v = [[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0], [x0, y0, z1], [x1, y0, z1], [x1, y1, z1],
         [x0, y1, z1]]
    quads = [(0, 1, 2, 3), (4, 5, 6, 7), (0, 1, 5, 4), (3, 2, 6, 7), (0, 3, 7, 4), (1, 2, 6, 5)]
    c = np.mean(v, axis=0)
    faces = []
    for q in quads:
        for tri in ((q[0], q[1], q[2]), (q[0], q[2], q[3])):
            p = np.array([v[i] for i in tri], np.float64)
            n = np.cross(p[1] - p[0], p[2] - p[0])
            out = np.dot(n, p.mean(0) - c) > 0
            faces.append(list(tri) if out == outward else [tri[0], tri[2], tri[1]])
    return {"vertices": v, "faces": faces, "material_id": material_id}


def emitting_sphere_material(bands, wl_min, wl_max):
    """The box scene's icosphere (material 5) made an emitter in bands >= 8
    only: a rough n = 1.4 opaque surface of albedo 0.5 with emission 3, so
    its 1280 faces are lights in those bands and the light lists, CDFs and
    light-tree powers differ by band."""
    se = (np.arange(bands) >= 8).astype(int)
    k = albedo_to_k(np.full(bands, 0.5), band_wavelengths(wl_min, wl_max, bands))
    return _mat(bands, n=1.4, k=k, r=0.2, tv=0, ss=1.0, se=se, em=3.0)


def lamps(bands=25, wl_min=400.0, wl_max=700.0, width=32, height=32, spp=16, depth=4, seed=1, n=8, h=0.3,
          size=0.1, extent=3.5, floor_roughness=None):
    """Many-light scene: n x n small downward-facing lamps (emission 10) at
    height h, spread evenly over an extent x extent square, above a diffuse
    floor 0.5 wider whose index matches the air (Fresnel 0), camera looking
    down at it.  With a large extent most lamps are far from any shading
    point (the sparse many-light case).  floor_roughness (default: the
    walls' roughness) 0 makes the floor Lambertian."""
    wl = band_wavelengths(wl_min, wl_max, bands)
    mats = [_mat(bands, n=1.0, tv=1),
            _mat(bands, n=1.0, k=albedo_to_k(np.full(bands, 0.7), wl),
                 r=WALL_ROUGHNESS if floor_roughness is None else floor_roughness, tv=0, ss=1.0),
            _mat(bands, n=1.0, se=1, em=10.0)]
    f = 0.5 * extent + 0.25
    meshes = [{"vertices": [[-f, 0, -f], [-f, 0, f], [f, 0, f], [f, 0, -f]], "faces": [[0, 1, 2], [0, 2, 3]],
               "material_id": 1}]
    for i in range(n):
        for j in range(n):
            x0 = -0.5 * extent + extent * (i + 0.5) / n - size / 2
            z0 = -0.5 * extent + extent * (j + 0.5) / n - size / 2
            meshes.append({"vertices": [[x0, h, z0], [x0 + size, h, z0], [x0 + size, h, z0 + size],
                                        [x0, h, z0 + size]], "faces": [[0, 1, 2], [0, 2, 3]], "material_id": 2})
    scene = {"materials": mats, "meshes": meshes,
             "camera": {"origin": [0.0, 2.5, 2.5], "lookat": [0.0, 0.0, 0.0], "up": [0.0, 1.0, 0.0]}}
    return RenderConfig("lamps", scene, bands, wl_min, wl_max, width, height, spp, depth, seed)


def subsurface_slab(w=0.9, sigma_s=5.0, bands=8, wl_min=400.0, wl_max=700.0, width=16, height=16, spp=256, depth=5,
                    seed=4, body=None):
    """A closed 20 x 20 x 5 slab (or `body`) of an opaque, index-matched,
    isotropically scattering material with single albedo w (sigma_a =
    sigma_s (1 - w) / w entered as k = sigma_a lambda / (4 pi)) inside a
    100-unit box that emits radiance 1 inward: the semi-infinite-medium
    setting of the subsurface walk."""
    wl = band_wavelengths(wl_min, wl_max, bands)
    sigma_a = sigma_s * (1.0 - w) / w
    k = sigma_a * wl * 1e-9 / (4.0 * math.pi)
    mats = [_mat(bands, n=1.0, tv=1),
            _mat(bands, n=1.0, k=k, r=0.0, tv=0, ss=sigma_s, g=0.0),
            _mat(bands, n=1.0, se=1, em=1.0)]
    if body is None:
        body = closed_box([-10, -5, -10], [10, 0, 10], 1)
    meshes = [body, closed_box([-50, -50, -50], [50, 50, 50], 2, outward=False)]
    cam = {"origin": [0.0, 3.0, 3.0], "lookat": [0.0, 0.0, 0.0], "up": [0.0, 1.0, 0.0]}
    scene = {"materials": mats, "meshes": meshes, "camera": cam}
    return RenderConfig("subsurface_slab", scene, bands, wl_min, wl_max, width, height, spp, depth, seed)


WORKLOADS = {
    "box_icosphere": box_icosphere,
    "heightfield_1m": heightfield_1m,
    "volume_fluor": volume_fluor,
    "inverse_volume": inverse_volume,
}
