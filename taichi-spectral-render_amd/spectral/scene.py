"""Scene dict -> flat arrays, with the reference's setup_scene semantics.

The scene API is the reference's ``create_scene()`` dict
(Xyzonox/taichi-spectral-render main.py:89-247):

* ``materials[i]``: ``refractive_index``, ``extinction_coefficient``,
  ``roughness``, ``scattering_coefficient``, ``emission``,
  ``anistropy_factor`` (sic) -- per-band floats; ``is_true_volume``,
  ``surface_emission`` -- per-band ints; ``eem_id`` -- int.  Missing keys
  default to zeros (main.py:545-561).  Material 0 is the ambient medium
  camera rays start in (main.py:1523).
* ``meshes[j]``: ``vertices``, ``faces``, ``material_id`` and optional
  ``vertex_normals`` (main.py:518-537).
* ``camera``: ``origin``, ``lookat``, ``up`` (main.py:242-246).

One addition: the reference fills its global ``reemission_matrix`` field
from ``EEMMatrix`` objects outside the dict (main.py:1684).  Here a scene
may carry ``eems`` -- a list of ``EEMMatrix`` objects or a ``{id: matrix}``
dict -- and the table is built from it.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

F32_KEYS = ("refractive_index", "extinction_coefficient", "roughness", "scattering_coefficient", "emission",
            "anistropy_factor")  # main.py:40-45
I32_KEYS = ("is_true_volume", "surface_emission")  # main.py:49-50


@dataclass
class FlatScene:
    """Everything the C ABI's ``sr_upload_scene`` takes, as numpy arrays."""

    bands: int
    wl_min: float
    wl_max: float
    tris_v: np.ndarray  # [N, 3, 3] f32 vertex positions
    tris_n: np.ndarray  # [N, 3, 3] f32 per-vertex shading normals
    tri_mat: np.ndarray  # [N] i32
    mat_f32: np.ndarray  # [M, 6, B] f32
    mat_i32: np.ndarray  # [M, 2, B] i32
    mat_eem: np.ndarray  # [M] i32
    eem: np.ndarray  # [E, B, B] f32; E counts the unused id 0
    cam: np.ndarray  # [9] f32: origin, lookat, up

    @property
    def n_tris(self):
        return int(self.tri_mat.shape[0])

    @property
    def n_mats(self):
        return int(self.mat_eem.shape[0])

    @property
    def n_eem(self):
        return int(self.eem.shape[0])


def face_normals_f32(v0, v1, v2):
    """Per-face normal and area in f32 as main.py:526,535-536 computes them:
    ``normal = cross(v1 - v0, v2 - v0)``, ``area = |normal| / 2`` and the
    normal divided by ``2 * area`` when ``area > 1e-6``."""
    e1 = (v1 - v0).astype(np.float32)
    e2 = (v2 - v0).astype(np.float32)
    n = np.empty_like(e1)
    n[:, 0] = e1[:, 1] * e2[:, 2] - e1[:, 2] * e2[:, 1]
    n[:, 1] = e1[:, 2] * e2[:, 0] - e1[:, 0] * e2[:, 2]
    n[:, 2] = e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0]
    sq = (n[:, 0] * n[:, 0] + n[:, 1] * n[:, 1]) + n[:, 2] * n[:, 2]
    area = np.sqrt(sq).astype(np.float32) / np.float32(2.0)
    ok = area > np.float32(1e-6)
    n[ok] = n[ok] / (area[ok] * np.float32(2.0))[:, None]
    return n, area


def _material_tables(materials, bands):
    M = len(materials)
    mf = np.zeros((M, 6, bands), np.float32)
    mi = np.zeros((M, 2, bands), np.int32)
    me = np.zeros((M,), np.int32)
    for i, m in enumerate(materials):
        for p, key in enumerate(F32_KEYS):
            vals = m.get(key, None)
            if vals is not None:
                mf[i, p, :] = np.asarray(vals, dtype=np.float64)[:bands]
        for p, key in enumerate(I32_KEYS):
            vals = m.get(key, None)
            if vals is not None:
                mi[i, p, :] = np.asarray(vals)[:bands].astype(np.int32)
        me[i] = int(m.get("eem_id", 0))
    return mf, mi, me


def _eem_table(scene, bands, max_id):
    eems = scene.get("eems", None)
    entries = {}
    if eems is None:
        pass
    elif isinstance(eems, dict):
        entries = {int(k): np.asarray(v, np.float64) for k, v in eems.items()}
    else:
        for e in eems:
            entries[int(e.id)] = np.asarray(e.matrix, np.float64)
    E = max([max_id] + list(entries.keys())) + 1
    tab = np.zeros((E, bands, bands), np.float32)
    for k, v in entries.items():
        if v.shape != (bands, bands):
            raise ValueError(f"EEM {k} has shape {v.shape}, expected {(bands, bands)}")
        tab[k] = v
    return tab


def scene_dimensions(scene, bands, verbose=False):
    """get_scene_dimensions, main.py:398-435."""
    mats = scene.get("materials", [])
    n_tris = sum(len(m.get("faces", [])) for m in scene.get("meshes", []))
    lights = [0] * bands
    for mesh in scene.get("meshes", []):
        se = mats[mesh["material_id"]].get("surface_emission", [False] * bands)
        if any(se):
            for i in range(bands):
                if se[i]:
                    lights[i] += len(mesh.get("faces", []))
    dims = {"num_materials": len(mats), "num_triangles": n_tris, "num_bvh_nodes": max(0, 2 * n_tris - 1),
            "max_num_lights_per_band": max(lights) if lights else 0}
    if verbose:
        print("--- Scene Dimensions ---")
        print(f"Materials: {dims['num_materials']}")
        print(f"Triangles: {dims['num_triangles']}")
        print(f"Max BVH Nodes: {dims['num_bvh_nodes']}")
        print(f"Max Light Tris in any band: {dims['max_num_lights_per_band']}")
        print("------------------------")
    return dims


def flatten_scene(scene, bands, wl_min, wl_max):
    """The host half of setup_scene (main.py:509-561, 629-631): triangles with
    per-vertex shading normals (the face normal when a mesh has none,
    main.py:527-534), material tables with zero defaults, EEM table, camera.

    BVH, per-face normals/areas and the per-band light CDF are built by the
    library on upload (``sr_upload_scene``), with the rules of
    main.py:526-536 and 590-627.
    """
    materials = scene["materials"]
    vs, ns, ms = [], [], []
    for mesh in scene["meshes"]:
        verts = np.asarray(mesh["vertices"], dtype=np.float32).reshape(-1, 3)
        faces = np.asarray(mesh["faces"], dtype=np.int64).reshape(-1, 3)
        mat_id = int(mesh["material_id"])
        if not 0 <= mat_id < len(materials):
            raise ValueError(f"mesh material_id {mat_id} out of range [0, {len(materials)})")
        # main.py:596 indexes surface_emission directly: a material without
        # it cannot carry triangles in the reference (KeyError there too).
        if "surface_emission" not in materials[mat_id]:
            raise KeyError("surface_emission")
        if faces.size and (faces.min() < 0 or faces.max() >= len(verts)):
            raise ValueError("face index out of range")
        tv = verts[faces]  # [F, 3, 3]
        if "vertex_normals" in mesh:
            vn = np.asarray(mesh["vertex_normals"], dtype=np.float32).reshape(-1, 3)
            tn = vn[faces]
        else:
            fn, _ = face_normals_f32(tv[:, 0], tv[:, 1], tv[:, 2])
            tn = np.repeat(fn[:, None, :], 3, axis=1)
        vs.append(tv)
        ns.append(tn)
        ms.append(np.full((len(faces),), mat_id, np.int32))
    tris_v = np.ascontiguousarray(np.concatenate(vs) if vs else np.zeros((0, 3, 3), np.float32), dtype=np.float32)
    tris_n = np.ascontiguousarray(np.concatenate(ns) if ns else np.zeros((0, 3, 3), np.float32), dtype=np.float32)
    tri_mat = np.ascontiguousarray(np.concatenate(ms) if ms else np.zeros((0,), np.int32), dtype=np.int32)
    mf, mi, me = _material_tables(materials, bands)
    eem = _eem_table(scene, bands, int(me.max()) if len(me) else 0)
    cam = scene["camera"]
    cam_arr = np.array(list(cam["origin"]) + list(cam["lookat"]) + list(cam["up"]), np.float32)
    return FlatScene(bands, float(wl_min), float(wl_max), tris_v, tris_n, tri_mat, mf, mi, me, eem, cam_arr)


def camera_basis(cam):
    """update_camera, main.py:633-640, in f32."""
    o = np.asarray(cam[0:3], np.float32)
    la = np.asarray(cam[3:6], np.float32)
    up = np.asarray(cam[6:9], np.float32)

    def nrm(v):
        return (v / np.float32(np.sqrt(np.float32((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2])))).astype(np.float32)

    w = nrm(o - la)
    u = nrm(np.cross(up, w).astype(np.float32))
    v = np.cross(w, u).astype(np.float32)
    return o, u, v, w


def pixel_center_rays(flat: FlatScene, width, height):
    """Deterministic primary rays through pixel centres, as visualize_normals
    generates them (main.py:1633-1641).  Returns (orig[n,3], dir[n,3]) in
    the accumulator's (i, j) order, i.e. ray index i * height + j."""
    o, u, v, w = camera_basis(flat.cam)
    i = np.arange(width, dtype=np.float32)
    j = np.arange(height, dtype=np.float32)
    ii, jj = np.meshgrid(i, j, indexing="ij")
    uu = (ii + np.float32(0.5)) / np.float32(width)
    vv = (jj + np.float32(0.5)) / np.float32(height)
    aspect = np.float32(width / height)
    d = (u[None, None, :] * ((uu * 2 - 1) * aspect)[..., None] + v[None, None, :] * (vv * 2 - 1)[..., None]
         - w[None, None, :] * np.float32(2.0)).astype(np.float32)
    d = d / np.sqrt((d * d).sum(-1, keepdims=True)).astype(np.float32)
    d = d.reshape(-1, 3).astype(np.float32)
    orig = np.broadcast_to(o, d.shape).astype(np.float32).copy()
    return orig, np.ascontiguousarray(d)
