"""ctypes wrapper of the CPU restatement (oracle/oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the checker and by bench.py's cpu_baseline leg; never by the product.
"""
from __future__ import annotations

import contextlib
import ctypes
import os
from pathlib import Path

import numpy as np

# ORACLE_LIB: another build of the same checker (the sanitizer build of `make asan`)
LIB_PATH = Path(os.environ.get("ORACLE_LIB") or Path(__file__).resolve().parent / "_build" / "liboracle.so")
_vp = ctypes.c_void_p
_LIB = None
_LIBS = {}


def load():
    global _LIB
    if _LIB is None:
        _LIB = _open(LIB_PATH)
    return _LIB


@contextlib.contextmanager
def planted(k):
    """Every oracle call inside the block goes to the variant built with
    planted mistake k (OR_PLANT in oracle.cpp; `make -C oracle plants`)."""
    global _LIB
    saved = load()
    _LIB = _open(Path(__file__).resolve().parent / "_build" / f"liboracle_plant{int(k)}.so")
    try:
        yield
    finally:
        _LIB = saved


def _open(path):
    key = str(path)
    if key in _LIBS:
        return _LIBS[key]
    if not Path(path).exists():
        raise RuntimeError(f"oracle not built: {path} missing (make -C oracle)")
    lib = ctypes.CDLL(key)
    lib.or_scene_create.restype = _vp
    lib.or_scene_create.argtypes = [ctypes.c_int, ctypes.c_float, ctypes.c_float, _vp, _vp, _vp, ctypes.c_int, _vp,
                                    _vp, _vp, ctypes.c_int, _vp, ctypes.c_int, _vp, ctypes.c_int]
    lib.or_scene_destroy.argtypes = [_vp]
    lib.or_scene_light_tree.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp, ctypes.c_int, _vp]
    lib.or_scene_info.argtypes = [_vp, _vp, _vp]
    lib.or_trace.argtypes = [_vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp]
    lib.or_render.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int,
                              ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _vp]
    lib.or_philox.argtypes = [_vp, _vp, _vp]
    lib.or_uniforms.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int,
                                _vp]
    lib.or_fresnel.restype = ctypes.c_float
    lib.or_fresnel.argtypes = [ctypes.c_float] * 5
    lib.or_ggx_D.restype = ctypes.c_float
    lib.or_ggx_D.argtypes = [ctypes.c_float, ctypes.c_float]
    lib.or_ggx_G.restype = ctypes.c_float
    lib.or_ggx_G.argtypes = [_vp, _vp, _vp, ctypes.c_float]
    lib.or_hg_eval.restype = ctypes.c_float
    lib.or_hg_eval.argtypes = [ctypes.c_float, ctypes.c_float]
    lib.or_sample.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_int, _vp, ctypes.c_float, _vp]
    lib.or_shadow.restype = ctypes.c_float
    lib.or_shadow.argtypes = [_vp, _vp, _vp, ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.or_nee.argtypes = [_vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                           ctypes.c_int, _vp]
    lib.or_lt_sample.argtypes = [ctypes.c_int, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp,
                                 _vp]
    lib.or_lt_pdf.argtypes = [ctypes.c_int, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp]
    lib.or_oren_nayar.restype = ctypes.c_float
    lib.or_oren_nayar.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float, _vp, _vp, _vp]
    lib.or_scatter_surface.argtypes = [_vp, ctypes.c_float, _vp, _vp, ctypes.c_uint64, ctypes.c_int, _vp]
    lib.or_eval_bsdf.argtypes = [_vp, ctypes.c_float, ctypes.c_int, _vp, _vp, _vp, _vp]
    lib.or_volume_events.argtypes = [_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_int, _vp,
                                     _vp, _vp, _vp]
    _LIBS[key] = lib
    return lib


def _p(a):
    return a.ctypes.data_as(_vp)


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


class OracleScene:
    """The oracle's view of a FlatScene (reference-shaped binary BVH, or brute force)."""

    def __init__(self, flat, brute=False):
        lib = load()
        self.flat = flat
        self._keep = [flat.tris_v, flat.tris_n, flat.tri_mat, flat.mat_f32, flat.mat_i32, flat.mat_eem, flat.eem,
                      flat.cam]
        self.h = lib.or_scene_create(flat.bands, flat.wl_min, flat.wl_max, _p(flat.tris_v), _p(flat.tris_n),
                                     _p(flat.tri_mat), flat.n_tris, _p(flat.mat_f32), _p(flat.mat_i32),
                                     _p(flat.mat_eem), flat.n_mats, _p(flat.eem), flat.n_eem, _p(flat.cam),
                                     1 if brute else 0)

    def __del__(self):
        if getattr(self, "h", None):
            load().or_scene_destroy(self.h)
            self.h = None

    def light_tree(self, nodes=None, power=None, keys=None):
        """NEE through a light tree (the device's arrays: nodes as raw
        64-byte records, power [B][n] f32, keys {face, leaf} int32 pairs);
        no arguments return to the area-CDF pick."""
        if nodes is None:
            load().or_scene_light_tree(self.h, 0, 0, None, None, 0, None)
            return
        nodes = np.ascontiguousarray(nodes, np.uint8)
        power = _f32(power)
        keys = np.ascontiguousarray(keys, np.int32)
        n = nodes.nbytes // 64
        if load().or_scene_light_tree(self.h, 1, n, _p(nodes), _p(power), keys.size // 2, _p(keys)) != 0:
            raise ValueError("malformed light tree arrays")
        self._tree = (nodes, power, keys)

    def info(self):
        n = np.zeros(1, np.int32)
        m = np.zeros(1, np.int32)
        load().or_scene_info(self.h, _p(n), _p(m))
        return {"nodes": int(n[0]), "max_lights": int(m[0])}

    def trace(self, orig, dirs):
        o = _f32(orig)
        d = _f32(dirs)
        n = o.shape[0]
        t = np.empty(n, np.float32)
        p = np.empty(n, np.int32)
        uv = np.empty((n, 2), np.float32)
        load().or_trace(self.h, n, _p(o), _p(d), _p(t), _p(p), _p(uv))
        return t, p, uv

    def render(self, width, height, seed, first_block, n_blocks, depth, accum=None, cols=None, stream=0):
        """Running-mean accumulator [W, H, B] after blocks [first, first+n)."""
        B = self.flat.bands
        if accum is None:
            accum = np.zeros((width, height, B), np.float32)
        i0, i1 = (0, width) if cols is None else cols
        stats = np.zeros(10, np.float64)
        load().or_render(self.h, width, height, int(seed), int(stream), int(first_block), int(n_blocks), int(depth),
                         int(i0), int(i1), _p(accum), _p(stats))
        return accum, {"segments": int(stats[0]), "traversals": int(stats[1]), "nonfinite": int(stats[2]),
                       "culled_traversals": int(stats[3]), "culled_nonfinite": int(stats[4]),
                       "nf_btdf_eta1": int(stats[5]), "nf_btdf_eta1_last": int(stats[6]),
                       "nf_other_estimator": int(stats[7]), "nf_direct": int(stats[8]), "nf_indirect": int(stats[9])}

    def shadow(self, o, d, max_dist, tri, vol, band):
        return float(load().or_shadow(self.h, _p(_f32(o)), _p(_f32(d)), float(max_dist), int(tri), int(vol),
                                      int(band)))

    def nee(self, pos, dirn, nrm, surf, vol, band, seed, n):
        out = np.zeros(3, np.float64)
        load().or_nee(self.h, _p(_f32(pos)), _p(_f32(dirn)), _p(_f32(nrm)), int(surf), int(vol), int(band), int(seed),
                      int(n), _p(out))
        return out

    def volume_events(self, pos, dirn, vol, band, seed, n):
        tr = np.empty(n, np.float32)
        ac = np.empty(n, np.int32)
        es = np.empty(n, np.float32)
        nb = np.empty(n, np.int32)
        load().or_volume_events(self.h, _p(_f32(pos)), _p(_f32(dirn)), int(vol), int(band), int(seed), int(n), _p(tr),
                                _p(ac), _p(es), _p(nb))
        return tr, ac, es, nb


def lt_sample(nodes, power, band, x, u):
    """Light-tree picks for points x [k, 3] and uniforms u [k]: (leaf, prob)."""
    nodes = np.ascontiguousarray(nodes, np.uint8)
    power = _f32(power)
    x = _f32(x).reshape(-1, 3)
    u = _f32(u).reshape(-1)
    k = x.shape[0]
    leaf = np.empty(k, np.int32)
    prob = np.empty(k, np.float32)
    n = nodes.nbytes // 64
    load().or_lt_sample(n, _p(nodes), _p(power), power.size // n, int(band), k, _p(x), _p(u), _p(leaf), _p(prob))
    return leaf, prob


def lt_pdf(nodes, power, band, x, leaves):
    """Probability that the light-tree pick from x [k, 3] returns leaves [k]."""
    nodes = np.ascontiguousarray(nodes, np.uint8)
    power = _f32(power)
    x = _f32(x).reshape(-1, 3)
    lv = np.ascontiguousarray(leaves, np.int32).reshape(-1)
    k = x.shape[0]
    prob = np.empty(k, np.float32)
    n = nodes.nbytes // 64
    load().or_lt_pdf(n, _p(nodes), _p(power), power.size // n, int(band), k, _p(x), _p(lv), _p(prob))
    return prob


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, np.uint32)
    k = np.ascontiguousarray(key, np.uint32)
    o = np.zeros(4, np.uint32)
    load().or_philox(_p(c), _p(k), _p(o))
    return o


def uniforms(seed, c0, c1, c2, n):
    out = np.empty(n, np.float32)
    load().or_uniforms(int(seed), int(c0), int(c1), int(c2), int(n), _p(out))
    return out


def fresnel(c, ni, ki, nt, kt):
    return float(load().or_fresnel(c, ni, ki, nt, kt))


def ggx_D(nh, r):
    return float(load().or_ggx_D(nh, r))


def ggx_G(n, v, l, r):
    return float(load().or_ggx_G(_p(_f32(n)), _p(_f32(v)), _p(_f32(l)), r))


def hg_eval(c, g):
    return float(load().or_hg_eval(c, g))


def sample(which, seed, n, axis, param=0.0):
    """which: 0 cosine hemisphere, 1 HG, 2 GGX half vector, 3 unit sphere."""
    out = np.empty((n, 3), np.float32)
    load().or_sample(int(which), int(seed), int(n), _p(_f32(axis)), float(param), _p(out))
    return out


def oren_nayar(r, NdotL, NdotV, l, v, n):
    """The Oren-Nayar factor A + B cos+ sin(alpha) tan(beta) shared by
    scatter_surface and evaluate_bsdf."""
    return float(load().or_oren_nayar(float(r), float(NdotL), float(NdotV), _p(_f32(l)), _p(_f32(v)), _p(_f32(n))))


def _mat(n_i, k_i, n_t, k_t, r, ss, tv):
    return np.array([n_i, k_i, n_t, k_t, r, ss, tv], np.float32)


def scatter_surface(mat, wl_nm, dirn, nrm, seed, n):
    """n samples of scatter_surface (probe scene: medium 0 = (n_i, k_i),
    surface 1 = (n_t, k_t, roughness, sigma_s, is_true_volume)); returns
    dict of direction [n,3], estimator, pdf, did_reflect, half vector."""
    out = np.empty((n, 9), np.float32)
    load().or_scatter_surface(_p(_mat(*mat)), float(wl_nm), _p(_f32(dirn)), _p(_f32(nrm)), int(seed), int(n), _p(out))
    return {"dir": out[:, 0:3], "est": out[:, 3], "pdf": out[:, 4], "reflect": out[:, 5] > 0.5, "h": out[:, 6:9]}


def eval_bsdf(mat, wl_nm, V, L, nrm):
    """evaluate_bsdf for (V[i], L[i]) about nrm: (bsdf [n], pdf [n])."""
    V = _f32(V).reshape(-1, 3)
    L = _f32(L).reshape(-1, 3)
    out = np.empty((V.shape[0], 2), np.float32)
    load().or_eval_bsdf(_p(_mat(*mat)), float(wl_nm), V.shape[0], _p(V), _p(L), _p(_f32(nrm)), _p(out))
    return out[:, 0], out[:, 1]
