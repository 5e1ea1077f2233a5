"""Host-side mirror of the reference's render protocol on the HIP library.

Xyzonox/taichi-spectral-render drives its kernels from a GUI loop
(main.py:1826-1837): ``cast_rays(band, block)`` once per band,
``average_and_reset_block(n)`` and ``color_rays()`` once per block of B
launches, ``clear_pixel_space()`` and ``update_camera()`` on a camera change
(main.py:1812-1823).  ``Renderer`` keeps those operations and names, but one
``render(n_blocks)`` call runs whole blocks (all B bands of every pixel) in a
single persistent launch and folds them into the running mean.

Device memory is held as torch tensors on the current HIP device; torch is
only the allocator and stream provider here.
"""
from __future__ import annotations

import ctypes
import hashlib
import json

import numpy as np
import torch

from . import _lib
from .scene import FlatScene, flatten_scene, pixel_center_rays, scene_dimensions

# SR_RENDER_* flags of include/spectral_render.h
SR_RENDER_LIGHT_TREE = 1
SR_RENDER_SUBSURFACE = 2
SR_RENDER_END_DEAD_PATHS = 4


def render_flags(light_tree=False, subsurface=False, end_dead_paths=False):
    """The SR_RENDER_* mask of Renderer.render's mode arguments."""
    return ((SR_RENDER_LIGHT_TREE if light_tree else 0) | (SR_RENDER_SUBSURFACE if subsurface else 0)
            | (SR_RENDER_END_DEAD_PATHS if end_dead_paths else 0))


CHECKPOINT_VERSION = 1


def check_checkpoint(meta, scene_sha256, bands, accum_shape):
    """Raise ValueError unless a checkpoint's metadata belongs to this scene
    (geometry, materials, camera and band grid, by digest) and its
    accumulator has the [W, H, B] shape it records."""
    if meta.get("version") != CHECKPOINT_VERSION:
        raise ValueError(f"checkpoint version {meta.get('version')!r}, expected {CHECKPOINT_VERSION}")
    if meta.get("scene_sha256") != scene_sha256:
        raise ValueError("checkpoint was rendered from another scene (geometry, materials, camera or band grid)")
    if meta.get("bands") != bands:
        raise ValueError(f"checkpoint has {meta.get('bands')} bands, the renderer {bands}")
    want = (meta.get("width"), meta.get("height"), meta.get("bands"))
    if tuple(accum_shape) != want:
        raise ValueError(f"checkpoint accumulator shape {tuple(accum_shape)} is not {want}")
    if not isinstance(meta.get("blocks_done"), int) or meta["blocks_done"] < 0:
        raise ValueError("checkpoint blocks_done must be a non-negative integer")
    if meta["blocks_done"] > 0 and not isinstance(meta.get("render"), dict):
        raise ValueError("checkpoint of a started render carries no render arguments")


class Renderer:
    """One scene, one accumulator.

    ``scene`` is a reference-style scene dict or a ``FlatScene``.  ``build``
    chooses where the scene is compiled (normals, light lists, BVH):
    ``"host"`` (SAH BVH built in C++ on the CPU, ``sr_upload_scene``) or
    ``"device"`` (the same records built on the GPU, ``sr_upload_scene_gpu``).
    """

    def __init__(self, scene, spectral_bands, wavelength_min, wavelength_max, width=None, height=None,
                 device="cuda", build="host"):
        if build not in ("host", "device"):
            raise ValueError(f"build must be 'host' or 'device', not {build!r}")
        self.build = build
        self.bands = int(spectral_bands)
        self.wl_min = float(wavelength_min)
        self.wl_max = float(wavelength_max)
        self.wavelength_step = (self.wl_max - self.wl_min) / (self.bands - 1)
        self.device = torch.device(device)
        if isinstance(scene, FlatScene):
            self.flat = scene
        else:
            self.dims = scene_dimensions(scene, self.bands)
            self.flat = flatten_scene(scene, self.bands, self.wl_min, self.wl_max)
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _lib.call("sr_create", self.bands, self.wl_min, self.wl_max, ctypes.byref(h))
        self._h = h
        self._upload(self.flat)
        self.accum = None
        self.width = self.height = 0
        self.blocks_done = 0
        self._render_args = None  # the estimator the accumulated blocks belong to
        if width is not None:
            self.set_resolution(width, height if height is not None else width)

    # ------------------------------------------------------------ lifecycle
    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.load().sr_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _bind_stream(self):
        """Point the handle at torch's current stream (one ctypes call only
        when that stream changed since the last call)."""
        st = torch.cuda.current_stream(self.device).cuda_stream
        if st != getattr(self, "_bound_stream", None):
            _lib.call("sr_set_stream", self._h, ctypes.c_void_p(st))
            self._bound_stream = st

    def _upload(self, f: FlatScene):
        self._bind_stream()
        if self.build == "device":
            dv = torch.from_numpy(np.ascontiguousarray(f.tris_v, np.float32)).to(self.device)
            dn = torch.from_numpy(np.ascontiguousarray(f.tris_n, np.float32)).to(self.device)
            dm = torch.from_numpy(np.ascontiguousarray(f.tri_mat, np.int32)).to(self.device)
            _lib.call("sr_upload_scene_gpu", self._h, _lib.ptr(dv), _lib.ptr(dn), _lib.ptr(dm), f.n_tris,
                      _lib.ptr(f.mat_f32), _lib.ptr(f.mat_i32), _lib.ptr(f.mat_eem), f.n_mats, _lib.ptr(f.eem),
                      f.n_eem, _lib.ptr(f.cam))
            return
        _lib.call("sr_upload_scene", self._h, _lib.ptr(f.tris_v), _lib.ptr(f.tris_n), _lib.ptr(f.tri_mat), f.n_tris,
                  _lib.ptr(f.mat_f32), _lib.ptr(f.mat_i32), _lib.ptr(f.mat_eem), f.n_mats, _lib.ptr(f.eem),
                  f.n_eem, _lib.ptr(f.cam))

    def update_materials(self, mat_f32, mat_i32=None, mat_eem=None, eem=None):
        """Replace material tables (same shapes) without rebuilding the BVH."""
        f = self.flat
        mf = np.ascontiguousarray(mat_f32, np.float32)
        mi = np.ascontiguousarray(f.mat_i32 if mat_i32 is None else mat_i32, np.int32)
        me = np.ascontiguousarray(f.mat_eem if mat_eem is None else mat_eem, np.int32)
        et = np.ascontiguousarray(f.eem if eem is None else eem, np.float32)
        self._bind_stream()
        _lib.call("sr_update_materials", self._h, _lib.ptr(mf), _lib.ptr(mi), _lib.ptr(me), mf.shape[0],
                  _lib.ptr(et), et.shape[0])
        self.flat.mat_f32, self.flat.mat_i32, self.flat.mat_eem, self.flat.eem = mf, mi, me, et

    def scene_info(self):
        n = ctypes.c_int()
        d = ctypes.c_int()
        m = ctypes.c_int()
        b = ctypes.c_longlong()
        _lib.call("sr_scene_info", self._h, ctypes.byref(n), ctypes.byref(d), ctypes.byref(m), ctypes.byref(b))
        return {"bvh_nodes": n.value, "bvh4_depth": d.value, "max_lights_per_band": m.value,
                "device_bytes": b.value}

    SCENE_ARRAYS = {"nodes": 0, "tris": 1, "shade": 2, "lights": 3, "lverts": 4, "light_count": 5, "light_area": 6,
                    "light_tree": 7, "light_power": 8, "light_keys": 9}

    def scene_array(self, name):
        """The compiled scene array `name` (SCENE_ARRAYS) as raw bytes
        (numpy uint8), read back from the device."""
        which = self.SCENE_ARRAYS[name]
        size = ctypes.c_size_t()
        _lib.call("sr_scene_array", self._h, which, None, 0, ctypes.byref(size))
        out = np.empty(size.value, dtype=np.uint8)
        _lib.call("sr_scene_array", self._h, which, out.ctypes.data_as(ctypes.c_void_p), out.nbytes,
                  ctypes.byref(size))
        return out

    # ------------------------------------------------------------ reference protocol
    def set_resolution(self, width, height):
        self.width, self.height = int(width), int(height)
        self.accum = torch.zeros((self.width, self.height, self.bands), dtype=torch.float32, device=self.device)
        self.blocks_done = 0
        self._render_args = None

    def clear_pixel_space(self):
        """clear_pixel_space + counter reset (main.py:1661-1674, 1819-1822)."""
        if self.accum is not None:
            self.accum.zero_()
        self.blocks_done = 0
        self._render_args = None

    def update_camera(self, origin, lookat, up):
        """update_camera (main.py:633-640) followed by the reset the reference
        performs on every camera change (main.py:1816-1822)."""
        cam = np.array(list(origin) + list(lookat) + list(up), np.float32)
        _lib.call("sr_set_camera", self._h, _lib.ptr(cam))
        self.flat.cam = cam
        self.clear_pixel_space()

    def render(self, n_blocks=1, max_depth=7, seed=0, stream=0, light_tree=False, subsurface=False,
               end_dead_paths=False, path_totals=None):
        """Run blocks [blocks_done, blocks_done + n_blocks) and fold them into
        the running spectral mean ``accum`` [W, H, B] (cast_rays x B +
        average_and_reset_block per block, main.py:1826-1836).
        ``light_tree``: NEE picks lights through the light tree instead of
        the reference's area CDF (the same expected image wherever the
        reference's NEE and BSDF strategies estimate the same integral; lower
        variance with many small lights).  ``subsurface``: opaque materials with a
        scattering coefficient scatter by a random walk inside their (closed)
        mesh instead of the single-scatter Oren-Nayar lobe.  ``end_dead_paths``:
        a path ends once its throughput is exactly 0 instead of running to
        ``max_depth`` (the same image; fewer segments executed and counted).
        ``path_totals``: a float32 tensor on the render's device with
        n_blocks x W x H x B elements that receives each path's summed
        contributions (NaN where fluorescence moved some of them to another
        band) for the adjoint; not with ``light_tree`` or ``subsurface``."""
        if self.accum is None:
            raise RuntimeError("set_resolution() first")
        self._bind_stream()
        flags = render_flags(light_tree, subsurface, end_dead_paths)
        if path_totals is None:
            _lib.call("sr_render", self._h, self.width, self.height, int(seed), int(stream), self.blocks_done,
                      int(n_blocks), int(max_depth), flags, _lib.ptr(self.accum))
        else:
            n = int(n_blocks) * self.width * self.height * self.bands
            if (path_totals.dtype != torch.float32 or path_totals.device != self.accum.device
                    or not path_totals.is_contiguous() or path_totals.numel() != n):
                raise ValueError(f"path_totals must be a contiguous float32 tensor of {n} elements on "
                                 f"{self.accum.device}")
            _lib.call("sr_render_record", self._h, self.width, self.height, int(seed), int(stream),
                      self.blocks_done, int(n_blocks), int(max_depth), flags, _lib.ptr(self.accum),
                      _lib.ptr(path_totals))
        self.blocks_done += int(n_blocks)
        self._render_args = {"seed": int(seed), "stream": int(stream), "max_depth": int(max_depth),
                             "light_tree": bool(light_tree), "subsurface": bool(subsurface),
                             "end_dead_paths": bool(end_dead_paths)}
        return self.accum

    # ------------------------------------------------------------ checkpoint / resume
    # No counterpart in the reference: its accumulators live only in device
    # fields and are cleared on every camera or mode change (main.py:1661-1674,
    # 1812-1823).  Paths draw their uniforms from counters keyed by (seed,
    # stream, pixel, block, band, segment), so blocks [n, n + k) rendered
    # after a resume are the blocks an uninterrupted render would have run.
    def scene_digest(self):
        """sha256 of the uploaded scene (triangles, normals, materials, EEM
        tables, camera) and of the band grid."""
        h = hashlib.sha256()
        f = self.flat
        for a in (f.tris_v, f.tris_n, f.tri_mat, f.mat_f32, f.mat_i32, f.mat_eem, f.eem, f.cam):
            a = np.ascontiguousarray(a)
            h.update(f"{a.dtype.str}{a.shape}".encode())
            h.update(a.tobytes())
        h.update(np.array([self.bands, self.wl_min, self.wl_max], np.float64).tobytes())
        return h.hexdigest()

    def checkpoint(self, path):
        """Write the running mean, blocks_done and the render arguments of
        the accumulated blocks to ``path`` (numpy .npz, no pickled objects)."""
        if self.accum is None:
            raise RuntimeError("set_resolution() first")
        meta = {"version": CHECKPOINT_VERSION, "blocks_done": int(self.blocks_done), "width": self.width,
                "height": self.height, "bands": self.bands, "scene_sha256": self.scene_digest(),
                "render": self._render_args}
        with open(path, "wb") as fh:
            np.savez(fh, accum=self.accum.detach().cpu().numpy(), meta=np.array(json.dumps(meta)))

    def resume(self, path):
        """Load a checkpoint of this scene: the resolution, running mean and
        blocks_done are restored, and the render arguments of its blocks are
        returned (pass them to ``render`` to continue the same estimate;
        None if no block was rendered).  ValueError if the checkpoint
        belongs to another scene or band grid."""
        with np.load(path, allow_pickle=False) as z:
            meta = json.loads(str(z["meta"]))
            acc = np.ascontiguousarray(z["accum"], np.float32)
        check_checkpoint(meta, self.scene_digest(), self.bands, acc.shape)
        self.set_resolution(meta["width"], meta["height"])
        self.accum.copy_(torch.from_numpy(acc))
        self.blocks_done = meta["blocks_done"]
        self._render_args = meta["render"]
        return None if self._render_args is None else dict(self._render_args)

    def color_rays(self, colour_system, exposure=1.0, gamma=2.0, out_width=None, out_height=None):
        """color_rays (main.py:1609-1629) -> [out_w, out_h, 3] tensor."""
        ow = self.width if out_width is None else int(out_width)
        oh = self.height if out_height is None else int(out_height)
        cmf = np.ascontiguousarray(colour_system.cmf, np.float32)
        T = np.ascontiguousarray(colour_system.T, np.float32)
        norm = float(np.float32(np.sum(colour_system.cmf[:, 1]) * self.wavelength_step))
        out = torch.empty((ow, oh, 3), dtype=torch.float32, device=self.device)
        self._bind_stream()
        _lib.call("sr_resolve", self._h, _lib.ptr(self.accum), self.width, self.height, _lib.ptr(cmf), _lib.ptr(T),
                  norm, float(exposure), float(gamma), ow, oh, _lib.ptr(out))
        return out

    def visualize_normals(self, out_width=None, out_height=None):
        """visualize_normals (main.py:1632-1660) -> [out_w, out_h, 3] tensor:
        shading normal * 0.5 + 0.5 of the pixel-centre camera ray, black on a
        miss, at the render resolution nearest-upsampled to the display."""
        ow = self.width if out_width is None else int(out_width)
        oh = self.height if out_height is None else int(out_height)
        out = torch.empty((ow, oh, 3), dtype=torch.float32, device=self.device)
        self._bind_stream()
        _lib.call("sr_render_normals", self._h, self.width, self.height, ow, oh, _lib.ptr(out))
        return out

    def to_rgba8(self, rgb):
        """Display conversion of a [w, h, 3] image in [0, 1] -> [w, h, 4] uint8."""
        rgb = rgb.contiguous()
        out = torch.empty(rgb.shape[:-1] + (4,), dtype=torch.uint8, device=rgb.device)
        self._bind_stream()
        _lib.call("sr_to_rgba8", self._h, _lib.ptr(rgb), int(rgb.numel() // 3), _lib.ptr(out))
        return out

    # ------------------------------------------------------------ measurement / parity
    def stats(self):
        s = ctypes.c_uint64()
        t = ctypes.c_uint64()
        n = ctypes.c_uint64()
        _lib.call("sr_stats", self._h, ctypes.byref(s), ctypes.byref(t), ctypes.byref(n))
        return {"segments": s.value, "traversals": t.value, "nonfinite": n.value}

    def last_kernel_ms(self):
        ms = ctypes.c_float()
        _lib.call("sr_last_kernel_ms", self._h, ctypes.byref(ms))
        return ms.value

    def sync(self):
        _lib.call("sr_sync", self._h)

    def trace(self, orig, dirs):
        """Closest hit for rays (numpy [n,3] f32 host arrays or cuda tensors).
        Returns (t, prim, uv) of the same kind."""
        if isinstance(orig, torch.Tensor):
            n = orig.shape[0]
            t = torch.empty((n,), dtype=torch.float32, device=orig.device)
            p = torch.empty((n,), dtype=torch.int32, device=orig.device)
            uv = torch.empty((n, 2), dtype=torch.float32, device=orig.device)
            self._bind_stream()
            _lib.call("sr_trace_closest", self._h, n, _lib.ptr(orig.contiguous()), _lib.ptr(dirs.contiguous()),
                      _lib.ptr(t), _lib.ptr(p), _lib.ptr(uv), _lib.SR_DEVICE_PTRS)
            return t, p, uv
        o = np.ascontiguousarray(orig, np.float32)
        d = np.ascontiguousarray(dirs, np.float32)
        n = o.shape[0]
        t = np.empty((n,), np.float32)
        p = np.empty((n,), np.int32)
        uv = np.empty((n, 2), np.float32)
        self._bind_stream()
        _lib.call("sr_trace_closest", self._h, n, _lib.ptr(o), _lib.ptr(d), _lib.ptr(t), _lib.ptr(p), _lib.ptr(uv), 0)
        return t, p, uv

    def primary_rays(self, width=None, height=None):
        return pixel_center_rays(self.flat, width or self.width, height or self.height)
