"""ctypes binding of libspectral_render.so (the C ABI in include/spectral_render.h).

The library is built in-tree (``make -C taichi-spectral-render_amd``) and
loaded from ``spectral/_lib``.  There is no fallback: if it is missing every
entry point raises.  torch is imported before the library is loaded so both
share the one HIP runtime already mapped into the process.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import torch  # noqa: F401  (maps the HIP runtime first; see module docstring)

LIB_PATH = Path(os.environ.get("SPECTRAL_RENDER_LIB",
                               Path(__file__).resolve().parent / "_lib" / "libspectral_render.so"))

SR_DEVICE_PTRS = 1
ERRORS = {-1: "EINVAL", -2: "ESTATE", -3: "EHIP", -4: "ENOMEM", -5: "EOVERFLOW"}

_c_float_p = ctypes.POINTER(ctypes.c_float)
_c_int_p = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p

# (name, argtypes) of every exported entry point; restype is int unless noted.
SIGNATURES = {
    "sr_last_error": ([], ctypes.c_char_p),
    "sr_version": ([], ctypes.c_int),
    "sr_create": ([ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.POINTER(_vp)], ctypes.c_int),
    "sr_destroy": ([_vp], ctypes.c_int),
    "sr_set_stream": ([_vp, _vp], ctypes.c_int),
    "sr_upload_scene": ([_vp, _vp, _vp, _vp, ctypes.c_int, _vp, _vp, _vp, ctypes.c_int, _vp, ctypes.c_int, _vp],
                        ctypes.c_int),
    "sr_upload_scene_gpu": ([_vp, _vp, _vp, _vp, ctypes.c_int, _vp, _vp, _vp, ctypes.c_int, _vp, ctypes.c_int,
                             _vp], ctypes.c_int),
    "sr_set_camera": ([_vp, _vp], ctypes.c_int),
    "sr_render_normals": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp], ctypes.c_int),
    "sr_to_rgba8": ([_vp, _vp, ctypes.c_int64, _vp], ctypes.c_int),
    "sr_update_materials": ([_vp, _vp, _vp, _vp, ctypes.c_int, _vp, ctypes.c_int], ctypes.c_int),
    "sr_scene_info": ([_vp, _c_int_p, _c_int_p, _c_int_p, ctypes.POINTER(ctypes.c_longlong)], ctypes.c_int),
    "sr_scene_array": ([_vp, ctypes.c_int, _vp, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)], ctypes.c_int),
    "sr_trace_closest": ([_vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, ctypes.c_int], ctypes.c_int),
    "sr_render": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int, ctypes.c_int,
                   ctypes.c_int, ctypes.c_int, _vp], ctypes.c_int),
    "sr_render_noalloc": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int,
                           ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp], ctypes.c_int),
    "sr_render_adjoint": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int,
                           ctypes.c_int, ctypes.c_int, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                           ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp], ctypes.c_int),
    "sr_render_record": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int,
                          ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _vp], ctypes.c_int),
    "sr_render_adjoint_record": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int,
                                  ctypes.c_int, ctypes.c_int, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _vp],
                                 ctypes.c_int),
    "sr_resolve": ([_vp, _vp, ctypes.c_int, ctypes.c_int, _vp, _vp, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                    ctypes.c_int, ctypes.c_int, _vp], ctypes.c_int),
    "sr_stats": ([_vp, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64),
                  ctypes.POINTER(ctypes.c_uint64)], ctypes.c_int),
    "sr_last_kernel_ms": ([_vp, _c_float_p], ctypes.c_int),
    "sr_sync": ([_vp], ctypes.c_int),
    "sr_batch_create": ([ctypes.POINTER(_vp)], ctypes.c_int),
    "sr_batch_destroy": ([_vp], ctypes.c_int),
    "sr_batch_submit": ([_vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int,
                         ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp], ctypes.c_int),
    "sr_batch_clear": ([_vp], ctypes.c_int),
    "sr_batch_config": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "sr_batch_launch": ([_vp, ctypes.c_int], ctypes.c_int),
    "sr_batch_wait": ([_vp, _c_float_p, _vp, _vp, _vp], ctypes.c_int),
}

_LIB = None


class SpectralRenderError(RuntimeError):
    def __init__(self, fn, code, msg):
        super().__init__(f"{fn} failed with {ERRORS.get(code, code)}: {msg}")
        self.code = code


def load():
    """Load the library (once) and declare every signature."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not LIB_PATH.exists():
        raise RuntimeError(f"HIP library not built: {LIB_PATH} is missing (run __graft_entry__.build())")
    lib = ctypes.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | os.RTLD_GLOBAL)
    for name, (args, res) in SIGNATURES.items():
        fn = getattr(lib, name, None)
        if fn is None:  # an older build: the call will fail with AttributeError
            continue
        fn.argtypes = args
        fn.restype = res
    _LIB = lib
    return lib


def call(name, *args):
    """Call an int-returning entry point and raise on a negative status."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        raise SpectralRenderError(name, rc, lib.sr_last_error().decode())
    return rc


def ptr(a):
    """Address of a numpy array's or torch tensor's data as c_void_p."""
    if a is None:
        return None
    if isinstance(a, torch.Tensor):
        return ctypes.c_void_p(a.data_ptr())
    return a.ctypes.data_as(ctypes.c_void_p)
