"""The C-ABI library loads and exports every entry point include/*.h
declares; argument validation happens before any device work."""
import ctypes
import re
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        names |= set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(sr_\w+)\s*\(", h.read_text(), re.M))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    from spectral import _lib

    lib = _lib.load()
    names = declared_symbols()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) <= set(_lib.SIGNATURES), set(names) - set(_lib.SIGNATURES)


def test_version_and_argument_errors():
    from spectral import _lib

    lib = _lib.load()
    assert lib.sr_version() == 100
    h = ctypes.c_void_p()
    assert lib.sr_create(1, 400.0, 700.0, ctypes.byref(h)) == -1
    assert b"bands" in lib.sr_last_error()
    assert lib.sr_create(31, 700.0, 400.0, ctypes.byref(h)) == -1
    assert lib.sr_render(None, 8, 8, 0, 0, 0, 1, 4, 0, None) == -1
    assert lib.sr_trace_closest(None, 1, None, None, None, None, None, 0) == -1


def test_create_rejects_bad_bands_and_wavelengths():
    """sr_create validates its arguments before touching the device."""
    import ctypes

    from spectral import _lib

    lib = _lib.load()
    h = ctypes.c_void_p()
    for bands, lo, hi in ((1, 400.0, 700.0), (65, 400.0, 700.0), (31, 700.0, 400.0), (31, 0.0, 700.0)):
        assert lib.sr_create(bands, lo, hi, ctypes.byref(h)) == -1
        assert h.value is None
        msg = lib.sr_last_error().decode()
        assert ("bands" in msg) or ("wl_min" in msg)


def test_every_entry_point_rejects_a_null_handle():
    """Every entry point that takes a handle or a batch returns SR_EINVAL
    with a message for a null one, before any device call (all other
    arguments zero or null)."""
    from spectral import _lib

    lib = _lib.load()
    skip = {"sr_last_error", "sr_version", "sr_create", "sr_destroy", "sr_batch_create", "sr_batch_destroy"}
    for name, (argtypes, _) in _lib.SIGNATURES.items():
        if name in skip:
            continue
        args = [0.0 if t is ctypes.c_float else (0 if t in (ctypes.c_int, ctypes.c_int64, ctypes.c_uint64,
                                                            ctypes.c_uint32, ctypes.c_size_t) else None)
                for t in argtypes]
        assert getattr(lib, name)(*args) == -1, name
        assert b"null" in lib.sr_last_error(), name
