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
