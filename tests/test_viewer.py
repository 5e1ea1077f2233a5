"""Host logic of the headless interactive loop (spectral.viewer): the
camera basis of update_camera (main.py:633-640), the FPS camera controller
of the GUI loop (main.py:1757-1815) and the PNG writer."""
import math
import struct
import zlib

import numpy as np
import pytest

from spectral.viewer import CameraController, camera_basis, write_png


def test_camera_basis_orthonormal_right_handed():
    u, v, w = camera_basis([0, 0, 3], [0, 0, 0], [0, 1, 0])
    assert np.allclose(w, [0, 0, 1]) and np.allclose(u, [1, 0, 0]) and np.allclose(v, [0, 1, 0])
    u, v, w = camera_basis([1, 2, 3], [-2, 0.5, 1], [0, 1, 0])
    for a in (u, v, w):
        assert np.linalg.norm(a) == pytest.approx(1, abs=1e-6)
    assert abs(u @ v) < 1e-6 and abs(u @ w) < 1e-6 and abs(v @ w) < 1e-6
    assert np.allclose(np.cross(u, v), w, atol=1e-6)


def test_controller_moves_in_the_horizontal_plane():
    # camera looking down -z, tilted: forward is flattened to -z
    c = CameraController([0.0, 0.5, 2.0], [0.0, 0.0, 1.0], [0, 1, 0])
    assert c.press({"w"})
    assert np.allclose(c.pos, [0.0, 0.5, 1.95], atol=1e-7)
    # yaw starts at -pi/2: the new look-at is one unit down -z
    assert np.allclose(c.lookat - c.pos, [0.0, 0.0, -1.0], atol=1e-7)
    c.press({"d"})  # right = +x
    assert np.allclose(c.pos, [0.05, 0.5, 1.95], atol=1e-7)
    c.press({"e"})
    c.press({"q", "q"})  # a set: one step down
    assert c.pos[1] == pytest.approx(0.5, abs=1e-7)
    assert not c.press({"r"})  # not a camera key


def test_controller_turns_and_clamps_pitch():
    c = CameraController([0, 0, 0], [0, 0, -1], [0, 1, 0])
    c.press({"LEFT"})
    assert c.yaw == pytest.approx(-math.pi / 2 - 0.03)
    for _ in range(100):
        c.press({"UP"})
    assert c.pitch == pytest.approx(math.pi / 2 - 0.01)
    look = c.lookat - c.pos
    assert look[1] == pytest.approx(math.sin(math.pi / 2 - 0.01), abs=1e-6)


def _read_png(path):
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, w = 8, b"", 0
    while pos < len(data):
        n, tag = struct.unpack(">I4s", data[pos:pos + 8])
        body = data[pos + 8:pos + 8 + n]
        assert struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])[0] == zlib.crc32(tag + body) & 0xFFFFFFFF
        if tag == b"IHDR":
            w, h, depth, ctype = struct.unpack(">IIBB", body[:10])
            assert (depth, ctype) == (8, 6)
        elif tag == b"IDAT":
            idat += body
        pos += 12 + n
    raw = zlib.decompress(idat)
    rows = [raw[r * (1 + 4 * w) + 1:(r + 1) * (1 + 4 * w)] for r in range(h)]
    return np.frombuffer(b"".join(rows), np.uint8).reshape(h, w, 4)


def test_png_roundtrip_orientation(tmp_path):
    rng = np.random.default_rng(0)
    img = rng.integers(0, 256, (5, 3, 4), dtype=np.uint8)  # [x, y] with y up
    p = tmp_path / "f.png"
    write_png(p, img)
    back = _read_png(p)  # rows top to bottom
    assert back.shape == (3, 5, 4)
    assert np.array_equal(back[0, :, :], img[:, 2, :])  # top row = largest y
    assert np.array_equal(back[2, 4], img[4, 0])
