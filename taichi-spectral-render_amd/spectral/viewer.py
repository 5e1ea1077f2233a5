"""The reference's interactive loop (main.py:1735-1846) without a window.

``CameraController`` is the FPS-style camera of the loop: WASD moves in the
horizontal plane along the camera's forward / right vectors (y flattened,
main.py:1767-1775), E / Q move up / down, the arrow keys turn (yaw) and tilt
(pitch, clamped short of +-pi/2), and every change rebuilds the look-at
point from yaw / pitch (main.py:1807-1815).  The yaw starts at -pi/2
(looking down -z) whatever the scene camera's look-at point, as in the
reference, so the first change snaps the view to that heading.

``Session`` runs the loop one iteration per ``step()``: key events first
(``press``), then either one spectral block of path tracing folded into the
running mean and colour-resolved (render mode), or the normals view
(geometry mode, key 'r' toggles and resets); a camera change or mode switch
clears the accumulation.  The reference advances one band per GUI frame and
resolves once per completed block (main.py:1826-1836); a step here is one
block, the display granularity.  ``frame()`` returns the RGBA8 display
image, ``save_png`` writes it.
"""
from __future__ import annotations

import math
import struct
import zlib

import numpy as np
import torch

KEYS = ("w", "s", "a", "d", "e", "q", "LEFT", "RIGHT", "UP", "DOWN", "r", "ESCAPE")


def _norm(v):
    n = float(np.linalg.norm(v))
    return v / n if n > 1e-6 else v


def camera_basis(origin, lookat, up):
    """update_camera (main.py:633-640): w = normalize(origin - lookat),
    u = normalize(up x w), v = w x u."""
    w = np.asarray(origin, np.float32) - np.asarray(lookat, np.float32)
    w = w / np.float32(np.linalg.norm(w))
    u = np.cross(np.asarray(up, np.float32), w)
    u = u / np.float32(np.linalg.norm(u))
    v = np.cross(w, u)
    return u.astype(np.float32), v.astype(np.float32), w.astype(np.float32)


class CameraController:
    def __init__(self, origin, lookat, up, move_speed=0.05, rot_speed=0.03):
        self.pos = np.array(origin, dtype=np.float32)
        self.lookat = np.array(lookat, dtype=np.float32)
        self.up = np.array(up, dtype=np.float32)
        self.yaw = -math.pi / 2.0
        self.pitch = 0.0
        self.move_speed = move_speed
        self.rot_speed = rot_speed

    def press(self, held):
        """One key-press event with the keys in ``held`` down
        (main.py:1757-1822).  Returns True when the camera changed."""
        held = set(held)
        u, _, w = camera_basis(self.pos, self.lookat, self.up)
        forward = -w
        forward[1] = 0.0
        forward = _norm(forward)
        right = u.copy()
        right[1] = 0.0
        right = _norm(right)
        changed = False
        steps = {"w": forward, "s": -forward, "a": -right, "d": right}
        for k in ("w", "s", "a", "d"):
            if k in held:
                self.pos = (self.pos + steps[k] * self.move_speed).astype(np.float32)
                changed = True
        if "e" in held:
            self.pos[1] += self.move_speed
            changed = True
        if "q" in held:
            self.pos[1] -= self.move_speed
            changed = True
        turns = {"LEFT": ("yaw", -1), "RIGHT": ("yaw", 1), "UP": ("pitch", 1), "DOWN": ("pitch", -1)}
        for k, (attr, sgn) in turns.items():
            if k in held:
                setattr(self, attr, getattr(self, attr) + sgn * self.rot_speed)
                changed = True
        self.pitch = max(-math.pi / 2.0 + 0.01, min(math.pi / 2.0 - 0.01, self.pitch))
        if changed:
            look = np.array([math.cos(self.yaw) * math.cos(self.pitch), math.sin(self.pitch),
                             math.sin(self.yaw) * math.cos(self.pitch)], dtype=np.float32)
            self.lookat = (self.pos + look).astype(np.float32)
        return changed


class Session:
    def __init__(self, renderer, colour_system, spp_blocks, max_depth, seed=0, exposure=1.0, gamma=2.0,
                 display=None):
        self.r = renderer
        self.cs = colour_system
        self.spp_blocks = int(spp_blocks)
        self.depth = int(max_depth)
        self.seed = int(seed)
        self.exposure = float(exposure)
        self.gamma = float(gamma)
        self.display = display or (renderer.width, renderer.height)
        cam = renderer.flat.cam
        self.camera = CameraController(cam[0:3], cam[3:6], cam[6:9])
        self.mode = "render"
        self.running = True
        self._pixels = torch.zeros(self.display + (3,), dtype=torch.float32, device=renderer.device)
        renderer.clear_pixel_space()

    def press(self, *held):
        """One key-press event (main.py:1757): movement keys held, 'r' switches
        the mode, ESCAPE stops the loop."""
        if "ESCAPE" in held:
            self.running = False
        changed = self.camera.press(held)
        if "r" in held:
            self.mode = "geometry" if self.mode == "render" else "render"
            self.r.clear_pixel_space()
            self._pixels.zero_()
        if changed:
            c = self.camera
            self.r.update_camera(c.pos, c.lookat, c.up)  # also clears the accumulation
            self._pixels.zero_()
        return changed

    @property
    def finished(self):
        return self.mode == "render" and self.r.blocks_done >= self.spp_blocks

    def step(self):
        """One loop iteration (main.py:1825-1846); returns the RGBA8 frame."""
        ow, oh = self.display
        if self.mode == "render":
            if not self.finished:
                self.r.render(n_blocks=1, max_depth=self.depth, seed=self.seed)
                self._pixels = self.r.color_rays(self.cs, self.exposure, self.gamma, ow, oh)
        else:
            self._pixels = self.r.visualize_normals(ow, oh)
        return self.frame()

    def frame(self):
        return self.r.to_rgba8(self._pixels)

    def save_png(self, path, frame=None):
        write_png(path, (self.frame() if frame is None else frame).cpu().numpy())


def write_png(path, rgba):
    """Write a [w, h, 4] uint8 image indexed [x, y] with y up (the reference's
    pixel field layout) as an 8-bit RGBA PNG (rows top to bottom)."""
    img = np.ascontiguousarray(np.transpose(np.asarray(rgba, np.uint8), (1, 0, 2))[::-1])
    h, w = img.shape[:2]
    raw = b"".join(b"\x00" + img[y].tobytes() for y in range(h))

    def chunk(tag, data):
        return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)

    png = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 6, 0, 0, 0))
    png += chunk(b"IDAT", zlib.compress(raw, 6)) + chunk(b"IEND", b"")
    with open(path, "wb") as f:
        f.write(png)
