"""Checkpoint / resume of a render's running mean (no counterpart in the
reference, whose accumulators live only in device fields,
main.py:1661-1674): the file round trip and the checks that keep a
checkpoint to its own scene, on host state only (no library handle)."""
import copy

import numpy as np
import pytest
import torch

from spectral import configs
from spectral.render import CHECKPOINT_VERSION, Renderer, check_checkpoint
from spectral.scene import flatten_scene


@pytest.fixture(scope="module")
def box():
    cfg = configs.box_icosphere()
    return cfg, flatten_scene(cfg.scene, cfg.bands, cfg.wl_min, cfg.wl_max)


def _host_renderer(cfg, flat, W=6, H=4):
    r = Renderer.__new__(Renderer)  # the renderer's host state without a device handle
    r.flat = flat
    r.bands, r.wl_min, r.wl_max = cfg.bands, cfg.wl_min, cfg.wl_max
    r.device = torch.device("cpu")
    r.accum, r.width, r.height, r.blocks_done, r._render_args = None, 0, 0, 0, None
    r.set_resolution(W, H)
    return r


ARGS = {"seed": 11, "stream": 2, "max_depth": 8, "light_tree": False, "subsurface": False, "end_dead_paths": True}


def test_round_trip(box, tmp_path):
    cfg, flat = box
    a = _host_renderer(cfg, flat)
    a.accum.copy_(torch.rand(a.accum.shape, generator=torch.Generator().manual_seed(0)))
    a.blocks_done, a._render_args = 5, dict(ARGS)
    a.checkpoint(tmp_path / "ck.npz")
    b = _host_renderer(cfg, flat, W=3, H=3)  # the checkpoint's resolution replaces this one
    assert b.resume(tmp_path / "ck.npz") == ARGS
    assert (b.width, b.height, b.blocks_done) == (6, 4, 5)
    assert torch.equal(b.accum, a.accum)
    # the file holds plain arrays: it loads without unpickling
    with np.load(tmp_path / "ck.npz", allow_pickle=False) as z:
        assert z["accum"].dtype == np.float32 and z["accum"].shape == (6, 4, cfg.bands)


def test_a_checkpoint_refuses_another_scene(box, tmp_path):
    cfg, flat = box
    a = _host_renderer(cfg, flat)
    a.blocks_done, a._render_args = 1, dict(ARGS)
    a.checkpoint(tmp_path / "ck.npz")
    moved = copy.deepcopy(flat)
    moved.cam = moved.cam + np.float32(0.25)
    with pytest.raises(ValueError, match="another scene"):
        _host_renderer(cfg, moved).resume(tmp_path / "ck.npz")
    tinted = copy.deepcopy(flat)
    tinted.mat_f32 = tinted.mat_f32 * np.float32(0.5)
    with pytest.raises(ValueError, match="another scene"):
        _host_renderer(cfg, tinted).resume(tmp_path / "ck.npz")


def test_digest_covers_the_band_grid(box):
    cfg, flat = box
    a, b = _host_renderer(cfg, flat), _host_renderer(cfg, flat)
    assert a.scene_digest() == b.scene_digest()
    b.wl_max = cfg.wl_max + 1.0
    assert a.scene_digest() != b.scene_digest()


def test_check_checkpoint_fields():
    meta = {"version": CHECKPOINT_VERSION, "blocks_done": 3, "width": 4, "height": 2, "bands": 5,
            "scene_sha256": "ab", "render": dict(ARGS)}
    check_checkpoint(meta, "ab", 5, (4, 2, 5))
    check_checkpoint(dict(meta, blocks_done=0, render=None), "ab", 5, (4, 2, 5))  # nothing rendered yet
    for bad, match in [({"version": 0}, "version"), ({"scene_sha256": "cd"}, "another scene"),
                       ({"bands": 6}, "bands"), ({"width": 5}, "shape"), ({"blocks_done": -1}, "non-negative"),
                       ({"blocks_done": 2.0}, "non-negative"), ({"render": None}, "render arguments")]:
        with pytest.raises(ValueError, match=match):
            check_checkpoint(dict(meta, **bad), "ab", 5, (4, 2, 5))
