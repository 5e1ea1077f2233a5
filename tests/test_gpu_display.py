"""Display path on the GPU: the normals view (visualize_normals,
main.py:1632-1660) against brute-force hits, the RGBA8 display conversion,
and the headless interactive loop (main.py:1735-1846) against direct calls."""
import numpy as np
import pytest
import torch

from oracle.oracle import OracleScene
from spectral import configs
from spectral.data import srgb_colour_system
from spectral.render import Renderer
from spectral.scene import flatten_scene, pixel_center_rays
from spectral.viewer import Session

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def box(gpu):
    cfg = configs.box_icosphere()
    flat = flatten_scene(cfg.scene, cfg.bands, cfg.wl_min, cfg.wl_max)
    return cfg, flat


def test_normals_view_matches_brute_force(box, gpu):
    cfg, flat = box
    W, H = 48, 32
    r = Renderer(flat, cfg.bands, cfg.wl_min, cfg.wl_max, width=W, height=H, device=gpu)
    img = r.visualize_normals(96, 64).cpu().numpy()  # 2x nearest upsample
    o, d = pixel_center_rays(flat, W, H)
    t, prim, uv = OracleScene(flat, brute=True).trace(o, d)
    n = flat.tris_n.astype(np.float64)  # [N, 3, 3] per-vertex normals
    exp = np.zeros((W * H, 3))
    hit = prim >= 0
    p = prim[hit]
    u, v = uv[hit, 0].astype(np.float64), uv[hit, 1].astype(np.float64)
    sn = n[p, 0] * (1 - u - v)[:, None] + n[p, 1] * u[:, None] + n[p, 2] * v[:, None]
    sn /= np.linalg.norm(sn, axis=1, keepdims=True)
    exp[hit] = sn * 0.5 + 0.5
    exp = exp.reshape(W, H, 3)
    # the display index of color_rays / visualize_normals, in f32
    ri = (np.arange(96, dtype=np.float32) / np.float32(96) * np.float32(W)).astype(int)
    rj = (np.arange(64, dtype=np.float32) / np.float32(64) * np.float32(H)).astype(int)
    up = exp[ri][:, rj]
    close = np.abs(img - up).max(-1) <= 1e-5
    assert close.mean() >= 0.995
    miss = ~hit.reshape(W, H)[ri][:, rj]
    assert (img[miss] == 0).all()
    # the upsampled view is the render-grid view gathered, bit for bit (one
    # ray per render pixel, as the reference's pixels_geometry)
    same = r.visualize_normals(W, H).cpu().numpy()
    assert np.array_equal(img, same[ri][:, rj])


def test_rgba8_closed_form(box, gpu):
    cfg, flat = box
    r = Renderer(flat, cfg.bands, cfg.wl_min, cfg.wl_max, device=gpu)
    x = torch.tensor(np.random.default_rng(1).uniform(-0.5, 1.5, (7, 5, 3)), dtype=torch.float32)
    x[0, 0, 0] = float("nan")
    x[1, 1] = torch.tensor([0.5 / 255, 1.5 / 255, 254.5 / 255])
    out = r.to_rgba8(x.to(gpu)).cpu().numpy()
    xn = np.nan_to_num(x.numpy(), nan=0.0)
    exp = np.floor(np.clip(xn, 0, 1).astype(np.float32) * np.float32(255) + np.float32(0.5)).astype(np.uint8)
    assert np.array_equal(out[..., :3], exp)
    assert (out[..., 3] == 255).all()


def test_session_matches_direct_calls(box, gpu):
    cfg, flat = box
    W = H = 16
    cs = srgb_colour_system(cfg.wl_min, cfg.wl_max, cfg.bands)
    s = Session(Renderer(flat, cfg.bands, cfg.wl_min, cfg.wl_max, width=W, height=H, device=gpu), cs,
                spp_blocks=3, max_depth=cfg.depth, seed=4)
    for _ in range(5):
        f = s.step()
    assert s.finished and s.r.blocks_done == 3
    ref = Renderer(flat, cfg.bands, cfg.wl_min, cfg.wl_max, width=W, height=H, device=gpu)
    ref.render(n_blocks=3, max_depth=cfg.depth, seed=4)
    assert np.allclose(s.r.accum.cpu().numpy(), ref.accum.cpu().numpy(), rtol=1e-5, atol=1e-7)
    want = ref.to_rgba8(ref.color_rays(cs, 1.0, 2.0, W, H)).cpu().numpy()
    assert np.abs(f.cpu().numpy().astype(int) - want.astype(int)).max() <= 1
    # a camera move restarts the accumulation from the moved camera
    assert s.press("w")
    assert s.r.blocks_done == 0
    s.step()
    assert s.r.blocks_done == 1
    # 'r': geometry mode shows the normals view
    s.press("r")
    g = s.step().cpu().numpy()
    assert np.array_equal(g, s.r.to_rgba8(s.r.visualize_normals(W, H)).cpu().numpy())
