"""Image parity of the persistent HIP path kernel (sr_render) against the CPU
restatement, plus closed-form images, the CMF resolve and error paths.

Matched RNG keys: both sides draw the same Philox stream per path, so most
paths are identical; differences come only from last-ulp differences of
transcendental functions that flip a branch.  The north-star gate: every
pixel-band within 3 sigma of the Monte Carlo standard error (sigma from the
CPU side's per-chunk spread) and mean relative error below 1% at 4096 spp.
Independent RNG streams: a plain two-sample z test per pixel-band.
"""
import numpy as np
import pytest

from oracle.oracle import OracleScene
from spectral import configs
from spectral.scene import flatten_scene

pytestmark = pytest.mark.gpu


def _renderer(flat, cfg, w, h, gpu):
    from spectral.render import Renderer

    return Renderer(flat, cfg.bands, cfg.wl_min, cfg.wl_max, width=w, height=h, device=gpu)


def _chunks_oracle(ora, w, h, seed, depth, n_chunks, per_chunk, stream=0):
    out = []
    for c in range(n_chunks):
        acc = np.zeros((w, h, ora.flat.bands), np.float32)
        # blocks [c*per, (c+1)*per) folded into a fresh mean: start from the
        # running-mean identity with first_block = c*per
        a, _ = ora.render(w, h, seed, c * per_chunk, per_chunk, depth, accum=acc, stream=stream)
        out.append(a * ((c + 1) * per_chunk) / per_chunk)  # undo the (first + n) normalisation
    return np.stack(out)


def _chunks_gpu(r, seed, depth, n_chunks, per_chunk, stream=0):
    out = []
    for c in range(n_chunks):
        r.accum.zero_()
        r.blocks_done = c * per_chunk
        r.render(n_blocks=per_chunk, max_depth=depth, seed=seed, stream=stream)
        out.append(r.accum.cpu().numpy() * ((c + 1) * per_chunk) / per_chunk)
    return np.stack(out)


@pytest.fixture(scope="module")
def box_small():
    cfg = configs.box_icosphere()
    flat = flatten_scene(cfg.scene, cfg.bands, cfg.wl_min, cfg.wl_max)
    return cfg, flat, OracleScene(flat)


def test_matched_seed_parity_4096spp(box_small, gpu):
    cfg, flat, ora = box_small
    W = H = 12
    n_chunks, per = 8, 512  # 4096 spp
    cpu = _chunks_oracle(ora, W, H, cfg.seed, cfg.depth, n_chunks, per)
    r = _renderer(flat, cfg, W, H, gpu)
    g = _chunks_gpu(r, cfg.seed, cfg.depth, n_chunks, per)
    mc, mg = cpu.mean(0), g.mean(0)
    sig = cpu.std(0, ddof=1) / np.sqrt(n_chunks)
    diff = np.abs(mg - mc)
    ok = diff <= 3 * sig + 1e-6 * np.abs(mc) + 1e-7
    assert ok.mean() >= 0.997, f"{(~ok).sum()} of {ok.size} pixel-bands beyond 3 sigma"
    lit = mc > 1e-3
    mre = (diff[lit] / mc[lit]).mean()
    assert mre < 0.01, f"mean relative error {mre:.4f}"


def test_counters_match_cpu(box_small, gpu):
    cfg, flat, ora = box_small
    W = H = 16
    _, st0 = ora.render(W, H, cfg.seed, 0, 16, cfg.depth)
    r = _renderer(flat, cfg, W, H, gpu)
    r.render(n_blocks=16, max_depth=cfg.depth, seed=cfg.seed)
    st = r.stats()
    assert abs(st["segments"] - st0["segments"]) <= 0.005 * st0["segments"]
    # the kernel does not trace shadow walks that cannot contribute (the
    # oracle traces them and counts them as culled)
    assert st0["culled_traversals"] > 0
    traced = st0["traversals"] - st0["culled_traversals"]
    assert abs(st["traversals"] - traced) <= 0.005 * traced
    nf = st0["nonfinite"] - st0["culled_nonfinite"]
    assert abs(st["nonfinite"] - nf) <= 0.02 * max(nf, 50)


@pytest.mark.parametrize("name,mat_id,roughness", [("volume_fluor", 6, None), ("box_icosphere", 5, 0.0017)])
def test_nonfinite_counts_match_cpu(name, mat_id, roughness, gpu):
    # The non-finite terms come from leaving a rough dielectric (a 0/0 BTDF
    # estimator, main.py:1176-1185); the device drops and counts them as the
    # oracle does.  With the sphere made smooth there are none on either side.
    cfg = getattr(configs, name)()
    scene = dict(cfg.scene)
    if roughness is not None:
        scene["materials"] = [dict(m) for m in scene["materials"]]
        scene["materials"][mat_id]["roughness"] = [roughness] * cfg.bands
    flat = flatten_scene(scene, cfg.bands, cfg.wl_min, cfg.wl_max)
    W = H = 16
    _, st0 = OracleScene(flat).render(W, H, cfg.seed, 0, 16, cfg.depth)
    r = _renderer(flat, cfg, W, H, gpu)
    r.render(n_blocks=16, max_depth=cfg.depth, seed=cfg.seed)
    st = r.stats()
    assert abs(st["segments"] - st0["segments"]) <= 0.005 * st0["segments"]
    nf = st0["nonfinite"] - st0["culled_nonfinite"]
    if roughness is None:
        assert nf > 1000
        assert abs(st["nonfinite"] - nf) <= 0.02 * nf
    else:
        assert nf == 0 and st["nonfinite"] == 0


def test_independent_streams_statistical(box_small, gpu):
    cfg, flat, ora = box_small
    W = H = 8
    n_chunks, per = 8, 256
    cpu = _chunks_oracle(ora, W, H, cfg.seed, cfg.depth, n_chunks, per, stream=0)
    r = _renderer(flat, cfg, W, H, gpu)
    g = _chunks_gpu(r, cfg.seed, cfg.depth, n_chunks, per, stream=1)
    s = np.sqrt(cpu.var(0, ddof=1) / n_chunks + g.var(0, ddof=1) / n_chunks)
    m = s > 0
    z = (g.mean(0) - cpu.mean(0))[m] / s[m]
    # t(7)-distributed per cell; |t| > 3.5 has probability ~1%
    assert (np.abs(z) > 3.5).mean() < 0.03
    assert abs(z.mean()) < 0.15


def _light_only(gpu):
    B = 4
    mats = [configs._mat(B, n=1.0, tv=1), configs._mat(B, se=1, em=10.0)]
    light = {"vertices": [[-0.3, 1, -0.3], [-0.3, 1, 0.3], [0.3, 1, 0.3], [0.3, 1, -0.3]],
             "faces": [[0, 2, 1], [0, 3, 2]], "material_id": 1}
    scene = {"materials": mats, "meshes": [light],
             "camera": {"origin": [0, 0, 0], "lookat": [0, 1, 0], "up": [0, 0, 1]}}
    cfg = configs.RenderConfig("light", scene, B, 400.0, 700.0, 16, 16, 3, 4, 1)
    flat = flatten_scene(scene, B, 400.0, 700.0)
    return cfg, flat, _renderer(flat, cfg, 16, 16, gpu)


def test_light_visible_pixels_exact(gpu):
    cfg, flat, r = _light_only(gpu)
    acc = r.render(n_blocks=3, max_depth=4, seed=1).cpu().numpy()
    expect = np.float32(np.float32(np.float32(1 / (1 + 1e-6)) * 10.0) * np.float32(1 / (1 + 1e-6)))
    assert np.allclose(acc[6:10, 6:10], expect, rtol=2e-6)
    assert (acc[0:2] == 0).all() and (acc[:, 0:2] == 0).all()


def test_zero_light_scene_black(gpu):
    cfg = configs.box_icosphere()
    scene = dict(cfg.scene)
    scene["materials"] = [dict(m) for m in scene["materials"]]
    scene["materials"][4]["surface_emission"] = [0] * cfg.bands
    flat = flatten_scene(scene, cfg.bands, cfg.wl_min, cfg.wl_max)
    r = _renderer(flat, cfg, 16, 16, gpu)
    acc = r.render(n_blocks=2, max_depth=6, seed=0)
    assert int((acc != 0).sum()) == 0
    assert r.stats()["segments"] > 0


def test_running_mean_blocks_split(box_small, gpu):
    cfg, flat, ora = box_small
    r = _renderer(flat, cfg, 8, 8, gpu)
    a = r.render(n_blocks=6, max_depth=cfg.depth, seed=3).clone()
    r.clear_pixel_space()
    for _ in range(3):
        r.render(n_blocks=2, max_depth=cfg.depth, seed=3)
    assert np.allclose(a.cpu().numpy(), r.accum.cpu().numpy(), rtol=1e-5, atol=1e-7)


# x2 upsampled (the reference's RESOLUTION_SCALE, main.py:16); a ratio that
# is not an integer, with the render grid larger than one tile; the display
# equal to the render grid (a partial last tile)
@pytest.mark.parametrize("W,H,OW,OH", [(10, 6, 20, 12), (23, 17, 61, 40), (37, 29, 37, 29)])
def test_resolve_matches_color_rays_closed_form(box_small, gpu, W, H, OW, OH):
    import torch

    from spectral.data import srgb_colour_system

    cfg, flat, _ = box_small
    r = _renderer(flat, cfg, W, H, gpu)
    rng = np.random.default_rng(0)
    acc = rng.uniform(0, 2, (W, H, cfg.bands)).astype(np.float32)
    r.accum.copy_(torch.from_numpy(acc))
    cs = srgb_colour_system(cfg.wl_min, cfg.wl_max, cfg.bands)
    out = r.color_rays(cs, exposure=1.0, gamma=2.0, out_width=OW, out_height=OH).cpu().numpy()
    # color_rays (main.py:1609-1629) in float64
    step = (cfg.wl_max - cfg.wl_min) / (cfg.bands - 1)
    norm = np.sum(cs.cmf[:, 1]) * step
    ri = (np.arange(OW) / OW * W).astype(int)
    rj = (np.arange(OH) / OH * H).astype(int)
    a = acc[ri][:, rj].astype(np.float64)
    xyz = a @ cs.cmf * (step / norm)
    rgb = xyz @ cs.T.T
    ldr = np.power(rgb / (rgb + 1.0), 0.5)
    fin = np.isfinite(ldr)
    assert np.allclose(out[fin], ldr[fin], rtol=1e-4, atol=1e-5)
    assert np.array_equal(np.isnan(out), ~fin)
    if (OW, OH) != (W, H):
        # an upsampled display is the render-grid resolve gathered, bit for bit
        base = r.color_rays(cs, exposure=1.0, gamma=2.0).cpu().numpy()
        ri32 = (np.arange(OW, dtype=np.float32) / np.float32(OW) * np.float32(W)).astype(int)
        rj32 = (np.arange(OH, dtype=np.float32) / np.float32(OH) * np.float32(H)).astype(int)
        assert np.array_equal(out, base[ri32][:, rj32], equal_nan=True)


@pytest.mark.parametrize("OW,OH", [(33, 9), (66, 18)])
def test_resolve_two_bands(gpu, OW, OH):
    """The fewest bands a handle takes (2), through the LDS-staged resolve
    tile over a partial last tile, at the render grid and upsampled."""
    import torch

    from spectral.data import srgb_colour_system

    cfg = configs.box_icosphere(bands=2)
    flat = flatten_scene(cfg.scene, cfg.bands, cfg.wl_min, cfg.wl_max)
    W, H = 33, 9
    r = _renderer(flat, cfg, W, H, gpu)
    acc = np.random.default_rng(1).uniform(0, 2, (W, H, 2)).astype(np.float32)
    r.accum.copy_(torch.from_numpy(acc))
    cs = srgb_colour_system(cfg.wl_min, cfg.wl_max, 2)
    out = r.color_rays(cs, exposure=1.0, gamma=2.0, out_width=OW, out_height=OH).cpu().numpy()
    step = cfg.wl_max - cfg.wl_min
    ri = (np.arange(OW) / OW * W).astype(int)
    rj = (np.arange(OH) / OH * H).astype(int)
    xyz = acc[ri][:, rj].astype(np.float64) @ cs.cmf * (step / (np.sum(cs.cmf[:, 1]) * step))
    rgb = xyz @ cs.T.T
    with np.errstate(invalid="ignore"):  # a negative channel is NaN, as in the kernel
        ldr = np.power(rgb / (rgb + 1.0), 0.5)
    fin = np.isfinite(ldr)
    assert fin.any()
    assert np.allclose(out[fin], ldr[fin], rtol=1e-4, atol=1e-5)
    assert np.array_equal(np.isnan(out), ~fin)


def test_upsampled_resolve_after_resizing(box_small, gpu):
    """One handle, render grids of growing size: the display-size scratch
    image grows with them and every upsampled resolve is the gathered
    render-grid resolve."""
    import torch

    from spectral.data import srgb_colour_system

    cfg, flat, _ = box_small
    r = _renderer(flat, cfg, 8, 6, gpu)
    cs = srgb_colour_system(cfg.wl_min, cfg.wl_max, cfg.bands)
    rng = np.random.default_rng(1)
    for W, H, OW, OH in ((8, 6, 16, 12), (20, 14, 40, 28), (9, 7, 30, 11)):
        r.set_resolution(W, H)
        r.accum.copy_(torch.from_numpy(rng.uniform(0, 2, (W, H, cfg.bands)).astype(np.float32)))
        up = r.color_rays(cs, out_width=OW, out_height=OH).cpu().numpy()
        base = r.color_rays(cs).cpu().numpy()
        ri = (np.arange(OW, dtype=np.float32) / np.float32(OW) * np.float32(W)).astype(int)
        rj = (np.arange(OH, dtype=np.float32) / np.float32(OH) * np.float32(H)).astype(int)
        assert np.array_equal(up, base[ri][:, rj], equal_nan=True), (W, H, OW, OH)


def test_upload_rejects_bad_material_ids(gpu):
    from spectral._lib import SpectralRenderError
    from spectral.render import Renderer

    cfg = configs.box_icosphere()
    flat = flatten_scene(cfg.scene, cfg.bands, cfg.wl_min, cfg.wl_max)
    flat.tri_mat = flat.tri_mat.copy()
    flat.tri_mat[3] = 99
    with pytest.raises(SpectralRenderError, match="EINVAL"):
        Renderer(flat, cfg.bands, cfg.wl_min, cfg.wl_max, device=gpu)


def test_render_argument_limits(gpu):
    """sr_render rejects what its kernels cannot index, before any launch:
    more than 2^31 - 1 work items (W x H x n_blocks x ceil(bands / 8): a
    work item is up to 8 bands of one pixel and block) in one call, negative
    blocks or depth, stream ids beyond 24 bits (the RNG counter field) and
    a zero resolution; a call without a scene is out of order.  Zero blocks
    is a no-op that leaves the accumulator untouched."""
    import ctypes

    from spectral import _lib
    from spectral._lib import SpectralRenderError, ptr
    from spectral.render import Renderer

    cfg = configs.box_icosphere()
    flat = flatten_scene(cfg.scene, cfg.bands, cfg.wl_min, cfg.wl_max)
    r = Renderer(flat, cfg.bands, cfg.wl_min, cfg.wl_max, width=8, height=8, device=gpu)
    acc = ptr(r.accum)
    bad = [
        ((8, 8, 0, 0, 0, (1 << 31) // (64 * -(-cfg.bands // 8)) + 1, 4, 0), "2\\^31"),
        ((8, 8, 0, 0, -1, 1, 4, 0), "block range"),
        ((8, 8, 0, 0, 0, -1, 4, 0), "block range"),
        ((8, 8, 0, 0, 0, 1, -1, 0), "max_depth"),
        ((8, 8, 0, 1 << 24, 0, 1, 4, 0), "24 bits"),
        ((0, 8, 0, 0, 0, 1, 4, 0), "resolution"),
    ]
    for args, msg in bad:
        with pytest.raises(SpectralRenderError, match=msg):
            _lib.call("sr_render", r._h, *args, acc)
    r.accum.fill_(7.0)
    _lib.call("sr_render", r._h, 8, 8, 0, 0, 0, 0, 4, 0, acc)
    assert bool((r.accum == 7.0).all())
    h = ctypes.c_void_p()
    _lib.call("sr_create", cfg.bands, float(cfg.wl_min), float(cfg.wl_max), ctypes.byref(h))
    try:
        with pytest.raises(SpectralRenderError, match="ESTATE"):
            _lib.call("sr_render", h, 8, 8, 0, 0, 0, 1, 4, 0, acc)
    finally:
        _lib.call("sr_destroy", h)


def test_empty_scene_renders_black(gpu):
    B = 4
    scene = {"materials": [configs._mat(B, n=1.0, tv=1)], "meshes": [],
             "camera": {"origin": [0, 0, 1], "lookat": [0, 0, 0], "up": [0, 1, 0]}}
    cfg = configs.RenderConfig("empty", scene, B, 400.0, 700.0, 8, 8, 1, 4, 0)
    flat = flatten_scene(scene, B, 400.0, 700.0)
    r = _renderer(flat, cfg, 8, 8, gpu)
    acc = r.render(n_blocks=1, max_depth=4, seed=0)
    assert int((acc != 0).sum()) == 0
    assert r.stats()["segments"] == 8 * 8 * B  # one escaping segment per path


def test_volume_fluorescence_matched_seed_parity(gpu):
    """The HG medium + fluorescent slab scene (volume fluorescence shifts
    bands) at reduced resolution: matched-seed parity as above, the same
    gate at 4096 blocks."""
    cfg = configs.volume_fluor()
    flat = flatten_scene(cfg.scene, cfg.bands, cfg.wl_min, cfg.wl_max)
    ora = OracleScene(flat)
    W = H = 12
    n_chunks, per = 8, 512  # 4096 spp
    cpu = _chunks_oracle(ora, W, H, cfg.seed, cfg.depth, n_chunks, per)
    r = _renderer(flat, cfg, W, H, gpu)
    g = _chunks_gpu(r, cfg.seed, cfg.depth, n_chunks, per)
    mc, mg = cpu.mean(0), g.mean(0)
    sig = cpu.std(0, ddof=1) / np.sqrt(n_chunks)
    diff = np.abs(mg - mc)
    ok = diff <= 3 * sig + 1e-6 * np.abs(mc) + 1e-7
    assert ok.mean() >= 0.997, f"{(~ok).sum()} of {ok.size} pixel-bands beyond 3 sigma"
    lit = mc > 1e-3
    assert (diff[lit] / mc[lit]).mean() < 0.01
    # fluorescence moved energy into longer bands: some band > 0 receives
    # light although the light is flat (the check is on the matched sum)
    assert mg.sum() == pytest.approx(mc.sum(), rel=0.01)


def test_many_lights_matched_seed_parity(gpu):
    """The icosphere of the box scene made an emitter in bands >= 8 only:
    1282 lights in those bands, 2 below, so the per-band light lists, CDFs
    and slots differ by band.  The GPU picks lights by binary search over
    the CDF, the oracle by the reference's linear scan (main.py:1310-1313);
    both pick the same light for the same uniform, so matched seeds agree."""
    cfg = configs.box_icosphere()
    B = cfg.bands
    cfg.scene["materials"][5] = configs.emitting_sphere_material(B, cfg.wl_min, cfg.wl_max)
    flat = flatten_scene(cfg.scene, B, cfg.wl_min, cfg.wl_max)
    ora = OracleScene(flat)
    W = H = 8
    n_chunks, per = 8, 256
    cpu = _chunks_oracle(ora, W, H, cfg.seed, cfg.depth, n_chunks, per)
    r = _renderer(flat, cfg, W, H, gpu)
    assert r.scene_info()["max_lights_per_band"] == 1282
    g = _chunks_gpu(r, cfg.seed, cfg.depth, n_chunks, per)
    mc, mg = cpu.mean(0), g.mean(0)
    sig = cpu.std(0, ddof=1) / np.sqrt(n_chunks)
    diff = np.abs(mg - mc)
    ok = diff <= 3 * sig + 1e-6 * np.abs(mc) + 1e-7
    assert ok.mean() >= 0.997, f"{(~ok).sum()} of {ok.size} pixel-bands beyond 3 sigma"
    lit = mc > 1e-3
    assert (diff[lit] / mc[lit]).mean() < 0.01


@pytest.mark.parametrize("make", [
    lambda: configs.box_icosphere(width=32, height=32, spp=4),
    lambda: configs.heightfield_1m(width=24, height=16, spp=3, nx=64, nz=32),
    lambda: configs.volume_fluor(width=16, height=16, spp=2),
], ids=["box_icosphere", "heightfield", "volume_fluor"])
def test_end_dead_paths_keeps_the_image(gpu, make):
    """SR_RENDER_END_DEAD_PATHS ends a path once its throughput is exactly 0:
    the same image (each path's later contributions are all 0), fewer
    segments, and none of the non-finite 0 * inf products the kernel drops;
    alone and as a batch job."""
    from spectral.batch import Batch

    cfg = make()
    flat = flatten_scene(cfg.scene, cfg.bands, cfg.wl_min, cfg.wl_max)
    full = _renderer(flat, cfg, cfg.width, cfg.height, gpu)
    ref = full.render(n_blocks=cfg.spp, max_depth=cfg.depth, seed=cfg.seed).cpu().numpy()
    st_ref = full.stats()
    cut = _renderer(flat, cfg, cfg.width, cfg.height, gpu)
    got = cut.render(n_blocks=cfg.spp, max_depth=cfg.depth, seed=cfg.seed, end_dead_paths=True).cpu().numpy()
    st = cut.stats()
    assert np.allclose(got, ref, rtol=1e-5, atol=1e-7)
    assert 0 < st["segments"] < st_ref["segments"]
    assert st["nonfinite"] <= st_ref["nonfinite"]
    bat = _renderer(flat, cfg, cfg.width, cfg.height, gpu)
    b = Batch()
    b.submit(bat, cfg.spp, cfg.depth, cfg.seed, end_dead_paths=True)
    out = b.run(graph=False)
    assert np.allclose(bat.accum.cpu().numpy(), got, rtol=1e-5, atol=1e-7)
    assert int(out["segments"][0]) == st["segments"]
