"""Batched multi-scene launch (sr_batch_*): one persistent kernel over
several scenes must reproduce each scene's own sr_render result (same RNG
keys per path), with and without hipGraph capture/replay."""
import numpy as np
import pytest

from spectral import configs
from spectral.scene import flatten_scene

pytestmark = pytest.mark.gpu


def _scenes():
    c0 = configs.box_icosphere(width=32, height=32, spp=4)
    c1 = configs.heightfield_1m(width=24, height=16, spp=3, nx=64, nz=32)
    c2 = configs.volume_fluor(width=16, height=16, spp=2)
    out = []
    for c in (c0, c1, c2):
        out.append((c, flatten_scene(c.scene, c.bands, c.wl_min, c.wl_max)))
    return out


def _renderers(gpu):
    from spectral.render import Renderer

    return [(c, Renderer(f, c.bands, c.wl_min, c.wl_max, width=c.width, height=c.height, device=gpu))
            for c, f in _scenes()]


def test_batch_matches_sequential(gpu):
    from spectral.batch import Batch

    seq = _renderers(gpu)
    ref, segs = [], []
    for c, r in seq:
        ref.append(r.render(n_blocks=c.spp, max_depth=c.depth, seed=c.seed).cpu().numpy())
        segs.append(r.stats()["segments"])
    bat = _renderers(gpu)
    b = Batch()
    for c, r in bat:
        b.submit(r, c.spp, c.depth, c.seed)
    for graph in (False, True, True):
        for _, r in bat:
            r.clear_pixel_space()
        out = b.run(graph=graph)
        for k, (c, r) in enumerate(bat):
            got = r.accum.cpu().numpy()
            assert np.allclose(got, ref[k], rtol=1e-5, atol=1e-6), (graph, c.name)
        assert list(out["segments"]) == segs
        lat = out["latency_ms"]
        assert np.all(lat > 0) and lat.max() <= out["total_ms"] * 1.05 + 0.5


def test_batch_tables_in_lds_and_in_global_memory_agree(gpu):
    """256-lane blocks keep every job's small scene tables in LDS (the
    block's waves switch jobs together); 64-lane blocks get a quarter of that
    LDS share, too little for these scenes' tables, so the same batch reads
    them from global memory.  Both give each scene the same image and the
    same segment count."""
    from spectral.batch import Batch

    # the materials table alone ([M][B] records of 32 B) of the box scene
    # outgrows a 64-lane block's 3,328-byte share of the 13,312-byte budget
    c0, f0 = _scenes()[0]
    assert f0.mat_f32.shape[0] * c0.bands * 32 > 13312 * 64 // 256
    outs = []
    for block in (256, 64):
        bat = _renderers(gpu)
        b = Batch()
        b.set_config(block, 100, 64)
        for c, r in bat:
            r.clear_pixel_space()
            b.submit(r, c.spp, c.depth, c.seed)
        out = b.run(graph=False)
        outs.append(([r.accum.cpu().numpy() for _, r in bat], list(out["segments"])))
    (img_l, seg_l), (img_g, seg_g) = outs
    assert seg_l == seg_g
    for a, g in zip(img_l, img_g):
        assert np.allclose(a, g, rtol=1e-5, atol=1e-6)


def test_batch_config_and_tuner(gpu):
    from spectral.batch import Batch, tune

    bat = _renderers(gpu)
    spec = [(r, 1, c.depth, c.seed) for c, r in bat]
    best, timings = tune(spec, blocks=(128, 256), grid_pcts=(100,), chunks=(64, 512), reps=1)
    assert best in timings and len(timings) == 4
    b = Batch()
    with pytest.raises(Exception):
        b.set_config(96, 100, 64)
    b.set_config(*best)
    for c, r in bat:
        r.clear_pixel_space()
        b.submit(r, 1, c.depth, c.seed)
    with pytest.raises(Exception):
        b.submit(bat[0][1], 1, 4, 0)  # a handle may appear once per batch
    out = b.run()
    assert all(s > 0 for s in out["segments"])


def test_batch_with_render_modes(gpu):
    """A light-tree job, a subsurface job and a plain job in one launch (the
    walk kernel, walking only in the job that asks for it) reproduce their
    own sr_render calls with the same flags."""
    from spectral.batch import Batch
    from spectral.render import Renderer

    cfgs = [(configs.lamps(width=16, height=16), {"light_tree": True}),
            (configs.subsurface_slab(width=16, height=16), {"subsurface": True}),
            (configs.box_icosphere(width=16, height=16), {})]

    def make():
        out = []
        for c, kw in cfgs:
            f = flatten_scene(c.scene, c.bands, c.wl_min, c.wl_max)
            out.append((c, kw, Renderer(f, c.bands, c.wl_min, c.wl_max, width=c.width, height=c.height, device=gpu)))
        return out

    ref, segs = [], []
    for c, kw, r in make():
        ref.append(r.render(n_blocks=4, max_depth=c.depth, seed=c.seed, **kw).cpu().numpy())
        segs.append(r.stats()["segments"])
    bat = make()
    b = Batch()
    for c, kw, r in bat:
        b.submit(r, 4, c.depth, c.seed, **kw)
    out = b.run()
    for k, (c, kw, r) in enumerate(bat):
        assert np.allclose(r.accum.cpu().numpy(), ref[k], rtol=1e-5, atol=1e-6), c.name
    assert list(out["segments"]) == segs
