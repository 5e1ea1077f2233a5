"""Host-runtime contracts of the C ABI that span several calls: a stream
change orders the handle's work (sr_set_stream), a batch launched again
continues from where its scenes are (sr_batch_* with Batch.run) and sees a
changed camera, sr_update_materials refuses a change of surface_emission
(the light lists were built from it), RenderFunction's backward replays
the paths of its own forward even after another forward ran, and a render
resumed from a checkpoint in a fresh renderer continues the same estimate."""
import ctypes

import numpy as np
import pytest
import torch

from spectral import _lib, configs
from spectral.render import Renderer
from spectral.scene import flatten_scene

pytestmark = pytest.mark.gpu


def _box(W=24, H=16):
    c = configs.box_icosphere(width=W, height=H)
    return c, flatten_scene(c.scene, c.bands, c.wl_min, c.wl_max)


def _new(c, f, gpu):
    return Renderer(f, c.bands, c.wl_min, c.wl_max, width=c.width, height=c.height, device=gpu)


def test_stream_switch_orders_work(gpu):
    c, f = _box()
    ref = _new(c, f, gpu)
    ref.render(n_blocks=8, max_depth=c.depth, seed=c.seed)
    want = ref.accum.cpu().numpy()
    r = _new(c, f, gpu)
    r.render(n_blocks=4, max_depth=c.depth, seed=c.seed)  # default stream, not waited for
    side = torch.cuda.Stream(device=gpu)
    with torch.cuda.stream(side):
        r.render(n_blocks=4, max_depth=c.depth, seed=c.seed)  # same block sums, other stream
        got_dev = r.accum.clone()
    side.synchronize()
    got = got_dev.cpu().numpy()
    assert np.allclose(got, want, rtol=1e-5, atol=1e-6)
    # and back to the default stream
    r.render(n_blocks=4, max_depth=c.depth, seed=c.seed)
    ref.render(n_blocks=4, max_depth=c.depth, seed=c.seed)
    assert np.allclose(r.accum.cpu().numpy(), ref.accum.cpu().numpy(), rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("graph", [False, True])
def test_batch_run_twice_continues(gpu, graph):
    from spectral.batch import Batch

    c, f = _box()
    seq = _new(c, f, gpu)
    seq.render(n_blocks=3, max_depth=c.depth, seed=c.seed)
    seq.render(n_blocks=3, max_depth=c.depth, seed=c.seed)
    r = _new(c, f, gpu)
    b = Batch()
    b.submit(r, 3, c.depth, c.seed)
    b.run(graph)
    b.run(graph)
    assert r.blocks_done == 6
    assert np.allclose(r.accum.cpu().numpy(), seq.accum.cpu().numpy(), rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("graph", [False, True])
def test_batch_sees_camera_change(gpu, graph):
    """A camera set through the C ABI after the job table (and graph) were
    made: the next launch renders the new view."""
    from spectral.batch import Batch

    c, f = _box()
    cam = np.array([0.05, 0.1, 1.6, 0.0, -0.1, 0.0, 0.0, 1.0, 0.0], np.float32)
    want = _new(c, f, gpu)
    _lib.call("sr_set_camera", want._h, _lib.ptr(cam))
    want.render(n_blocks=2, max_depth=c.depth, seed=c.seed)
    r = _new(c, f, gpu)
    b = Batch()
    b.submit(r, 2, c.depth, c.seed)
    b.launch(graph)
    b.wait()
    _lib.call("sr_set_camera", r._h, _lib.ptr(cam))
    r.accum.zero_()
    b.launch(graph)  # same submission: the batch re-fills its job table
    b.wait()
    assert np.allclose(r.accum.cpu().numpy(), want.accum.cpu().numpy(), rtol=1e-5, atol=1e-6)


def test_batch_after_resolution_change(gpu):
    """set_resolution reallocates the accumulator: the next run renders into
    the new one (Batch.run submits again), at the new size."""
    from spectral.batch import Batch

    c, f = _box()
    r = _new(c, f, gpu)
    b = Batch()
    b.submit(r, 2, c.depth, c.seed)
    b.run()
    r.set_resolution(40, 32)
    r2 = Renderer(f, c.bands, c.wl_min, c.wl_max, width=40, height=32, device=gpu)
    r2.render(n_blocks=2, max_depth=c.depth, seed=c.seed)
    b.run()
    assert r.accum.shape == (40, 32, c.bands)
    assert np.allclose(r.accum.cpu().numpy(), r2.accum.cpu().numpy(), rtol=1e-5, atol=1e-6)


def test_update_materials_refuses_emission_change(gpu):
    c, f = _box()
    r = _new(c, f, gpu)
    mi = f.mat_i32.copy()
    mi[1, 1, 3] = 1  # a wall starts to emit in band 3
    with pytest.raises(_lib.SpectralRenderError, match="surface_emission"):
        r.update_materials(f.mat_f32, mat_i32=mi)
    # other properties may change
    mf = f.mat_f32.copy()
    mf[1, 2, :] = 0.3
    r.update_materials(mf)
    r.render(n_blocks=1, max_depth=c.depth, seed=c.seed)
    assert np.isfinite(r.accum.cpu().numpy()).all()


def test_autograd_backward_replays_its_own_forward(gpu):
    from spectral.autograd import DifferentiableScene, inverse_volume_spec, render

    cfg = configs.inverse_volume(width=16, height=16, spp=16, depth=8)
    flat = flatten_scene(cfg.scene, cfg.bands, cfg.wl_min, cfg.wl_max)
    ds = DifferentiableScene(flat, inverse_volume_spec(), cfg.width, cfg.height, cfg.spp, cfg.depth, cfg.seed,
                             device=gpu)
    th0 = ds.initial_params()
    d = torch.rand((cfg.width, cfg.height, cfg.bands), generator=torch.Generator().manual_seed(3)).to(gpu)
    a = torch.tensor(th0, dtype=torch.float32, requires_grad=True)
    (render(a, ds) * d).sum().backward()
    alone = a.grad.clone()
    b1 = torch.tensor(th0, dtype=torch.float32, requires_grad=True)
    b2 = torch.tensor(th0 * 0.8, dtype=torch.float32, requires_grad=True)
    out1 = render(b1, ds)
    out2 = render(b2, ds)  # uploads other parameters
    (out1 * d).sum().backward()
    assert torch.allclose(b1.grad, alone, rtol=1e-4, atol=1e-6)
    (out2 * d).sum().backward()
    assert not torch.allclose(b2.grad, alone, rtol=1e-3, atol=1e-6)


def test_resume_continues_the_same_estimate(gpu, tmp_path):
    c, f = _box()
    a = _new(c, f, gpu)
    a.render(n_blocks=3, max_depth=c.depth, seed=11, stream=1)
    a.checkpoint(tmp_path / "ck.npz")
    a.render(n_blocks=2, max_depth=c.depth, seed=11, stream=1)
    b = Renderer(f, c.bands, c.wl_min, c.wl_max, device=gpu)  # a fresh handle, no resolution yet
    args = b.resume(tmp_path / "ck.npz")
    assert args == {"seed": 11, "stream": 1, "max_depth": c.depth, "light_tree": False, "subsurface": False,
                    "end_dead_paths": False}
    assert (b.width, b.height, b.blocks_done) == (c.width, c.height, 3)
    b.render(n_blocks=2, **args)
    assert b.blocks_done == 5
    # blocks 3 and 4 are the same paths in both; only the order of the float
    # atomics into the block sums can differ
    np.testing.assert_allclose(b.accum.cpu().numpy(), a.accum.cpu().numpy(), rtol=1e-5, atol=1e-7)
    assert float(a.accum.abs().sum()) > 0.0
