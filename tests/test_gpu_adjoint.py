"""Path-replay adjoint (sr_render_adjoint) against central finite differences
of the forward kernel with common random numbers (same seed on every
render), h = 1e-3, on the volume/fluorescence inverse-render scene at
reduced resolution.  The reference has no autodiff, so finite differences of
the forward estimator are the oracle for d E[image] / d theta."""
import numpy as np
import pytest
import torch

from spectral import configs
from spectral.autograd import DifferentiableScene, inverse_volume_spec, render
from spectral.scene import flatten_scene

pytestmark = pytest.mark.gpu

H_FD = 1e-3


@pytest.fixture(scope="module")
def setup(gpu):
    cfg = configs.inverse_volume(width=24, height=24, spp=256, depth=12, seed=3)
    flat = flatten_scene(cfg.scene, cfg.bands, cfg.wl_min, cfg.wl_max)
    ds = DifferentiableScene(flat, inverse_volume_spec(), cfg.width, cfg.height, cfg.spp, cfg.depth, cfg.seed,
                             device=gpu)
    theta0 = ds.initial_params()
    gen = torch.Generator(device="cpu").manual_seed(7)
    d_acc = torch.rand((cfg.width, cfg.height, cfg.bands), generator=gen).to(gpu)
    ds.set_params(theta0)
    ds.render()
    g_adj = ds.adjoint(d_acc).cpu().numpy().astype(np.float64)
    return ds, theta0, d_acc, g_adj


def _loss(ds, theta, d_acc):
    ds.set_params(theta)
    acc = ds.render()
    return float((acc.double() * d_acc.double()).sum())


def _fd(ds, theta0, d_acc, i):
    tp = theta0.copy()
    tm = theta0.copy()
    if theta0[i] > H_FD:
        tp[i] += H_FD
        tm[i] -= H_FD
        return (_loss(ds, tp, d_acc) - _loss(ds, tm, d_acc)) / (2 * H_FD)
    tp[i] += H_FD
    return (_loss(ds, tp, d_acc) - _loss(ds, tm, d_acc)) / H_FD


def test_param_count_and_roundtrip(setup):
    ds, theta0, _, _ = setup
    assert ds.n_params == 126
    ds.set_params(theta0)
    assert np.allclose(ds.initial_params(), theta0, rtol=1e-5, atol=1e-9)


def test_albedo_gradient_matches_fd(setup):
    ds, theta0, d_acc, g_adj = setup
    sl = slice(0, 93)
    idx = range(0, 93, 4)
    fd = np.array([_fd(ds, theta0, d_acc, i) for i in idx])
    ga = g_adj[sl][list(idx)]
    rel = np.linalg.norm(ga - fd) / np.linalg.norm(fd)
    assert rel < 0.05, (rel, ga[:6], fd[:6])


def test_scalar_gradients_match_fd(setup):
    ds, theta0, d_acc, g_adj = setup
    for name, i in (("roughness", 93), ("g", 94)):
        fd = _fd(ds, theta0, d_acc, i)
        assert abs(g_adj[i] - fd) <= 0.05 * abs(fd) + 1e-3 * np.abs(g_adj).max(), (name, g_adj[i], fd)


def test_sigma_a_gradient_matches_fd(setup):
    ds, theta0, d_acc, g_adj = setup
    idx = list(range(95, 126, 3))
    fd = np.array([_fd(ds, theta0, d_acc, i) for i in idx])
    ga = g_adj[idx]
    rel = np.linalg.norm(ga - fd) / np.linalg.norm(fd)
    assert rel < 0.05, (rel, ga, fd)


def test_autograd_function(setup):
    ds, theta0, d_acc, g_adj = setup
    th = torch.tensor(theta0, dtype=torch.float32, requires_grad=True)
    out = render(th, ds)
    (out * d_acc).sum().backward()
    assert th.grad is not None and th.grad.shape == (126,)
    assert np.allclose(th.grad.numpy(), g_adj, rtol=1e-4, atol=1e-6)
