"""Path-replay adjoint (sr_render_adjoint) against central finite differences
of the forward kernel with common random numbers (every render of a
comparison uses the same seed and stream), on the volume/fluorescence
inverse-render scene at reduced resolution.  The reference has no autodiff,
so finite differences of the forward estimator are the oracle for
d E[image] / d theta.

Albedo enters only as a factor: central differences with h = 1e-3 on one
stream are precise, and the adjoint must agree within 5%.  Roughness, g and
sigma_a also move the sampled directions, free-flight distances and branch
choices; with common random numbers a perturbation flips a branch in a
fraction ~h of paths, and each flip contributes O(path radiance / h), so the
h = 1e-3 difference is dominated by noise.  For those the comparison uses
several independent streams and a larger step (1e-2 for roughness, 3e-2 for
g and sigma_a; the curvature bias at these steps is below the tolerance),
and requires agreement within 5% plus three combined standard errors."""
import numpy as np
import pytest
import torch

from spectral import configs
from spectral.autograd import DifferentiableScene, inverse_volume_spec, render
from spectral.scene import flatten_scene

pytestmark = pytest.mark.gpu

STREAMS = 6


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
    g_adj = np.stack([ds.adjoint(d_acc, stream=s).cpu().numpy().astype(np.float64) for s in range(STREAMS)])
    return ds, theta0, d_acc, g_adj


def _loss(ds, theta, d_acc, stream):
    ds.set_params(theta)
    acc = ds.render(stream=stream)
    return float((acc.double() * d_acc.double()).sum())


def _fd(ds, theta0, d_acc, i, h, stream):
    tp = theta0.copy()
    tm = theta0.copy()
    tp[i] += h
    tm[i] -= h
    return (_loss(ds, tp, d_acc, stream) - _loss(ds, tm, d_acc, stream)) / (2 * h)


def _agree(adj, fd):
    """Means over streams agree within 5% plus 3 combined standard errors."""
    a, f = adj.mean(0), fd.mean(0)
    se = np.sqrt(adj.var(0, ddof=1) / len(adj) + fd.var(0, ddof=1) / len(fd))
    return np.linalg.norm(a - f) <= 0.05 * np.linalg.norm(f) + 3 * np.linalg.norm(se), (a, f, se)


def test_param_count_and_roundtrip(setup):
    ds, theta0, _, _ = setup
    assert ds.n_params == 126
    ds.set_params(theta0)
    assert np.allclose(ds.initial_params(), theta0, rtol=1e-5, atol=1e-9)


def test_albedo_gradient_matches_fd_h1e3(setup):
    ds, theta0, d_acc, g_adj = setup
    idx = list(range(0, 93, 4))
    fd = np.array([_fd(ds, theta0, d_acc, i, 1e-3, 0) for i in idx])
    ga = g_adj[0, idx]
    rel = np.linalg.norm(ga - fd) / np.linalg.norm(fd)
    assert rel < 0.05, (rel, ga[:6], fd[:6])


@pytest.mark.parametrize("name,i,h", [("roughness", 93, 1e-2), ("g", 94, 3e-2)])
def test_scalar_gradient_matches_fd(setup, name, i, h):
    ds, theta0, d_acc, g_adj = setup
    fd = np.array([[_fd(ds, theta0, d_acc, i, h, s)] for s in range(STREAMS)])
    ok, info = _agree(g_adj[:, [i]], fd)
    assert ok, (name, info)


def test_sigma_a_gradient_matches_fd(setup):
    ds, theta0, d_acc, g_adj = setup
    idx = [i for i in range(95, 126) if theta0[i] > 0.1][::2]
    fd = np.array([[_fd(ds, theta0, d_acc, i, 3e-2, s) for i in idx] for s in range(STREAMS)])
    ok, info = _agree(g_adj[:, idx], fd)
    assert ok, info


def test_autograd_function(setup):
    ds, theta0, d_acc, g_adj = setup
    th = torch.tensor(theta0, dtype=torch.float32, requires_grad=True)
    out = render(th, ds)
    (out * d_acc).sum().backward()
    assert th.grad is not None and th.grad.shape == (126,)
    assert np.allclose(th.grad.numpy(), g_adj[0], rtol=1e-4, atol=1e-6)
