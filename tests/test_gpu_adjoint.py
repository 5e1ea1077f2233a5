"""Path-replay adjoint (sr_render_adjoint) against central finite differences
of the forward kernel with common random numbers (every render of a
comparison uses the same seed and stream), on the volume/fluorescence
inverse-render scene at reduced resolution.  The reference has no autodiff,
so finite differences of the forward estimator are the oracle for
d E[image] / d theta.

Albedo enters only as a factor: central differences with h = 1e-3 on one
stream are precise, and all 93 albedo gradients must agree within 5%.
Roughness, g and sigma_a also move the sampled directions, free-flight
distances and branch choices; with common random numbers a perturbation
flips a branch in a fraction ~h of paths, each flip contributing
O(path radiance / h), so one stream's difference is noisy.  For those the
comparison averages 160-224 independent streams (64x64, 256 blocks each) of
both the adjoint and the difference (h = 1e-2 for roughness, 3e-2 for g, a
3e-2 relative step of all 31 sigma_a together), and requires the means to
agree within 5% with no noise allowance: the test first checks that three
standard errors of the difference of the means are below 2.5% of the
derivative, so a 5% bound is a bound on the bias."""
import numpy as np
import pytest
import torch

from spectral import configs
from spectral.autograd import DifferentiableScene, inverse_volume_spec, render
from spectral.scene import flatten_scene

pytestmark = pytest.mark.gpu

STREAMS = {"roughness": 160, "g": 224, "sigma_a": 160}  # g's differences are the noisiest
NOISE_BOUND = 0.025  # 3 standard errors, relative to the derivative
BIAS_BOUND = 0.05


def _scene(gpu, W, spp, fluorescent_medium=False):
    cfg = configs.inverse_volume(width=W, height=W, spp=spp, depth=12, seed=3)
    if fluorescent_medium:
        # The slab's fluorescent light never leaves the slab under the
        # reference's rules (its inner surface sees n_vol = n_surf, so the
        # rough BTDF gives 0 or 0/0 there): the index-matched medium sphere
        # gets the EEM as well, so that some paths move light to other bands.
        cfg.scene["materials"][5]["eem_id"] = 1
    flat = flatten_scene(cfg.scene, cfg.bands, cfg.wl_min, cfg.wl_max)
    ds = DifferentiableScene(flat, inverse_volume_spec(), cfg.width, cfg.height, cfg.spp, cfg.depth, cfg.seed,
                             device=gpu)
    gen = torch.Generator(device="cpu").manual_seed(7)
    d_acc = torch.rand((cfg.width, cfg.height, cfg.bands), generator=gen).to(gpu)
    return ds, d_acc


@pytest.fixture(scope="module")
def setup(gpu):
    ds, d_acc = _scene(gpu, 24, 256)
    theta0 = ds.initial_params()
    ds.set_params(theta0)
    g_adj = ds.adjoint(d_acc, stream=0).cpu().numpy().astype(np.float64)[None]
    return ds, theta0, d_acc, g_adj


@pytest.fixture(scope="module")
def streams(gpu):
    ds, d_acc = _scene(gpu, 64, 256)
    theta0 = ds.initial_params()
    ds.set_params(theta0)
    n = max(STREAMS.values())
    g_adj = np.stack([ds.adjoint(d_acc, stream=s).cpu().numpy().astype(np.float64) for s in range(n)])
    return ds, theta0, d_acc, g_adj


def _loss(ds, theta, d_acc, stream):
    ds.set_params(theta)
    acc = ds.render(stream=stream)
    return float((acc.double() * d_acc.double()).sum())


def _fd(ds, theta0, d_acc, v, h, stream):
    """Central difference along v with step h (common random numbers)."""
    return (_loss(ds, theta0 + h * v, d_acc, stream) - _loss(ds, theta0 - h * v, d_acc, stream)) / (2 * h)


def test_param_count_and_roundtrip(setup):
    ds, theta0, _, _ = setup
    assert ds.n_params == 126
    ds.set_params(theta0)
    assert np.allclose(ds.initial_params(), theta0, rtol=1e-5, atol=1e-9)


def test_albedo_gradient_matches_fd_h1e3(setup):
    ds, theta0, d_acc, g_adj = setup
    fd = np.array([_fd(ds, theta0, d_acc, np.eye(126)[i], 1e-3, 0) for i in range(93)])
    ga = g_adj[0, :93]
    rel = np.linalg.norm(ga - fd) / np.linalg.norm(fd)
    worst = np.max(np.abs(ga - fd) / np.maximum(np.abs(fd), 1e-3 * np.abs(fd).max()))
    print(f"albedo (93 parameters, h = 1e-3): relative error {rel:.2e}, worst single {worst:.2e}")
    assert rel < BIAS_BOUND, (rel, ga[:6], fd[:6])


def _direction(theta0, name):
    v = np.zeros(126)
    if name == "roughness":
        v[93] = 1.0
    elif name == "g":
        v[94] = 1.0
    else:  # all sigma_a together, each scaled by its own value
        v[95:] = theta0[95:]
        v /= np.linalg.norm(v)
    return v


@pytest.mark.parametrize("name,h", [("roughness", 1e-2), ("g", 3e-2), ("sigma_a", 3e-2)])
def test_gradient_matches_fd_without_noise_allowance(streams, name, h):
    ds, theta0, d_acc, g_adj = streams
    v = _direction(theta0, name)
    n = STREAMS[name]
    adj = g_adj[:n] @ v
    fd = np.array([_fd(ds, theta0, d_acc, v, h, s) for s in range(n)])
    a, f = adj.mean(), fd.mean()
    se = np.sqrt(adj.var(ddof=1) / n + fd.var(ddof=1) / n)
    rel, noise = abs(a - f) / abs(f), 3 * se / abs(f)
    print(f"{name}: adjoint {a:.5g}, finite difference {f:.5g} (h = {h}), relative error {rel:.4f}, "
          f"3 standard errors {noise:.4f} ({n} streams)")
    assert noise <= NOISE_BOUND, (name, noise)
    assert rel <= BIAS_BOUND, (name, rel)


def test_autograd_function(setup):
    ds, theta0, d_acc, g_adj = setup
    th = torch.tensor(theta0, dtype=torch.float32, requires_grad=True)
    out = render(th, ds)
    (out * d_acc).sum().backward()
    assert th.grad is not None and th.grad.shape == (126,)
    assert np.allclose(th.grad.numpy(), g_adj[0], rtol=1e-4, atol=1e-6)


def test_autograd_without_record_replays_twice(setup):
    """A render whose path totals would exceed record_limit_bytes records
    nothing, and backward runs the two-pass adjoint: the gradient of the
    setup fixture's own two-pass call."""
    ds, theta0, d_acc, g_adj = setup
    limit = ds.record_limit_bytes
    ds.record_limit_bytes = 0
    try:
        assert ds.new_path_totals() is None
        th = torch.tensor(theta0, dtype=torch.float32, requires_grad=True)
        out = render(th, ds)
        (out * d_acc).sum().backward()
    finally:
        ds.record_limit_bytes = limit
    assert np.allclose(th.grad.numpy(), g_adj[0], rtol=1e-5, atol=1e-7)


def test_path_totals_are_the_block_images(gpu):
    """sr_render_record writes one total per path at [block][i][j][band]: on
    the box scene (no fluorescence, so every total is finite and is that
    path's whole contribution) each block's totals are the image of that
    block rendered alone, and recording leaves the image unchanged."""
    from spectral.render import Renderer

    W = H = 16
    nb = 3
    cfg = configs.box_icosphere(width=W, height=H, spp=nb, depth=8)
    flat = flatten_scene(cfg.scene, cfg.bands, cfg.wl_min, cfg.wl_max)
    B = cfg.bands
    r = Renderer(flat, B, cfg.wl_min, cfg.wl_max, width=W, height=H, device=gpu)
    tot = torch.full((nb * W * H * B,), -1.0, device=gpu)
    r.render(n_blocks=nb, max_depth=8, seed=5, path_totals=tot)
    img_rec = r.accum.clone()
    r.clear_pixel_space()
    r.render(n_blocks=nb, max_depth=8, seed=5)
    torch.testing.assert_close(img_rec, r.accum, rtol=1e-6, atol=1e-7)
    tot = tot.view(nb, W, H, B)
    assert torch.isfinite(tot).all() and (tot >= 0).all()
    assert (tot > 0).float().mean().item() > 0.3
    for b in range(nb):
        r.clear_pixel_space()
        r.blocks_done = b  # block b alone: accum = block image / (b + 1)
        r.render(n_blocks=1, max_depth=8, seed=5)
        torch.testing.assert_close(tot[b], r.accum * (b + 1), rtol=1e-5, atol=1e-7)


def test_recorded_adjoint_equals_two_pass(gpu):
    """The adjoint given the forward's path totals replays each path whose
    contributions stayed in its band once; paths that fluoresced (NaN
    totals) are replayed twice as before.  The gradient is the two-pass one
    up to float rounding (the first pass summed dL/dacc x c_k term by term,
    the record gives dL/dacc x sum c_k)."""
    ds, d_acc = _scene(gpu, 24, 64, fluorescent_medium=True)
    ds.set_params(ds.initial_params())
    tot = ds.new_path_totals()
    ds.render(path_totals=tot)
    nan_frac = torch.isnan(tot).float().mean().item()
    assert 0.0 < nan_frac < 0.5, nan_frac  # the medium's fluorescence moves some paths' light
    # paths that added no contribution are recorded as -0.0 and not replayed
    skipped = ((tot == 0) & torch.signbit(tot)).float().mean().item()
    g2 = ds.adjoint(d_acc).double()
    t2 = ds.r.last_kernel_ms()
    g1 = ds.adjoint(d_acc, path_totals=tot).double()
    t1 = ds.r.last_kernel_ms()
    rel = ((g1 - g2).norm() / g2.norm()).item()
    print(f"recorded vs two-pass adjoint: relative difference {rel:.2e}, {nan_frac:.3f} of the paths fluoresced, "
          f"{skipped:.3f} added nothing, {t1:.1f} vs {t2:.1f} ms")
    assert rel < 1e-5
    torch.testing.assert_close(g1, g2, rtol=1e-4, atol=1e-6 * g2.abs().max().item())


def test_record_refuses_light_tree_and_subsurface(gpu):
    from spectral._lib import SpectralRenderError

    ds, _ = _scene(gpu, 8, 2)
    tot = ds.new_path_totals()
    for kw in ({"light_tree": True}, {"subsurface": True}):
        with pytest.raises(SpectralRenderError, match="SR_RENDER_END_DEAD_PATHS"):
            ds.r.render(n_blocks=2, max_depth=4, path_totals=tot, **kw)
    with pytest.raises(ValueError):
        ds.r.render(n_blocks=1, max_depth=4, path_totals=tot)  # sized for 2 blocks
