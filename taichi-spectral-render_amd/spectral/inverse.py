"""Inverse rendering on the volume/fluorescence scene: recover material
parameters from an image by gradient descent through RenderFunction.

The reference has no autodiff and no optimisation loop (no grad/Tape in
main.py); this is the use the differentiable renderer exists for.  The
recipe of the inverse-render workload:

* target: the scene rendered with its known parameters (theta*) at seed 3,
  with many spectral blocks so the target itself carries little noise;
* start: every parameter scaled by 1 +- 30% (signs drawn with seed 4) and
  clipped to its valid range;
* each iteration renders `spp` blocks on a fresh Philox stream (the
  adjoint replays exactly those paths), takes the loss and its gradient
  d loss / d accum on the host side of torch, runs the path-replay adjoint
  and takes one Adam step, then projects the parameters back into range.

Loss: the squared relative difference of tile means -- the spectral
accumulator averaged over tile x tile pixel blocks, per band -- so the
Monte Carlo noise of a 64-block forward (which no parameter can fit)
averages out while every parameter's effect on the image (wall colours per
band, the medium's brightness and shape) stays visible.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import configs
from .autograd import DifferentiableScene, inverse_volume_spec, render
from .scene import flatten_scene

# valid ranges of the parameter groups (albedo of the reference's
# s / (s + 4 pi k / lambda + 1e-6) mapping, GGX roughness, HG g, sigma_a)
BOUNDS = {"albedo": (0.02, 0.98), "roughness": (0.01, 1.0), "g": (-0.95, 0.95), "sigma_a": (0.0, 1e3)}
# Adam step sizes per group (the step Adam takes is about this per iteration)
LEARNING_RATES = {"albedo": 0.02, "roughness": 0.005, "g": 0.02, "sigma_a": 0.05}


def parameter_groups(ds: DifferentiableScene):
    """{group: index array} of the flat parameter vector."""
    B = ds.B
    g = {"albedo": [], "roughness": [], "g": [], "sigma_a": []}
    for key, off in ds.offsets.items():
        if isinstance(key, tuple):
            g["albedo"].extend(range(off, off + B))
        elif key == "sigma_a":
            g["sigma_a"].extend(range(off, off + B))
        else:
            g[key].append(off)
    return {k: np.asarray(v, np.int64) for k, v in g.items() if v}


def project(theta, groups):
    out = np.array(theta, np.float64)
    for name, idx in groups.items():
        lo, hi = BOUNDS[name]
        out[idx] = np.clip(out[idx], lo, hi)
    return out


def perturb(theta, groups, rel=0.3, seed=4):
    """theta * (1 + rel * s) with random signs s (seed), projected."""
    rng = np.random.default_rng(seed)
    s = rng.choice([-1.0, 1.0], size=len(theta))
    return project(np.asarray(theta, np.float64) * (1.0 + rel * s), groups)


def tile_means(acc, tile):
    """[W, H, B] -> [W / tile, H / tile, B] means (W, H multiples of tile)."""
    W, H, B = acc.shape
    return acc.reshape(W // tile, tile, H // tile, tile, B).mean(dim=(1, 3))


@dataclass
class History:
    loss: list = field(default_factory=list)
    param_error: list = field(default_factory=list)  # {group: relative L2 error}
    seconds: list = field(default_factory=list)


class InverseRender:
    """Gradient-descent recovery of the 126 parameters of the inverse-render
    scene (3 wall albedo spectra, slab roughness, medium g and sigma_a)."""

    def __init__(self, width=256, height=256, spp=64, depth=12, target_spp=None, tile=4, target_seed=3,
                 init_seed=4, rel=0.3, device="cuda", learning_rates=None):
        if width % tile or height % tile:
            raise ValueError("width and height must be multiples of tile")
        cfg = configs.inverse_volume(width=width, height=height, spp=spp, depth=depth, seed=target_seed)
        self.cfg = cfg
        self.tile = int(tile)
        flat = flatten_scene(cfg.scene, cfg.bands, cfg.wl_min, cfg.wl_max)
        self.ds = DifferentiableScene(flat, inverse_volume_spec(), width, height, spp, depth, target_seed, device=device)
        self.groups = parameter_groups(self.ds)
        self.theta_true = self.ds.initial_params()
        self.lr = dict(LEARNING_RATES if learning_rates is None else learning_rates)
        # target: the true parameters, seed 3, more blocks than a forward
        self.ds.set_params(self.theta_true)
        self.ds.spp = int(target_spp if target_spp is not None else 8 * spp)
        self.target = tile_means(self.ds.render(stream=0).clone(), self.tile)
        self.ds.spp = int(spp)
        # relative error per tile and band; the floor keeps dark tiles (whose
        # relative noise is largest) from dominating
        floor = 0.1 * float(self.target.mean())
        self.weight = 1.0 / (self.target * self.target + floor * floor)
        self.theta0 = perturb(self.theta_true, self.groups, rel, init_seed)
        self.theta = self.theta0.copy()
        self._m = np.zeros_like(self.theta)
        self._v = np.zeros_like(self.theta)
        self._t = 0
        self._stream = 1  # stream 0 rendered the target
        self.history = History()

    def loss_of(self, acc):
        d = tile_means(acc, self.tile) - self.target
        return (self.weight * d * d).sum()

    def param_error(self, theta=None):
        th = self.theta if theta is None else theta
        out = {}
        for name, idx in self.groups.items():
            ref = self.theta_true[idx]
            out[name] = float(np.linalg.norm(th[idx] - ref) / max(np.linalg.norm(ref), 1e-12))
        return out

    def step(self):
        """One forward + adjoint on a fresh stream and one Adam step;
        returns the loss of this iteration's forward."""
        t0 = time.time()
        theta = torch.tensor(self.theta, dtype=torch.float32, requires_grad=True)
        acc = render(theta, self.ds, self._stream)
        loss = self.loss_of(acc)
        loss.backward()
        g = theta.grad.detach().cpu().numpy().astype(np.float64)
        self._stream += 1
        # Adam, step size per group
        b1, b2, eps = 0.9, 0.999, 1e-12
        self._t += 1
        self._m = b1 * self._m + (1 - b1) * g
        self._v = b2 * self._v + (1 - b2) * g * g
        mh = self._m / (1 - b1 ** self._t)
        vh = self._v / (1 - b2 ** self._t)
        upd = mh / (np.sqrt(vh) + eps)
        for name, idx in self.groups.items():
            self.theta[idx] -= self.lr[name] * upd[idx]
        self.theta = project(self.theta, self.groups)
        lv = float(loss.detach())
        self.history.loss.append(lv)
        self.history.param_error.append(self.param_error())
        self.history.seconds.append(time.time() - t0)
        return lv

    def run(self, iterations, log=None):
        for k in range(iterations):
            lv = self.step()
            if log:
                pe = self.history.param_error[-1]
                log(f"iter {k}: loss {lv:.5g} " + " ".join(f"{n} {e:.4f}" for n, e in pe.items()))
        return self.history

