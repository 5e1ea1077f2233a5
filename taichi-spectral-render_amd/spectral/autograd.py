"""Differentiable rendering: a torch.autograd.Function over the HIP forward
kernel and its path-replay adjoint (sr_render_adjoint).

The reference has no autodiff (no grad/Tape anywhere in main.py); the
parameters follow how its shading uses the material tables:

* ``albedo[m]`` (per band) of Oren-Nayar materials: the diffuse albedo
  s / (s + 4 pi k / lambda + 1e-6) of main.py:1070, 1395; set through k.
* ``roughness`` of one material (GGX D, G and the half-vector sampler,
  main.py:900-978).
* ``g`` of one medium (Henyey-Greenstein, main.py:758-792).
* ``sigma_a`` (per band) of one medium: 4 pi k / lambda (main.py:1203).

The flat parameter vector is [albedo rows..., roughness, g, sigma_a].
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .render import Renderer
from .scene import FlatScene

PROP_K, PROP_R, PROP_SS, PROP_G = 1, 2, 3, 5  # main.py:40-45


@dataclass
class ParamSpec:
    albedo_mats: list = field(default_factory=list)
    rough_mat: int = -1
    g_mat: int = -1
    sigma_mat: int = -1

    def layout(self, bands):
        off = {}
        i = 0
        for m in self.albedo_mats:
            off[("albedo", m)] = i
            i += bands
        if self.rough_mat >= 0:
            off["roughness"] = i
            i += 1
        if self.g_mat >= 0:
            off["g"] = i
            i += 1
        if self.sigma_mat >= 0:
            off["sigma_a"] = i
            i += bands
        return off, i


class DifferentiableScene:
    """A Renderer whose selected material properties are driven by a flat
    parameter tensor."""

    def __init__(self, flat: FlatScene, spec: ParamSpec, width, height, spp, depth, seed=0, device="cuda"):
        self.flat = flat
        self.spec = spec
        self.B = flat.bands
        self.r = Renderer(flat, flat.bands, flat.wl_min, flat.wl_max, width=width, height=height, device=device)
        self.spp, self.depth, self.seed = int(spp), int(depth), int(seed)
        self.offsets, self.n_params = spec.layout(self.B)
        self.lam_m = np.linspace(flat.wl_min, flat.wl_max, self.B) * 1e-9
        self.base_f32 = flat.mat_f32.copy()
        self.alb_off = np.full((flat.n_mats,), -1, np.int32)
        for m in spec.albedo_mats:
            self.alb_off[m] = self.offsets[("albedo", m)]
        self.current_params = None  # the parameters last uploaded by set_params
        # RenderFunction records the forward's path totals (4 bytes per path)
        # up to this size; larger renders replay every path twice
        self.record_limit_bytes = 4 << 30

    def initial_params(self):
        """Parameter vector read back from the current material tables."""
        f = self.base_f32.astype(np.float64)
        th = np.zeros(self.n_params)
        for m in self.spec.albedo_mats:
            s = f[m, PROP_SS]
            absorb = 4 * math.pi * f[m, PROP_K] / self.lam_m
            th[self.offsets[("albedo", m)]:][:self.B] = s / (s + absorb + 1e-6)
        if self.spec.rough_mat >= 0:
            th[self.offsets["roughness"]] = f[self.spec.rough_mat, PROP_R, 0]
        if self.spec.g_mat >= 0:
            th[self.offsets["g"]] = f[self.spec.g_mat, PROP_G, 0]
        if self.spec.sigma_mat >= 0:
            th[self.offsets["sigma_a"]:][:self.B] = 4 * math.pi * f[self.spec.sigma_mat, PROP_K] / self.lam_m
        return th

    def set_params(self, theta):
        th = np.asarray(theta, np.float64)
        f = self.base_f32.astype(np.float64)
        for m in self.spec.albedo_mats:
            a = th[self.offsets[("albedo", m)]:][:self.B]
            s = f[m, PROP_SS]
            f[m, PROP_K] = (s * (1.0 - a) / a - 1e-6) * self.lam_m / (4 * math.pi)
        if self.spec.rough_mat >= 0:
            f[self.spec.rough_mat, PROP_R, :] = th[self.offsets["roughness"]]
        if self.spec.g_mat >= 0:
            f[self.spec.g_mat, PROP_G, :] = th[self.offsets["g"]]
        if self.spec.sigma_mat >= 0:
            sa = th[self.offsets["sigma_a"]:][:self.B]
            f[self.spec.sigma_mat, PROP_K] = sa * self.lam_m / (4 * math.pi)
        self.r.update_materials(f.astype(np.float32))
        self.current_params = th.copy()

    def new_path_totals(self):
        """A buffer for the path totals of one render (one float per path),
        or None when it would exceed ``record_limit_bytes``."""
        n = self.spp * self.r.width * self.r.height * self.B
        if n * 4 > self.record_limit_bytes:
            return None
        return torch.empty((n,), dtype=torch.float32, device=self.r.accum.device)

    def render(self, stream=0, path_totals=None):
        """Fresh render of `spp` blocks; returns the spectral mean [W, H, B].
        Paths end once their throughput is exactly 0 (the same image; the
        adjoint replays them the same way).  ``path_totals`` (from
        new_path_totals) receives each path's total for ``adjoint``."""
        self.r.clear_pixel_space()
        self.r.render(n_blocks=self.spp, max_depth=self.depth, seed=self.seed, stream=stream, end_dead_paths=True,
                      path_totals=path_totals)
        return self.r.accum

    def adjoint(self, d_accum, stream=0, flags=0, path_totals=None):
        """Gradient of sum(d_accum * render()) with respect to the parameters.
        With the ``path_totals`` of that render, each path whose
        contributions all stayed in its band is replayed once instead of
        twice (the same gradient up to float rounding)."""
        d = d_accum.contiguous().to(torch.float32)
        want = (self.r.width, self.r.height, self.B)
        if tuple(d.shape) != want or d.device.type != "cuda":
            raise ValueError(f"d_accum must be a cuda tensor of shape {want}, got {tuple(d.shape)} on {d.device}")
        grad = torch.zeros((self.n_params,), dtype=torch.float32, device=d.device)
        sp = self.spec
        self.r._bind_stream()
        args = [self.r._h, self.r.width, self.r.height, ctypes.c_uint64(self.seed), ctypes.c_uint32(stream), 0,
                self.spp, self.depth, _lib.ptr(d), _lib.ptr(self.alb_off), sp.rough_mat,
                self.offsets.get("roughness", -1), sp.g_mat, self.offsets.get("g", -1), sp.sigma_mat,
                self.offsets.get("sigma_a", -1), self.n_params, int(flags)]
        if path_totals is None:
            _lib.call("sr_render_adjoint", *args, _lib.ptr(grad))
        else:
            n = self.spp * self.r.width * self.r.height * self.B
            if path_totals.numel() != n or path_totals.dtype != torch.float32 or path_totals.device != d.device:
                raise ValueError(f"path_totals must be a float32 tensor of {n} elements on {d.device}")
            _lib.call("sr_render_adjoint_record", *args, _lib.ptr(path_totals), _lib.ptr(grad))
        return grad


class RenderFunction(torch.autograd.Function):
    """accum = render(theta) on Philox stream `stream`; backward runs the
    path-replay adjoint of that same render (same parameters, seed and
    stream)."""

    @staticmethod
    def forward(ctx, theta, scene: DifferentiableScene, stream=0):
        th = theta.detach().cpu().numpy().copy()
        scene.set_params(th)
        totals = scene.new_path_totals()
        out = scene.render(stream=stream, path_totals=totals).clone()
        ctx.scene = scene
        ctx.path_totals = totals  # this forward's paths: the adjoint replays them once
        ctx.theta = th  # the adjoint replays the paths of THESE parameters
        ctx.stream = int(stream)
        ctx.seed = scene.seed
        ctx.theta_device = theta.device
        ctx.theta_dtype = theta.dtype
        return out

    @staticmethod
    def backward(ctx, grad_out):
        scene = ctx.scene
        if scene.current_params is None or not np.array_equal(scene.current_params, ctx.theta):
            scene.set_params(ctx.theta)  # another forward ran on this scene since
        seed = scene.seed
        scene.seed = ctx.seed
        try:
            g = scene.adjoint(grad_out, stream=ctx.stream, path_totals=ctx.path_totals)
        finally:
            scene.seed = seed
        return g.to(device=ctx.theta_device, dtype=ctx.theta_dtype), None, None


def render(theta, scene: DifferentiableScene, stream=0):
    return RenderFunction.apply(theta, scene, int(stream))


def inverse_volume_spec():
    """The inverse-render parameters of the volume/fluorescence scene:
    3 wall albedo spectra, slab roughness, medium g and sigma_a."""
    return ParamSpec(albedo_mats=[1, 2, 3], rough_mat=6, g_mat=5, sigma_mat=5)
