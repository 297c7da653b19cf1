"""Gradient harness: a torch.autograd.Function over the HIP backend.

Forward renders `spp` progressive passes of the scene with the given material
colours (main_color of every material) and returns the mean linear image, the
Ref renderer's "clean" framebuffer.  Backward replays the same passes on the
device (include/rh_hip.h, rh_adjoint_pass): each lane regenerates its pixel's
path from the same iteration and Halton table, records the colour each vertex
adds and the material scaling it, and adds adj . S_m / main_color to the
material's gradient, S_m being the path's colour from that vertex on.  Memory
is O(pixels): one path per lane at a time, no tape between forward and
backward.  FP32 atomics in the gradient sum are the only run-to-run
difference.  Sampling decisions are held fixed (Russian roulette weights are
not differentiated; the refined Cornell workload turns roulette off).

This module is only the gradient harness: the renderer never needs torch.
"""
from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np
import torch

from . import capi as A


class DiffRenderer:
    """A HIP renderer + scene pair whose material colours are the parameters."""

    def __init__(self, lib: A.RayLib, renderer: A.Renderer, scene: A.Scene, width: int, height: int, spp: int,
                 n_materials: int, device: str = "cuda"):
        if not lib.has_adjoint:
            raise RuntimeError(f"{lib.path} has no adjoint entry points (HIP backend required)")
        self.lib, self.r, self.s = lib, renderer, scene
        self.w, self.h, self.spp = width, height, spp
        self.n_materials = n_materials
        self.device = device
        self.rays_forward = 0
        self.rays_adjoint = 0

    def set_colors(self, colors: torch.Tensor):
        rgb = colors.detach().to(torch.float32).cpu().numpy()
        self.r.set_material_colors(self.s, 0, rgb)

    def render(self, colors: torch.Tensor | None = None) -> torch.Tensor:
        """Mean linear image (h, w, 4) of passes 1..spp, on the device."""
        if colors is not None:
            self.set_colors(colors)
        self.r.clear()
        reg = A.Region(self.lib, 0, 0, self.w, self.h)
        for _ in range(self.spp):
            self.r.render(self.s, reg)
        img = torch.empty((self.h, self.w, 4), dtype=torch.float32, device=self.device)
        torch.cuda.synchronize()
        self.r.read_accum(img.data_ptr())
        return img

    def adjoint(self, grad_img: torch.Tensor) -> torch.Tensor:
        """d loss / d main_color (n_materials, 3) for d loss / d image grad_img
        ((h, w, 3) or (h, w, 4)), for the colours of the last render."""
        adj = torch.zeros((self.h, self.w, 4), dtype=torch.float32, device=self.device)
        adj[..., :grad_img.shape[-1]] = grad_img.to(torch.float32)
        adj[..., 3] = 0.0
        g = torch.zeros((self.n_materials, 4), dtype=torch.float32, device=self.device)
        torch.cuda.synchronize()
        reg = A.Region(self.lib, 0, 0, self.w, self.h)
        for _ in range(self.spp):
            self.r.adjoint(self.s, reg, adj.data_ptr(), g.data_ptr(), 1.0 / self.spp)
        self.r.synchronize()
        return g[:, :3].contiguous()


class RenderFunction(torch.autograd.Function):
    """image = render(colors); backward by path replay."""

    @staticmethod
    def forward(ctx, colors: torch.Tensor, dr: DiffRenderer):
        ctx.dr = dr
        return dr.render(colors)[..., :3].contiguous()

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        return ctx.dr.adjoint(grad_out.contiguous()), None


def render(colors: torch.Tensor, dr: DiffRenderer) -> torch.Tensor:
    return RenderFunction.apply(colors, dr)


def refined_cornell(lib: A.RayLib, k=17, width=512, height=512, spp=32, depth=6, albedo_seed=None, renderer=None):
    """DiffRenderer over the refined Cornell box: (dr, initial colours (n, 3) numpy)."""
    from . import scenes
    r = renderer or A.Renderer(lib, width, height, A.RENDERER_HIP)
    s = r.create_scene()
    kw = {} if albedo_seed is None else {"albedo_seed": albedo_seed}
    spec = scenes.cornell_refined(s, k=k, width=width, height=height, spp=spp, depth=depth, **kw)
    n = spec.info["materials"]
    view = s.view()
    colors = np.array(view["materials"]["main_color"], np.float32)
    return DiffRenderer(lib, r, s, width, height, spp, n), colors, spec


@dataclass
class DescentResult:
    steps: int
    seconds: float
    s_per_iter: float
    loss_first: float
    loss_last: float
    albedo_rmse_first: float
    albedo_rmse_last: float


def descend(dr: DiffRenderer, init: np.ndarray, target_img: torch.Tensor, target_albedo: np.ndarray, steps=200,
            lr=0.02, optimize=slice(None)) -> DescentResult:
    """Adam on the colours selected by `optimize` (the others fixed) against an
    L2 image loss; returns timing and convergence."""
    dev = dr.device
    full = torch.tensor(init, dtype=torch.float32, device=dev)
    mask = torch.zeros(full.shape[0], dtype=torch.bool, device=dev)
    mask[optimize] = True
    params = full[mask].clone().requires_grad_(True)
    opt = torch.optim.Adam([params], lr=lr)
    tgt_alb = torch.tensor(target_albedo, dtype=torch.float32, device=dev)[mask]

    def rmse():
        return float(torch.sqrt(torch.mean((params.detach() - tgt_alb) ** 2)))

    first_rmse = rmse()
    losses = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        opt.zero_grad(set_to_none=True)
        colors = full.clone()
        colors = colors.masked_scatter(mask[:, None].expand_as(colors), params)
        img = render(colors, dr)
        loss = torch.mean((img - target_img) ** 2)
        loss.backward()
        opt.step()
        with torch.no_grad():
            params.clamp_(0.01, 1.0)
        losses.append(loss.detach())
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return DescentResult(steps, dt, dt / max(steps, 1), float(losses[0]), float(losses[-1]), first_rmse, rmse())
