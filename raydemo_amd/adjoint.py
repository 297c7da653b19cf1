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

Each autograd forward records the scene state it rendered (material colours,
changed textures); a backward taken after the state changed (several renders
before one backward, or new colours set in between) first puts that state back,
so every image is differentiated at the parameters that made it.

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
                 n_materials: int, device: str = "cuda", tape: bool | None = False, tape_texels: bool = False):
        """tape: False (the default) differentiates by path replay, with
        memory O(pixels); True keeps every forward pass's vertex records on the
        device (the "adjoint_tape" option, O(pixels x passes x depth)) so that
        adjoint() needs no replay; None turns the tape on when the records of
        all passes take at most half of the free device memory.  tape_texels:
        the records include texel footprints (for texel gradients; 4x the
        memory)."""
        if not lib.has_adjoint:
            raise RuntimeError(f"{lib.path} has no adjoint entry points (HIP backend required)")
        self.lib, self.r, self.s = lib, renderer, scene
        self.w, self.h, self.spp = width, height, spp
        self.n_materials = n_materials
        self.device = device
        self.rays_forward = 0
        self.rays_adjoint = 0
        self._layout = None
        self._grad_textures = None
        # the parameter state now on the device: colours, textures changed
        # through set_texture (with their original texels), and a version
        self._colors = np.array(scene.view()["materials"]["main_color"], np.float32)
        self._textures: dict = {}
        self._tex_original: dict = {}
        self._version = 0
        if tape is None:
            vmax = renderer.get_option("max_bounces") + 1
            need = spp * width * height * (vmax * (64 if tape_texels else 16) + 4)
            tape = need <= 0.5 * torch.cuda.mem_get_info()[0]
        self.tape = bool(tape)
        self.r.set_option("adjoint_tape", spp if self.tape else 0)
        self.r.set_option("adjoint_tape_texels", 1 if (self.tape and tape_texels) else 0)

    def tex_layout(self) -> dict:
        """The scene's texture records, mip blocks and texel count (they do not
        change when texel values do)."""
        if self._layout is None:
            v = self.s.view()
            self._layout = {k: v[k] for k in ("textures", "tex_mips", "texels_count")}
        return self._layout

    def tape_bytes(self) -> int:
        """Device memory the tape of all passes takes (0 without the tape)."""
        if not self.tape:
            return 0
        vmax = self.r.get_option("max_bounces") + 1
        tex = self.r.get_option("adjoint_tape_texels")
        return self.spp * self.w * self.h * (vmax * (64 if tex else 16) + 4)

    def replay_bytes(self) -> int:
        """Device memory of the path replay's per-lane vertex records (held
        for one pass at a time): O(pixels)."""
        vmax = 17  # records per lane: at most max_bounces + 1 <= 17 vertices
        return self.w * self.h * (vmax * (16 + 48) + 4)

    def set_colors(self, colors: torch.Tensor | np.ndarray):
        if isinstance(colors, torch.Tensor):
            colors = colors.detach().to(torch.float32).cpu().numpy()
        rgb = np.array(colors, np.float32).reshape(-1, 3)
        self.r.set_material_colors(self.s, 0, rgb)
        if rgb.shape != self._colors.shape or not np.array_equal(rgb, self._colors):
            self._colors = rgb
            self._version += 1

    def state(self):
        """The parameter state now on the device, for restore()."""
        return (self._version, self._colors, dict(self._textures))

    def restore(self, state):
        """Puts back a state() taken earlier: the colours, and every texture
        changed since (to its texels then, or its original texels)."""
        version, colors, textures = state
        if version == self._version:
            return
        if colors.shape != self._colors.shape or not np.array_equal(colors, self._colors):
            self.set_colors(colors)
        for tex in set(textures) | set(self._textures):
            want = textures.get(tex, self._tex_original.get(tex))
            if want is not None and not np.array_equal(want, self._textures.get(tex)):
                self._set_texture_rgba(tex, want)
        self._version = version

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

    def adjoint(self, grad_img: torch.Tensor, texels: bool = False, textures=None):
        """d loss / d main_color (n_materials, 3) for d loss / d image grad_img
        ((h, w, 3) or (h, w, 4)), at the colours and texels now on the device
        (the autograd functions put back the state of their own render
        first); with
        texels, also d loss / d stored texel (n_texels, 4) on the device, for
        the texels of `textures` only when given (the others stay zero)."""
        want = tuple(textures) if (texels and textures) else None
        if want != self._grad_textures:
            self.r.set_gradient_textures(want)
            self._grad_textures = want
        adj = torch.zeros((self.h, self.w, 4), dtype=torch.float32, device=self.device)
        adj[..., :grad_img.shape[-1]] = grad_img.to(torch.float32)
        adj[..., 3] = 0.0
        g = torch.zeros((self.n_materials, 4), dtype=torch.float32, device=self.device)
        tg = None
        if texels:
            tg = torch.zeros((self.tex_layout()["texels_count"], 4), dtype=torch.float32, device=self.device)
        torch.cuda.synchronize()
        reg = A.Region(self.lib, 0, 0, self.w, self.h)
        for _ in range(self.spp):
            self.r.adjoint(self.s, reg, adj.data_ptr(), g.data_ptr(), 1.0 / self.spp,
                           tg.data_ptr() if tg is not None else None)
        self.r.synchronize()
        if texels:
            return g[:, :3].contiguous(), tg
        return g[:, :3].contiguous()

    def set_texture(self, tex: int, rgb: torch.Tensor):
        """Level-0 texels of texture tex from values in [0, 1] ((h, w, 3),
        alpha 255), quantised to 8 bits as the texture stores them."""
        q = torch.clamp(torch.round(rgb.detach().to(torch.float32) * 255.0), 0, 255).to(torch.uint8).cpu().numpy()
        rgba = np.concatenate([q, np.full(q.shape[:2] + (1,), 255, np.uint8)], axis=2)
        if not np.array_equal(rgba, self._textures.get(tex)):
            self._set_texture_rgba(tex, rgba)
            self._version += 1

    def _set_texture_rgba(self, tex: int, rgba: np.ndarray):
        if tex not in self._tex_original:
            self._tex_original[tex] = self._level0(tex)
        self.r.set_texture_data(self.s, tex, rgba)
        self._textures[tex] = rgba

    def _level0(self, tex: int) -> np.ndarray:
        """Texture tex's level-0 texels as stored (h, w, 4) uint8: the inside of
        its bordered level-0 block."""
        v = self.s.raw_view()
        m = self.tex_layout()["tex_mips"][tex][0]
        off, bw, bh = int(m["offset"]), int(m["bw"]), int(m["bh"])
        words = np.ctypeslib.as_array((A.c_uint32 * (off + bw * bh)).from_address(v.texels))[off:]
        block = words.reshape(bh, bw)[1:-1, 1:-1].copy()
        return block.view(np.uint8).reshape(bh - 2, bw - 2, 4)


def fold_texel_grad(view: dict, tex: int, tg: np.ndarray) -> np.ndarray:
    """Gradient of the stored texels of texture tex (its bordered blocks, one
    per mip level) onto its level-0 texels, (h, w, 4): a block texel (x, y)
    is level texel ((x-1) mod w, (y-1) mod h), and a level-k texel is the box
    average of 4 level-(k-1) texels, so its gradient spreads as 1/4 to each
    (the 8-bit rounding of the mips is taken as the identity)."""
    t = view["textures"][tex]
    w0, h0 = int(t["width"]) & 0x7FFF, int(t["height"])
    mips = view["tex_mips"][tex]
    levels = []
    for lev in range(len(mips)):
        if lev > 0 and mips[lev]["offset"] == mips[lev - 1]["offset"]:
            break
        bw, bh, off = int(mips[lev]["bw"]), int(mips[lev]["bh"]), int(mips[lev]["offset"])
        lw, lh = bw - 2, bh - 2
        block = tg[off:off + bw * bh].reshape(bh, bw, 4)
        g = np.zeros((lh, lw, 4), np.float64)
        ys = (np.arange(bh) - 1) % lh
        xs = (np.arange(bw) - 1) % lw
        np.add.at(g, (ys[:, None], xs[None, :]), block)
        levels.append(g)
    for lev in range(len(levels) - 1, 0, -1):
        up = np.repeat(np.repeat(levels[lev], 2, axis=0), 2, axis=1) / 4.0
        levels[lev - 1] += up
    assert levels[0].shape[:2] == (h0, w0)
    return levels[0]


def fold_texel_grad_torch(view: dict, tex: int, tg: torch.Tensor) -> torch.Tensor:
    """fold_texel_grad on the device: tg is the (texels, 4) gradient tensor,
    the result the (h, w, 3) level-0 gradient (float32) on tg's device."""
    t = view["textures"][tex]
    w0, h0 = int(t["width"]) & 0x7FFF, int(t["height"])
    mips = view["tex_mips"][tex]
    dev = tg.device
    levels = []
    for lev in range(len(mips)):
        if lev > 0 and mips[lev]["offset"] == mips[lev - 1]["offset"]:
            break
        bw, bh, off = int(mips[lev]["bw"]), int(mips[lev]["bh"]), int(mips[lev]["offset"])
        lw, lh = bw - 2, bh - 2
        block = tg[off:off + bw * bh, :3].reshape(bh * bw, 3)
        ys = (torch.arange(bh, device=dev) - 1) % lh
        xs = (torch.arange(bw, device=dev) - 1) % lw
        dst = (ys[:, None] * lw + xs[None, :]).reshape(-1)
        g = torch.zeros((lh * lw, 3), dtype=torch.float32, device=dev)
        g.index_add_(0, dst, block.to(torch.float32))
        levels.append(g.reshape(lh, lw, 3))
    for lev in range(len(levels) - 1, 0, -1):
        up = levels[lev].repeat_interleave(2, 0).repeat_interleave(2, 1) / 4.0
        levels[lev - 1] += up
    assert levels[0].shape[:2] == (h0, w0)
    return levels[0]


class RenderFunction(torch.autograd.Function):
    """image = render(colors); backward by path replay."""

    @staticmethod
    def forward(ctx, colors: torch.Tensor, dr: DiffRenderer):
        ctx.dr = dr
        img = dr.render(colors)[..., :3].contiguous()
        ctx.state = dr.state()
        return img

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        ctx.dr.restore(ctx.state)
        return ctx.dr.adjoint(grad_out.contiguous()), None


def render(colors: torch.Tensor, dr: DiffRenderer) -> torch.Tensor:
    return RenderFunction.apply(colors, dr)


class TextureRenderFunction(torch.autograd.Function):
    """image = render(level-0 texels of the given textures); backward by path
    replay, texel gradients folded onto level 0 (straight through the 8-bit
    quantisation of the texels)."""

    @staticmethod
    def forward(ctx, texels: torch.Tensor, dr: DiffRenderer, tex_ids: tuple):
        ctx.dr, ctx.tex_ids, ctx.param_device = dr, tex_ids, texels.device
        for k, t in enumerate(tex_ids):
            dr.set_texture(t, texels[k])
        img = dr.render()[..., :3].contiguous()
        ctx.state = dr.state()
        return img

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        dr = ctx.dr
        dr.restore(ctx.state)
        _, tg = dr.adjoint(grad_out.contiguous(), texels=True, textures=ctx.tex_ids)
        layout = dr.tex_layout()
        g = torch.stack([fold_texel_grad_torch(layout, t, tg) for t in ctx.tex_ids])
        return g.to(ctx.param_device), None, None


def render_textures(texels: torch.Tensor, dr: DiffRenderer, tex_ids: tuple) -> torch.Tensor:
    return TextureRenderFunction.apply(texels, dr, tex_ids)


def refined_cornell(lib: A.RayLib, k=17, width=512, height=512, spp=32, depth=6, albedo_seed=None, renderer=None,
                    tape=False):
    """DiffRenderer over the refined Cornell box: (dr, initial colours (n, 3) numpy)."""
    from . import scenes
    r = renderer or A.Renderer(lib, width, height, A.RENDERER_HIP)
    s = r.create_scene()
    kw = {} if albedo_seed is None else {"albedo_seed": albedo_seed}
    spec = scenes.cornell_refined(s, k=k, width=width, height=height, spp=spp, depth=depth, **kw)
    n = spec.info["materials"]
    view = s.view()
    colors = np.array(view["materials"]["main_color"], np.float32)
    return DiffRenderer(lib, r, s, width, height, spp, n, tape=tape), colors, spec


def textured_cornell(lib: A.RayLib, width=128, height=128, spp=8, depth=4, tex_size=64, mipmaps=True, textures=None,
                     renderer=None, tape=False):
    """DiffRenderer over the textured Cornell box: (dr, spec); spec.info["textures"]
    are the four differentiable textures."""
    from . import scenes
    r = renderer or A.Renderer(lib, width, height, A.RENDERER_HIP)
    s = r.create_scene()
    spec = scenes.textured_cornell(s, tex_size=tex_size, width=width, height=height, spp=spp, depth=depth,
                                   mipmaps=mipmaps, textures=textures)
    return DiffRenderer(lib, r, s, width, height, spp, len(s.view()["materials"]), tape=tape, tape_texels=True), spec


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
