"""Differentiable path tracing: torch.autograd.Function over the C ABI.

forward  = bn_set_params + bn_render_pt (PathTracingIntegrator.Render with the
           given material albedo / roughness and light emission)
backward = bn_backward_pt (path-replay adjoint of the same frame)

Tensors live on the host (the film crosses the boundary as a host buffer, as
it does for the reference's Film.Pixels); the rendering runs on the GPU.
"""
from __future__ import annotations

import numpy as np
import torch

from . import configs


class RenderPT(torch.autograd.Function):
    """film = RenderPT.apply(mat_params (n_mat, 4), light_emission (n_light, 3), device_scene, pt_params)."""

    @staticmethod
    def forward(ctx, mat_params, light_emission, ds, params):
        ds.set_params(mat_params.detach().cpu().numpy(), light_emission.detach().cpu().numpy())
        film = ds.render_pt(params)
        ctx.ds, ctx.params = ds, params
        ctx.state = (mat_params.detach().cpu().numpy().copy(), light_emission.detach().cpu().numpy().copy())
        ctx.gen = ds.param_gen
        return torch.from_numpy(film)

    @staticmethod
    def backward(ctx, grad_film):
        ds = ctx.ds
        if ds.param_gen != ctx.gen:                  # parameters changed since this forward
            ds.set_params(*ctx.state)
        gm, gl = ds.backward_pt(ctx.params, grad_film.detach().cpu().numpy().astype(np.float32))
        return torch.from_numpy(gm), torch.from_numpy(gl), None, None


def masked_l2(film, target):
    """Mean squared error over pixels finite in both images (paths whose light
    sample lands on their own emissive patch can be NaN, as in the reference)."""
    ok = torch.isfinite(film).all(dim=1) & torch.isfinite(target).all(dim=1)
    d = torch.where(ok[:, None], film - target, torch.zeros_like(film))
    return (d * d).sum() / max(int(ok.sum()) * 3, 1)


def optimise(ds, info, sc, params, steps=200, lr=1e-2, target=None, callback=None):
    """Inverse rendering of the config-3 scene: Adam on the patch materials
    and lights (albedo, roughness, emission) from the start point of
    configs.config3_params to match `target` (rendered from the truth when
    None).  Returns (final parameters, loss history)."""
    (tm, tl), (m0, l0), (mm, lm) = configs.config3_params(sc, info)
    if target is None:
        ds.set_params(tm, tl)
        target = torch.from_numpy(ds.render_pt(params))
    mat = torch.tensor(m0, requires_grad=True)
    light = torch.tensor(l0, requires_grad=True)
    mmask, lmask = torch.from_numpy(mm), torch.from_numpy(lm)
    opt = torch.optim.Adam([mat, light], lr=lr)
    hist = []
    for it in range(steps):
        opt.zero_grad()
        film = RenderPT.apply(mat, light, ds, params)
        loss = masked_l2(film, target)
        loss.backward()
        mat.grad *= mmask
        light.grad *= lmask
        opt.step()
        with torch.no_grad():                        # stay in the valid parameter range
            mat[:, :3].clamp_(0.0, 1.0)
            mat[:, 3].clamp_(0.01, 1.0)
            light.clamp_(min=0.0)
        hist.append(float(loss.detach()))
        if callback:
            callback(it, hist[-1])
    return (mat.detach().numpy(), light.detach().numpy()), hist
