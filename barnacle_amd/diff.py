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
    """Mean squared error over pixels finite in both images (a path can end
    in a NaN radiance under the reference's semantics, e.g. a self-hit on a
    surface that both reflects and emits)."""
    ok = torch.isfinite(film).all(dim=1) & torch.isfinite(target).all(dim=1)
    d = torch.where(ok[:, None], film - target, torch.zeros_like(film))
    return (d * d).sum() / max(int(ok.sum()) * 3, 1)


def parameter_errors(info, sc, mat, light, seen=None):
    """Mean absolute error of the optimised parameter sets against the truth
    of configs.config3: patch albedo, patch roughness, tile emission.  With
    `seen` (a boolean per patch: the patches the image constrains, e.g. those
    with a non-zero gradient at the start), also the errors over those
    patches, with roughness split into metallic patches (whose whole
    reflection is the GGX lobe) and dielectric ones (4% specular)."""
    (tm, tl), _, _ = configs.config3_params(sc, info)
    p = info["x-patches"]
    fm, fl, n = p["first_material"], p["first_light"], p["count"]
    m = np.asarray(mat)[fm:fm + n]
    t = tm[fm:fm + n]
    light_ = np.asarray(light)[fl:fl + n]
    out = {"albedo": float(np.abs(m[:, :3] - t[:, :3]).mean()),
           "roughness": float(np.abs(m[:, 3] - t[:, 3]).mean()),
           "emission": float(np.abs(light_ - tl[fl:fl + n]).mean())}
    if seen is not None:
        metal = np.array([sc.materials[fm + k]["metallic"] for k in range(n)]) > 0.5
        out["albedo_seen"] = float(np.abs(m[seen, :3] - t[seen, :3]).mean())
        for name, sel in (("roughness_seen_metallic", seen & metal), ("roughness_seen_dielectric", seen & ~metal)):
            out[name] = float(np.abs(m[sel, 3] - t[sel, 3]).mean()) if sel.any() else None
    return out


def seen_patches(ds, info, sc, params):
    """Patches the image constrains: a non-zero gradient of the L2 loss at
    the start point (patches facing away from the camera, or hidden, get
    none)."""
    (tm, tl), (m0, l0), _ = configs.config3_params(sc, info)
    ds.set_params(tm, tl)
    target = ds.render_pt(params)
    ds.set_params(m0, l0)
    film = ds.render_pt(params)
    d = (2.0 * (film - target)).astype(np.float32)
    d[~np.isfinite(d).all(axis=1)] = 0.0
    gm, _ = ds.backward_pt(params, d)
    p = info["x-patches"]
    fm, n = p["first_material"], p["count"]
    return np.abs(gm[fm:fm + n]).sum(axis=1) > 0


def optimise(ds, info, sc, params, steps=200, lr=1e-2, target=None, callback=None, fresh_samples=False,
             lr_decay=1.0):
    """Inverse rendering of the config-3 scene: Adam on the patch materials
    (albedo, roughness) and tile lights (emission) from the start point of
    configs.config3_params.

    Default: match the fixed `target` (rendered from the truth when None)
    with the same random streams every step.  fresh_samples=True draws new
    streams every step (frame id = step + 1) and renders that step's target
    from the truth with the same streams: the loss then compares images with
    common random numbers while its noise is independent from step to step,
    which is what path replay's gradient (of the expected image, sampling
    held fixed) needs to average out.  lr_decay multiplies the learning rate
    after every step.  Returns (final parameters, loss history).
    callback(step, loss, mat, light), if given, sees the parameters after
    every step."""
    (tm, tl), (m0, l0), (mm, lm) = configs.config3_params(sc, info)
    if target is None and not fresh_samples:
        ds.set_params(tm, tl)
        target = torch.from_numpy(ds.render_pt(params))
    mat = torch.tensor(m0, requires_grad=True)
    light = torch.tensor(l0, requires_grad=True)
    mmask, lmask = torch.from_numpy(mm), torch.from_numpy(lm)
    opt = torch.optim.Adam([mat, light], lr=lr)
    sched = torch.optim.lr_scheduler.ExponentialLR(opt, gamma=lr_decay)
    hist = []
    for it in range(steps):
        p = params
        if fresh_samples:
            p = type(params).from_buffer_copy(params)
            p.frame_id = it + 1
            ds.set_params(tm, tl)
            target = torch.from_numpy(ds.render_pt(p))
        opt.zero_grad()
        film = RenderPT.apply(mat, light, ds, p)
        loss = masked_l2(film, target)
        loss.backward()
        mat.grad *= mmask
        light.grad *= lmask
        opt.step()
        sched.step()
        with torch.no_grad():                        # stay in the valid parameter range
            mat[:, :3].clamp_(0.0, 1.0)
            mat[:, 3].clamp_(0.01, 1.0)
            light.clamp_(min=0.0)
        hist.append(float(loss.detach()))
        if callback:
            callback(it, hist[-1], mat.detach().numpy(), light.detach().numpy())
    return (mat.detach().numpy(), light.detach().numpy()), hist
