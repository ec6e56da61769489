"""Differentiable rendering: the renderer as a torch.autograd.Function.

forward(mat_params, emission) uploads the parameters and renders every view
of the context's scene; backward(dL/dimage) runs the path-replay adjoint
(gi_adjoint).  mat_params is (n_mats, 5): base colour rgb (diffuse colour
for Lambert materials), roughness, metallic; emission is (n_lights, 3) --
the lights' selection weights (gi_light.power) stay as uploaded.
"""
from __future__ import annotations

import numpy as np
import torch

from .context import Context


def apply_params(scene, mat_params: np.ndarray, emission: np.ndarray):
    mats = scene.mats.copy()
    lam = mats["principled"] == 0
    mats["color"] = mat_params[:, 0:3]
    mats["diffuse_color"][lam] = mat_params[lam, 0:3]
    mats["roughness"] = mat_params[:, 3]
    mats["metallic"] = mat_params[:, 4]
    lights = scene.lights.copy()
    lights["col"] = emission
    return mats, lights


def material_params(scene) -> np.ndarray:
    m = scene.mats
    col = np.where((m["principled"] == 0)[:, None], m["diffuse_color"], m["color"])
    return np.concatenate([col, m["roughness"][:, None], m["metallic"][:, None]], 1).astype(np.float32)


class RenderFunction(torch.autograd.Function):
    """forward renders with the given parameters; backward replays exactly
    that render.  The parameters, views and options of the forward pass are
    kept, and when the context has changed since (another forward, a
    set_materials / set_lights or render in between), backward uploads them
    again and re-renders before gi_adjoint, so the gradient always belongs
    to the image forward returned."""

    @staticmethod
    def forward(fctx, mat_params, emission, ctx: Context):
        mats, lights = apply_params(ctx.scene, mat_params.detach().cpu().numpy().astype(np.float32),
                                    emission.detach().cpu().numpy().astype(np.float32))
        ctx.set_materials(mats)
        ctx.set_lights(lights)
        ctx.render()
        img = np.concatenate([im.reshape(-1, 3) for im in ctx.images()])
        fctx.gi = ctx
        fctx.state = (mats, lights, ctx._views.copy(), ctx._opts.copy())
        fctx.generation = ctx.generation
        return torch.from_numpy(img).to(mat_params.device)

    @staticmethod
    def backward(fctx, grad_img):
        ctx = fctx.gi
        if ctx.generation != fctx.generation:
            mats, lights, views, opts = fctx.state
            ctx.set_materials(mats)
            ctx.set_lights(lights)
            ctx.render(views, opts)
            fctx.generation = ctx.generation
        dm, de = ctx.adjoint(grad_img.detach().cpu().numpy().astype(np.float32))
        return (torch.from_numpy(dm).to(grad_img.device), torch.from_numpy(de).to(grad_img.device), None)


def render(ctx: Context, mat_params: torch.Tensor, emission: torch.Tensor) -> torch.Tensor:
    """Differentiable image (all views concatenated, (pixels, 3))."""
    return RenderFunction.apply(mat_params, emission, ctx)
