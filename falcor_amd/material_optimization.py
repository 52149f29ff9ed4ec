"""Inverse rendering of material textures and roughness from WARDiffPathTracer gradients.

The reference's material-optimization driver (scripts/inv-rendering/material_optimization/
run_material_optimization.py, sphere_materials_example.py, diff_render_module.py) renders the
scene with the current parameters, takes the L2 image loss against a target rendering, runs
render_grad for the material gradients and steps the parameters with Adam (base colour lr 1e-2,
roughness 3e-3). Here the parameters are the per-texel albedo of every texture (the texture
gradients of the adjoint) and the roughness of every material, each iteration one 16-spp primal
launch and one 16-spp path-replay adjoint launch (falcor_amd.autograd.DiffRenderer).
"""
from __future__ import annotations

import numpy as np

from . import scene as S
from .autograd import DiffRenderer, material_params
from .render_graph import Scene, WARDiffPathTracer


def _torch():
    import torch
    return torch


class TextureRecovery:
    """Recovers albedo textures and roughness from a target image.

    target / init: the scene with the true and the initial materials (same geometry and texture
    layout). Adam with the reference's learning rates (albedo 1e-2, roughness 3e-3, eps 1e-6);
    parameters are clamped to [0, 1] (albedo) and [0.05, 1] (roughness, the layout's lower bound)
    after each step.
    """

    def __init__(self, target: S.SceneData, init: S.SceneData, width: int = 512, height: int = 512,
                 spp_target: int = 64, spp_primal: int = 16, spp_grad: int = 16, max_bounces: int = 1,
                 lr_albedo: float = 1e-2, lr_roughness: float = 3e-3, hash_size: int = 256):
        torch = _torch()
        dev = torch.device("cuda")
        pt = WARDiffPathTracer({"maxBounces": max_bounces})
        pt.set_scene(Scene(target))
        self.target = pt.render(width, height, spp_target, frame=100000)[..., :3].clone()
        del pt
        self.scene = Scene(init)
        self.renderer = DiffRenderer(self.scene, width, height, spp_primal, spp_grad, props={"maxBounces": max_bounces},
                                     hash_size=hash_size, texture_grads=True, seed=0)
        self.texels_true = torch.from_numpy(np.asarray(target.texels, np.float32)).to(dev)
        self.rough_true = torch.from_numpy(np.asarray(target.materials[:, S.M_ROUGH], np.float32)).to(dev)
        self.params0 = torch.from_numpy(material_params(init.materials)).to(dev)
        self.textured = torch.from_numpy((init.materials[:, S.M_TEX] >= 0) & (init.materials[:, S.M_TYPE] == S.MAT_STANDARD)).to(dev)
        self.texels = torch.from_numpy(np.asarray(init.texels, np.float32)).to(dev).requires_grad_(True)
        self.rough = self.params0[:, 4].clone().requires_grad_(True)
        self.opt = torch.optim.Adam([{"params": [self.texels], "lr": lr_albedo},
                                     {"params": [self.rough], "lr": lr_roughness}], eps=1e-6)
        self.seen = torch.zeros_like(self.texels, dtype=torch.bool)  # texels that have received a gradient
        self.history: list[dict] = []

    def step(self) -> dict:
        torch = _torch()
        self.opt.zero_grad()
        mats = self.params0.clone()
        mats = torch.cat([mats[:, :4], torch.where(self.textured, self.rough, mats[:, 4])[:, None], mats[:, 5:]], dim=1)
        img = self.renderer(materials=mats, texels=self.texels)
        loss = ((img - self.target) ** 2).mean()
        loss.backward()
        self.seen |= self.texels.grad != 0
        self.opt.step()
        with torch.no_grad():
            self.texels.clamp_(0.0, 1.0)
            self.rough.clamp_(0.05, 1.0)
        rec = dict(loss=float(loss.detach()), **self.errors())
        self.history.append(rec)
        return rec

    def errors(self, seen=None) -> dict:
        """Relative RMS errors of all texels, of the texels that have received a gradient so far (or
        of the mask `seen`: a texture is only seen where its object faces the camera) and of the
        textured materials' roughness."""
        torch = _torch()
        seen = self.seen if seen is None else seen
        with torch.no_grad():
            t = float(torch.linalg.norm(self.texels - self.texels_true) / torch.linalg.norm(self.texels_true))
            if bool(seen.any()):
                ts = float(torch.linalg.norm((self.texels - self.texels_true)[seen]) /
                           torch.linalg.norm(self.texels_true[seen]))
            else:
                ts = float("nan")
            m = self.textured
            r = float(torch.linalg.norm(self.rough[m] - self.rough_true[m]) / torch.linalg.norm(self.rough_true[m]))
        return dict(texel_rmse=t, seen_texel_rmse=ts, roughness_rmse=r)
