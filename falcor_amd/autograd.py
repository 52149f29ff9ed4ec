"""torch.autograd binding of the WARDiffPathTracer forward and adjoint passes.

`DiffRenderer(scene, ...)(materials=..., texels=..., positions=..., normals=...)` renders the
image (H, W, 3) with the given parameters and, on `backward`, runs the path-replay adjoint
(`fhp_pass_backward`) with dL/dI = the incoming gradient, returning the gradients of the
parameters that were passed in. It plays the role of the reference's DiffRenderModule drivers
(scripts/inv-rendering/material_optimization/diff_render_module.py,
scripts/inv-rendering/shape_optimization/diff_render_module.py) and render_primal / render_grad
(scripts/inv-rendering/common.py:80-114), with the spp samples of each pass in one launch.

Material parameters use the reference's 20-float layout per material
(StandardMaterialParamLayout.slang:32-41): baseColor 0-2, metallic 3, roughness 4, ior 5,
transmissionColor 6-8, diffuseTransmission 9, specularTransmission 10, emissive 11-13,
emissiveFactor 14 (the transmission entries 6-10 through the transmission lobes of materials with
transmission, and 9-10 through the lobe weights of every StandardMaterial).
"""
from __future__ import annotations

import numpy as np

from . import scene as S
from .render_graph import Scene, SceneGradients, WARDiffPathTracer

PARAM_COUNT = 20


def material_params(records: np.ndarray) -> np.ndarray:
    """Scene material records (20 floats each) -> reference parameter layout (20 floats each)."""
    r = np.asarray(records, np.float32)
    p = np.zeros((r.shape[0], PARAM_COUNT), np.float32)
    p[:, 0:3] = r[:, S.M_BASE:S.M_BASE + 3]
    p[:, 3] = r[:, S.M_METAL]
    p[:, 4] = r[:, S.M_ROUGH]
    p[:, 5] = r[:, S.M_IOR]
    p[:, 6:9] = r[:, S.M_TRANS:S.M_TRANS + 3]
    p[:, 9] = r[:, S.M_DIFF_TRANS]
    p[:, 10] = r[:, S.M_SPEC_TRANS]
    p[:, 11:14] = r[:, S.M_EMIT:S.M_EMIT + 3]
    p[:, 14] = r[:, S.M_EMIT_FACTOR]
    return p


def material_records(params: np.ndarray, like: np.ndarray) -> np.ndarray:
    """Reference parameter layout -> scene material records (other fields copied from `like`)."""
    p = np.asarray(params, np.float32)
    r = np.array(like, np.float32, copy=True)
    r[:, S.M_BASE:S.M_BASE + 3] = p[:, 0:3]
    r[:, S.M_METAL] = p[:, 3]
    r[:, S.M_ROUGH] = p[:, 4]
    r[:, S.M_IOR] = p[:, 5]
    r[:, S.M_TRANS:S.M_TRANS + 3] = p[:, 6:9]
    r[:, S.M_DIFF_TRANS] = p[:, 9]
    r[:, S.M_SPEC_TRANS] = p[:, 10]
    r[:, S.M_EMIT:S.M_EMIT + 3] = p[:, 11:14]
    r[:, S.M_EMIT_FACTOR] = p[:, 14]
    return r


def _fn():
    import torch

    class _Render(torch.autograd.Function):
        @staticmethod
        def forward(ctx, renderer, materials, texels, positions, normals):
            r = renderer
            if materials is not None:
                r.scene.set_material_records(material_records(materials.detach().cpu().numpy(), r.scene.materials))
            if texels is not None:
                r.scene.set_texels(texels.detach())
            if positions is not None:
                r.scene.set_mesh_vertices(r.mesh_id, positions.detach(), normals.detach() if normals is not None else None)
            img = r.pass_.render(r.width, r.height, r.spp_primal, frame=r._next_frame())[..., :3].clone()
            ctx.renderer = r
            ctx.wants = (materials is not None, texels is not None, positions is not None, normals is not None)
            return img

        @staticmethod
        def backward(ctx, grad):
            r = ctx.renderer
            g = r.grads
            r.pass_.scene_gradients = g
            r.pass_.dL_dI = grad.contiguous()
            g.clear_all_grads()
            r.pass_.backward(r.width, r.height, r.spp_grad, frame=r._next_frame())
            g.aggregate_all_grads()
            inv = 1.0 / r.spp_grad  # render_grad divides the raw sum by spp (common.py:111-112)
            wm, wt, wp, wn = ctx.wants
            gm = g.get_grads_buffer(g.MATERIAL).view(-1, PARAM_COUNT) * inv if wm else None
            gt = g.get_grads_buffer(g.TEXTURE) * inv if wt else None
            gp = g.get_grads_buffer(g.MESH_POSITION).view(-1, 3) * inv if wp else None
            gn = g.get_grads_buffer(g.MESH_NORMAL).view(-1, 3) * inv if wn else None
            return None, gm, gt, gp, gn

    return _Render


class DiffRenderer:
    """Differentiable WARDiffPathTracer renderer of one scene.

    props are WARDiffPathTracer properties; mesh_id selects the mesh whose vertices receive
    gradients (useWAR = 1, maxBounces up to 8 for vertex gradients).
    """

    def __init__(self, scene: Scene, width: int, height: int, spp_primal: int = 16, spp_grad: int = 16,
                 props: dict | None = None, mesh_id: int | None = None, hash_size: int = 256,
                 texture_grads: bool = True, seed: int = 0):
        self.scene = scene
        self.width, self.height = width, height
        self.spp_primal, self.spp_grad = spp_primal, spp_grad
        self.pass_ = WARDiffPathTracer(dict({"diffMode": "BackwardDiff"}, **(props or {})))
        self.pass_.set_scene(scene)
        self.mesh_id = mesh_id
        self.grads = SceneGradients(scene, hash_size=hash_size, texture_grads=texture_grads, mesh_id=mesh_id)
        if mesh_id is not None:
            self.pass_.set_mesh_to_optimize(mesh_id)
        self._frame = seed
        self._op = _fn()

    def _next_frame(self) -> int:
        f = self._frame
        self._frame += 1
        return f

    def __call__(self, materials=None, texels=None, positions=None, normals=None):
        if (positions is not None or normals is not None) and self.mesh_id is None:
            raise ValueError("vertex parameters need mesh_id")
        return self._op.apply(self, materials, texels, positions, normals)
