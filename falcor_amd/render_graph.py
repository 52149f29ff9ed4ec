"""Falcor-style render-pass / render-graph surface over libfhp.

Mirrors the parts of Falcor's Python API that MinimalPathTracer and WARDiffPathTracer users
touch: Testbed (Source/Falcor/Core/Testbed.cpp:703-768), RenderGraph.create_pass / add_edge /
mark_output / get_output (Source/Falcor/RenderGraph/RenderGraph.cpp:725-810), the pass property
dictionaries (MinimalPathTracer.cpp:61-89, WARDiffPathTracer.cpp:53-144) and the WAR pass's
scene_gradients / run_backward / dL_dI properties (WARDiffPathTracer.cpp:697-715).
All images are torch tensors on the current HIP device; the work runs in libfhp's kernels.
"""
from __future__ import annotations

import ctypes as C
import math
import warnings
from pathlib import Path

import numpy as np

from . import _lib
from .scene import MAT_STANDARD as S_MAT_STANDARD, MAT_STRIDE, SceneData


def _torch():
    import torch
    return torch


def _device():
    torch = _torch()
    if not torch.cuda.is_available():
        raise _lib.FhpError("no HIP device available: fhp passes run on the GPU only")
    return torch.device("cuda", torch.cuda.current_device())


# --------------------------------------------------------------------------- scene
class Material:
    """Material parameter layout (Material.cpp:409-431: PARAM_COUNT, get_material_param_layout)."""

    PARAM_COUNT = 20
    # StandardMaterialParamLayout.slang:32-41 (python names, offsets and sizes).
    LAYOUT = {"base_color": (0, 3), "metallic": (3, 1), "roughness": (4, 1), "ior": (5, 1),
              "transmission_color": (6, 3), "diffuse_transmission": (9, 1), "specular_transmission": (10, 1),
              "emissive_color": (11, 3), "emissive_factor": (14, 1)}


def get_material_param_layout(material_type: str = "Standard") -> dict:
    """{python name: {"offset", "size"}} of a material type's serialized parameters."""
    if material_type not in ("Standard", 0, S_MAT_STANDARD):
        raise ValueError(f"material type {material_type!r} has no parameter layout here")
    return {k: {"offset": o, "size": n} for k, (o, n) in Material.LAYOUT.items()}


class Scene:
    """Device-resident scene (BVH4, light collection, light BVH, materials, textures)."""

    def __init__(self, data: SceneData, use_compressed_hit_info: bool = False):
        """use_compressed_hit_info: 64-bit PackedHitInfo when the header fits
        (SceneBuilder::Flags::UseCompressedHitInfo)."""
        L = _lib.load()
        self.data = data
        # Own copies of the vertex arrays: vertex updates and node transforms write them, and the
        # caller's SceneData stays as it was given.
        self._arrays = dict(
            pos=np.array(data.positions, np.float32, order="C"), nrm=np.array(data.normals, np.float32, order="C"),
            uv=np.ascontiguousarray(data.texcoords, np.float32), idx=np.ascontiguousarray(data.indices, np.uint32),
            tm=np.ascontiguousarray(data.tri_material, np.uint32), mats=np.ascontiguousarray(data.materials, np.float32),
            td=np.ascontiguousarray(data.tex_desc, np.int32), tx=np.ascontiguousarray(data.texels, np.float32),
            cam=np.ascontiguousarray(data.camera_array(), np.float32))
        off = data.mesh_vertex_offsets
        if off is None:
            off = np.array([0, data.positions.shape[0]], np.uint32)
        self._arrays["moff"] = np.ascontiguousarray(off, np.uint32)
        a = self._arrays
        d = _lib.SceneDesc()
        d.num_vertices = a["pos"].shape[0]
        d.h_positions, d.h_normals, d.h_texcoords = a["pos"].ctypes.data, a["nrm"].ctypes.data, a["uv"].ctypes.data
        d.num_triangles = a["idx"].shape[0]
        d.h_indices, d.h_tri_material = a["idx"].ctypes.data, a["tm"].ctypes.data
        d.num_materials = a["mats"].shape[0]
        d.h_materials = a["mats"].ctypes.data
        d.num_textures = a["td"].shape[0]
        d.h_tex_desc = a["td"].ctypes.data if a["td"].size else None
        d.num_texels_floats = a["tx"].size
        d.h_texels = a["tx"].ctypes.data if a["tx"].size else None
        d.h_camera = a["cam"].ctypes.data
        d.num_meshes = a["moff"].shape[0] - 1
        d.h_mesh_vertex_offsets = a["moff"].ctypes.data
        d.flags = _lib.SCENE_USE_COMPRESSED_HIT_INFO if use_compressed_hit_info else 0
        h = C.c_void_p()
        _lib.check(L.fhp_scene_create(C.byref(d), C.byref(h)))
        self.handle = h
        self.materials = a["mats"].copy()
        # Incremented by every parameter change (materials, textures, vertices, lights, sampling
        # freeze); AccumulatePass restarts when it moves (the scene updates of AccumulatePass.cpp:141-147).
        self.version = 0
        if getattr(data, "analytic_lights", None):
            self.set_analytic_lights(data.analytic_lights)
        if data.env_map is not None:
            e = data.env_map
            self.set_env_map(e["texels"], e.get("intensity", 1.0), e.get("tint", (1, 1, 1)), e.get("rotation", (0, 0, 0)))
        nm = a["moff"].shape[0] - 1
        xf = list(getattr(data, "mesh_transforms", None) or [])
        self._rest_xf = [np.asarray(m, np.float64) for m in xf] if len(xf) == nm else [np.eye(4) for _ in range(nm)]

    def close(self):
        if getattr(self, "handle", None):
            _lib.load().fhp_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> dict:
        i = _lib.SceneInfo()
        _lib.check(_lib.load().fhp_scene_info(self.handle, C.byref(i)))
        out = {k: getattr(i, k) for k, _ in _lib.SceneInfo._fields_ if k != "camera_basis"}
        out["camera_basis"] = np.array(i.camera_basis[:], np.float32)
        return out

    # Material records (the scene description's 20-float layout, see scene.py).
    def get_material_records(self) -> np.ndarray:
        return self.materials.copy()

    def set_material_records(self, mats: np.ndarray):
        m = np.ascontiguousarray(mats, np.float32)
        if m.shape != self.materials.shape:
            raise ValueError(f"expected materials of shape {self.materials.shape}")
        _lib.check(_lib.load().fhp_scene_set_materials(self.handle, m.ctypes.data))
        self.materials = m.copy()
        self.version += 1

    # Scene.get_material_params / set_material_params (Scene.cpp:4426-4476, 4620-4643): serialized
    # StandardMaterialParamLayout values (Material.PARAM_COUNT floats per material) of a list of
    # material ids. `params` may be a torch tensor or numpy array to fill, as the reference fills
    # its parameter buffer.
    @staticmethod
    def _ids(material_ids) -> np.ndarray:
        if hasattr(material_ids, "detach"):
            material_ids = material_ids.detach().cpu().numpy()
        return np.ascontiguousarray(np.asarray(material_ids).reshape(-1), np.uint32)

    def get_material_params(self, material_ids, params=None):
        ids = self._ids(material_ids)
        out = np.zeros((ids.size, Material.PARAM_COUNT), np.float32)
        _lib.check(_lib.load().fhp_scene_get_material_params(self.handle, ids.size, ids.ctypes.data, out.ctypes.data))
        if params is None:
            return out
        if hasattr(params, "copy_"):
            torch = _torch()
            params.view(-1)[: out.size].copy_(torch.from_numpy(out.reshape(-1)))
        else:
            np.asarray(params).reshape(-1)[: out.size] = out.reshape(-1)
        return params

    def set_material_params(self, material_ids, params):
        ids = self._ids(material_ids)
        if hasattr(params, "detach"):
            params = params.detach().float().cpu().numpy()
        p = np.ascontiguousarray(np.asarray(params, np.float32).reshape(-1)[: ids.size * Material.PARAM_COUNT])
        if p.size != ids.size * Material.PARAM_COUNT:
            raise ValueError(f"expected {ids.size * Material.PARAM_COUNT} parameters")
        L = _lib.load()
        _lib.check(L.fhp_scene_set_material_params(self.handle, ids.size, ids.ctypes.data, p.ctypes.data))
        _lib.check(L.fhp_scene_get_materials(self.handle, self.materials.ctypes.data))
        self.version += 1

    # Scene::getMeshVerticesAndIndices / setMeshVertices (Scene.cpp:4314-4336)
    def mesh_vertex_range(self, mesh_id: int) -> tuple[int, int]:
        off = self._arrays["moff"]
        if not 0 <= mesh_id < off.shape[0] - 1:
            raise ValueError(f"mesh id {mesh_id} out of range")
        return int(off[mesh_id]), int(off[mesh_id + 1])

    def get_mesh_vertices_and_indices(self, mesh_id: int):
        """(positions [V, 3], normals [V, 3], indices [T, 3] local to the mesh) as numpy arrays."""
        v0, v1 = self.mesh_vertex_range(mesh_id)
        idx = self._arrays["idx"]
        sel = np.all((idx >= v0) & (idx < v1), axis=1)
        return (self._arrays["pos"][v0:v1].copy(), self._arrays["nrm"][v0:v1].copy(), (idx[sel] - v0).astype(np.uint32))

    def set_mesh_vertices(self, mesh_id: int, positions, normals=None, update_mode: str = "refit"):
        """Replaces the mesh's vertex positions (and normals) (Scene::setMeshVertices).

        Device tensors are applied on the GPU: update_mode "refit" (the reference's default BLAS
        update mode) refits the BVH in place, "rebuild" rebuilds it; a mesh with emissive
        triangles always rebuilds (its lights move). Host arrays rebuild the BVH and lights."""
        if update_mode not in ("refit", "rebuild"):
            raise ValueError("update_mode is 'refit' or 'rebuild'")
        v0, v1 = self.mesh_vertex_range(mesh_id)
        def check(x):
            if x.reshape(-1).shape[0] != 3 * (v1 - v0):
                raise ValueError(f"mesh {mesh_id} has {v1 - v0} vertices, got {x.reshape(-1).shape[0] // 3}")
        if hasattr(positions, "is_cuda") and positions.is_cuda:
            torch = _torch()
            p = positions.detach().to(torch.float32).contiguous()
            check(p)
            n = normals.detach().to(device=p.device, dtype=torch.float32).contiguous() if normals is not None else None
            if n is not None:
                check(n)
            _lib.check(_lib.load().fhp_scene_update_vertices_device(
                self.handle, mesh_id, p.data_ptr(), n.data_ptr() if n is not None else None,
                int(update_mode == "rebuild"), _lib.stream_ptr()))
            p = p.cpu().numpy().reshape(-1, 3)
            n = n.cpu().numpy().reshape(-1, 3) if n is not None else None
        else:
            def host(x):
                if hasattr(x, "detach"):
                    x = x.detach().float().cpu().numpy()
                x = np.ascontiguousarray(x, np.float32).reshape(-1, 3)
                check(x)
                return x
            p = host(positions)
            n = host(normals) if normals is not None else None
            _lib.check(_lib.load().fhp_scene_update_vertices(self.handle, mesh_id, p.ctypes.data,
                                                             n.ctypes.data if n is not None else None))
        self._arrays["pos"][v0:v1] = p
        if n is not None:
            self._arrays["nrm"][v0:v1] = n
        self._rest_xf[mesh_id] = np.eye(4)  # the new world-space vertices are the mesh's rest pose
        self.version += 1

    # Scene::updateNodeTransform (Scene.cpp:4282-4289) and the animation update it starts.
    def update_node_transform(self, mesh_id: int, transform):
        """Places mesh `mesh_id` with a new local-to-world transform (a 4x4 matrix or a
        scene.Transform), as animating its scene-graph node does: the mesh's local vertices are
        those it was built with (SceneBuilder transform undone), or the vertices last given to
        set_mesh_vertices. On the GPU the BVH is refitted, the emissive triangles' positions,
        normals and areas are recomputed (flux kept, LightCollection::updateTrianglePositions)
        and the light BVH is refitted (LightBVH::refit). Mirroring transforms are refused."""
        self.mesh_vertex_range(mesh_id)
        T = transform.matrix() if hasattr(transform, "matrix") else np.asarray(transform, np.float64)
        if T.shape != (4, 4):
            raise ValueError("transform must be a 4x4 matrix or a Transform")
        M = T @ np.linalg.inv(self._rest_xf[mesh_id])
        m = np.ascontiguousarray(M[:3, :4], np.float32)
        L = _lib.load()
        _lib.check(L.fhp_scene_update_node_transform(self.handle, mesh_id, m.ctypes.data, _lib.stream_ptr()))
        v0, v1 = self.mesh_vertex_range(mesh_id)
        p = np.zeros((v1 - v0, 3), np.float32)
        n = np.zeros((v1 - v0, 3), np.float32)
        _lib.check(L.fhp_scene_get_mesh_vertices(self.handle, mesh_id, p.ctypes.data, n.ctypes.data))
        self._arrays["pos"][v0:v1] = p
        self._arrays["nrm"][v0:v1] = n
        self.version += 1

    def set_analytic_lights(self, records):
        """Analytic point / directional lights (Scene::addLight): records of scene.point_light /
        scene.directional_light, [n, 16]; an empty list removes them."""
        r = np.ascontiguousarray(np.asarray(records, np.float32).reshape(-1, 16))
        _lib.check(_lib.load().fhp_scene_set_analytic_lights(self.handle, r.shape[0], r.ctypes.data if r.size else None))
        self.analytic_lights = r.copy()
        self.version += 1

    def set_env_map(self, texels=None, intensity: float = 1.0, tint=(1.0, 1.0, 1.0), rotation=(0.0, 0.0, 0.0),
                    use_light: bool = True, use_background: bool = True):
        """Environment map (EnvMap): [H, W, 3 or 4] lat-long radiance (array or image path), intensity,
        tint, rotation in degrees XYZ; it lights rays that leave the scene (use_light) and is the
        background of primary misses (use_background). None removes it."""
        if texels is None:
            _lib.check(_lib.load().fhp_scene_set_env_map(self.handle, None, 0, 0, None, 0))
        else:
            if isinstance(texels, (str, Path)):
                from .image_io import read_rgba
                texels = read_rgba(texels)
            t = np.ascontiguousarray(np.asarray(texels, np.float32)[..., :3])
            if t.ndim != 3 or t.shape[2] != 3:
                raise ValueError("env map texels must be [H, W, 3 or 4]")
            prm = np.array([intensity, *tint, *rotation], np.float32)
            flags = (1 if use_light else 0) | (2 if use_background else 0)
            _lib.check(_lib.load().fhp_scene_set_env_map(self.handle, t.ctypes.data, t.shape[1], t.shape[0],
                                                         prm.ctypes.data, flags))
        self.version += 1

    def get_env_importance(self):
        """The env map's importance levels (EnvMapSampler): a list of square float32 arrays from
        1 x 1 (the mean) to the finest level; empty without a sampleable env map."""
        L = _lib.load()
        dim = C.c_uint32()
        _lib.check(L.fhp_scene_get_env_importance(self.handle, C.byref(dim), None))
        if dim.value == 0:
            return []
        buf = np.empty((4 * dim.value ** 2 - 1) // 3, np.float32)
        _lib.check(L.fhp_scene_get_env_importance(self.handle, C.byref(dim), buf.ctypes.data))
        out, r = [], 1
        while r <= dim.value:
            o = (r * r - 1) // 3
            out.append(buf[o:o + r * r].reshape(r, r))
            r *= 2
        return out

    def env_sample(self, u):
        """EnvMapSampler::sample for [n, 2] uniforms (device tensor): [n, 8] (dir, pdf, Le, valid)."""
        torch = _torch()
        u = u.detach().to(device=_device(), dtype=torch.float32).contiguous()
        out = torch.empty((u.shape[0], 8), dtype=torch.float32, device=u.device)
        _lib.check(_lib.load().fhp_env_sample(self.handle, u.data_ptr(), u.shape[0], out.data_ptr(), _lib.stream_ptr()))
        return out

    def env_eval_pdf(self, dirs):
        """EnvMapSampler::evalPdf for [n, 3] world directions (device tensor): [n]."""
        torch = _torch()
        d = dirs.detach().to(device=_device(), dtype=torch.float32).contiguous()
        out = torch.empty((d.shape[0],), dtype=torch.float32, device=d.device)
        _lib.check(_lib.load().fhp_env_eval_pdf(self.handle, d.data_ptr(), d.shape[0], out.data_ptr(), _lib.stream_ptr()))
        return out

    def set_texels(self, texels):
        """Copies a device tensor of texels (same layout as SceneData.texels) into the scene."""
        torch = _torch()
        t = texels.detach().to(device=_device(), dtype=torch.float32).contiguous()
        _lib.check(_lib.load().fhp_scene_set_texels(self.handle, t.data_ptr(), t.numel(), _lib.stream_ptr()))
        self.version += 1

    def freeze_sampling(self, enable: bool = True):
        """Finite-difference mode: keep sampling/pdfs at the current parameters (see fhp.h)."""
        _lib.check(_lib.load().fhp_scene_freeze_sampling(self.handle, int(enable), _lib.stream_ptr()))
        self.version += 1

    def lights(self) -> np.ndarray:
        """Light collection records [num_lights, 20] (see fhp_scene_lights)."""
        n = self.info()["num_lights"]
        out = np.zeros((n, 20), np.float32)
        if n:
            _lib.check(_lib.load().fhp_scene_lights(self.handle, out.ctypes.data))
        return out

    def light_bvh(self):
        i = self.info()
        nodes = np.zeros((i["light_bvh_nodes"], 16), np.float32)
        tri = np.zeros(max(i["num_active_lights"], 1), np.uint32)
        masks = np.zeros((max(i["num_lights"], 1), 2), np.uint32)
        _lib.check(_lib.load().fhp_scene_light_bvh(self.handle, nodes.ctypes.data, tri.ctypes.data, masks.ctypes.data))
        return nodes, tri[: i["num_active_lights"]], masks[: i["num_lights"]]

    # Ray queries (TraceRay / RayQuery replacements).
    def set_camera(self, camera):
        """Replaces the scene's camera (a scene.Camera) for later renders."""
        c = np.ascontiguousarray(camera.pack(), np.float32)
        _lib.check(_lib.load().fhp_scene_set_camera(self.handle, c.ctypes.data))
        self.version += 1

    TRACE_ERR_STACK_OVERFLOW = 1

    def trace_errors(self, reset: bool = True) -> int:
        """Trace error bits raised by this scene's traversal kernels since the last reset (1: a
        ray needed more stack entries than the tree's worst case and ended without a hit)."""
        v = C.c_uint32()
        _lib.check(_lib.load().fhp_scene_trace_errors(self.handle, C.byref(v), int(reset), _lib.stream_ptr()))
        return int(v.value)

    def trace_closest(self, rays):
        """rays: float32 tensor [N, 8] on the device -> int32-bit tensor [N, 4] (tri, t, b1, b2 bits)."""
        torch = _torch()
        r = rays.contiguous()
        hits = torch.empty((r.shape[0], 4), dtype=torch.int32, device=r.device)
        _lib.check(_lib.load().fhp_trace_closest(self.handle, r.data_ptr(), r.shape[0], hits.data_ptr(), _lib.stream_ptr()))
        return hits

    def trace_any(self, rays):
        torch = _torch()
        r = rays.contiguous()
        occ = torch.empty((r.shape[0],), dtype=torch.int32, device=r.device)
        _lib.check(_lib.load().fhp_trace_any(self.handle, r.data_ptr(), r.shape[0], occ.data_ptr(), _lib.stream_ptr()))
        return occ


# --------------------------------------------------------------------------- passes
class RenderPass:
    kind = ""
    outputs: tuple = ()

    def __init__(self, props: dict | None = None, name: str | None = None):
        self.name = name or self.kind
        self.properties = dict(props or {})
        self.scene: Scene | None = None
        self.frame_count = 0

    def set_scene(self, scene: Scene | None):
        self.scene = scene
        self.frame_count = 0


class _NativePass(RenderPass):
    def __init__(self, props=None, name=None):
        super().__init__(props, name)
        keys, vals = [], []
        for k, v in self.properties.items():
            v = self._encode(k, v)
            keys.append(k.encode())
            vals.append(float(v))
        karr = (C.c_char_p * max(len(keys), 1))(*keys)
        varr = (C.c_double * max(len(vals), 1))(*vals)
        h = C.c_void_p()
        _lib.check(_lib.load().fhp_pass_create(self.kind.encode(), len(keys), karr, varr, C.byref(h)))
        w = _lib.last_warning()
        if w:
            warnings.warn(w)
        self.handle = h

    @staticmethod
    def _encode(key, value):
        if key == "diffMode" and isinstance(value, str):
            return {"Primal": 0, "BackwardDiff": 1, "ForwardDiffDebug": 2, "BackwardDiffDebug": 3}[value]
        if key == "diffVarName" and isinstance(value, str):
            return {"": 0, "CBOX_BUNNY_MATERIAL": 1, "CBOX_BUNNY_TRANSLATION": 2}[value]
        if key == "emissiveSampler" and isinstance(value, str):
            return {"Uniform": 0, "LightBVH": 1}[value]
        return float(value)

    def __del__(self):
        if getattr(self, "handle", None):
            try:
                _lib.load().fhp_pass_destroy(self.handle)
            except Exception:
                pass
            self.handle = None

    def set_property(self, key: str, value):
        _lib.check(_lib.load().fhp_pass_set_property(self.handle, key.encode(), float(self._encode(key, value))))
        w = _lib.last_warning()
        if w:
            warnings.warn(w)
        self.properties[key] = value

    def get_property(self, key: str) -> float:
        v = C.c_double()
        _lib.check(_lib.load().fhp_pass_get_property(self.handle, key.encode(), C.byref(v)))
        return v.value

    KERNEL_KINDS = ("trace", "shadow", "shade", "war_primal", "war_adjoint", "other")

    def kernel_times(self, reset: bool = True) -> dict:
        """Per-kind kernel time (ms) and launch counts recorded with HIP events (property timeKernels=1)."""
        ms = np.zeros(6, np.float64)
        n = np.zeros(6, np.uint64)
        _lib.check(_lib.load().fhp_pass_kernel_times(self.handle, ms.ctypes.data, n.ctypes.data, int(reset)))
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(self.KERNEL_KINDS)}

    def render(self, width: int, height: int, spp: int = 1, frame: int | None = None, out=None, stats: bool = False):
        """Renders one frame of `spp` samples per pixel into a [H, W, 4] float32 tensor."""
        torch = _torch()
        dev = _device()
        if out is None:
            out = torch.empty((height, width, 4), dtype=torch.float32, device=dev)
        f = self.frame_count if frame is None else frame
        st = np.zeros(8, np.uint64) if stats else None
        _lib.check(_lib.load().fhp_pass_execute(self.handle, self.scene.handle if self.scene else None, width, height, spp, f,
                                                out.data_ptr(), st.ctypes.data if stats else None, _lib.stream_ptr()))
        if frame is None:
            self.frame_count += 1
        return (out, st) if stats else out

    def execute(self, render_data: dict):
        out = render_data["color"]
        h, w = out.shape[0], out.shape[1]
        self.render(w, h, 1, out=out)

    def render_batch(self, views: list[dict], stats: bool = False):
        """Renders several views in one launch sequence (fhp_pass_execute_batch). Each view is a dict
        with scene, width, height and optional spp (1), frame (0), camera (a scene.Camera, default
        the scene's) and out ([H, W, 4] float32 tensor). Returns the output tensors (and stats).
        WARDiffPathTracer views share one scene, width, height and spp."""
        torch = _torch()
        dev = _device()
        arr = (_lib.View * len(views))()
        outs, cams = [], []
        for i, v in enumerate(views):
            w, h = int(v["width"]), int(v["height"])
            out = v.get("out")
            if out is None:
                out = torch.empty((h, w, 4), dtype=torch.float32, device=dev)
            outs.append(out)
            cam = v.get("camera")
            if cam is not None:
                cams.append(np.ascontiguousarray(cam.pack(), np.float32))
            arr[i].scene = v["scene"].handle
            arr[i].h_camera = cams[-1].ctypes.data if cam is not None else None
            arr[i].width, arr[i].height = w, h
            arr[i].spp, arr[i].frame = int(v.get("spp", 1)), int(v.get("frame", 0))
            arr[i].d_color = out.data_ptr()
        st = np.zeros(8, np.uint64) if stats else None
        _lib.check(_lib.load().fhp_pass_execute_batch(self.handle, len(views), arr, st.ctypes.data if stats else None,
                                                      _lib.stream_ptr()))
        return (outs, st) if stats else outs


class MinimalPathTracer(_NativePass):
    """MinimalPathTracer (MinimalPathTracer.cpp). Properties: maxBounces, computeDirect,
    useImportanceSampling; extensions useEmissiveNEE, useEnvNEE (env map light sampling),
    emissiveSampler ('Uniform'|'LightBVH'). With several light types enabled (env map, emissive,
    analytic) one is chosen per vertex with equal probability. tailPaths (default 65536, 0 off):
    a bounce depth with at most that many hits runs its paths to their end in one launch (same
    image, bit for bit), tailLanes (default 8) of them per wave."""

    kind = "MinimalPathTracer"
    outputs = ("color",)

    def set_input(self, channel: str, tensor):
        """Binds the "vbuffer" (PackedHitInfo per pixel, as VBufferRT writes it) or "viewW" input
        for later frames; None unbinds. The pass keeps a reference to the tensor."""
        if not hasattr(self, "_inputs"):
            self._inputs = {}
        ptr = None
        if tensor is not None:
            if not tensor.is_contiguous():
                raise ValueError(f"input '{channel}' must be contiguous")
            ptr = tensor.data_ptr()
        _lib.check(_lib.load().fhp_pass_set_input(self.handle, channel.encode(), ptr))
        if tensor is None:
            self._inputs.pop(channel, None)
        else:
            self._inputs[channel] = tensor

    def render(self, width: int, height: int, spp: int = 1, frame: int | None = None, out=None, stats: bool = False):
        for k, t in getattr(self, "_inputs", {}).items():
            want = (height, width)
            if tuple(t.shape[:2]) != want:
                raise ValueError(f"input '{k}' is {tuple(t.shape[:2])}, frame is {want}")
            if k == "vbuffer" and self.scene is not None:
                words = 2 if self.scene.info()["hit_info_compressed"] else 4
                if t.dim() != 3 or t.shape[2] != words or t.element_size() != 4:
                    raise ValueError(f"vbuffer must be [H, W, {words}] 32-bit words for this scene")
            if k == "viewW" and (t.dim() != 3 or t.shape[2] != 4 or t.dtype != _torch().float32):
                raise ValueError("viewW must be [H, W, 4] float32")
        return super().render(width, height, spp, frame=frame, out=out, stats=stats)

class VBufferRT(_NativePass):
    """VBufferRT (Source/RenderPasses/GBuffer/VBuffer/VBufferRT.cpp): per-pixel primary hits as
    PackedHitInfo. Properties samplePattern ('Center' | 'DirectX' | 'Halton' | 'Stratified'),
    sampleCount, useAlphaTest, adjustShadingNormals, forceCullMode, cull, useTraceRayInline, useDOF.
    `execute_frame(width, height)` returns {"vbuffer": [H, W, 2 or 4] int32 (the uint32 words),
    "depth": [H, W], "viewW": [H, W, 4], "mask": [H, W]} and advances the camera jitter pattern.
    The output tensors belong to the pass and are overwritten by the next frame of the same size
    (as render-pass output textures are), so bound consumers see stable device addresses."""

    kind = "VBufferRT"
    outputs = ("vbuffer", "depth", "viewW", "mask")
    SAMPLE_PATTERNS = {"Center": 0, "DirectX": 1, "Halton": 2, "Stratified": 3}
    CULL_MODES = {"None": 0, "Front": 1, "Back": 2}

    @classmethod
    def _encode(cls, key, value):
        if key == "samplePattern" and isinstance(value, str):
            return cls.SAMPLE_PATTERNS[value]
        if key == "cull" and isinstance(value, str):
            return cls.CULL_MODES[value]
        return float(value)

    def execute_frame(self, width: int, height: int, outputs=("vbuffer", "depth", "viewW", "mask")) -> dict:
        torch = _torch()
        dev = _device()
        if self.scene is None:
            raise _lib.FhpError("VBufferRT needs a scene")
        words = 2 if self.scene.info()["hit_info_compressed"] else 4
        shapes = {"vbuffer": ((height, width, words), torch.int32), "depth": ((height, width), torch.float32),
                  "viewW": ((height, width, 4), torch.float32), "mask": ((height, width), torch.float32)}
        if not hasattr(self, "_bufs"):
            self._bufs = {}
        out = {}
        for k, (shape, dt) in shapes.items():
            if k != "vbuffer" and k not in outputs:
                continue
            b = self._bufs.get(k)
            if b is None or tuple(b.shape) != shape or b.device != dev:
                b = self._bufs[k] = torch.empty(shape, dtype=dt, device=dev)
            out[k] = b
        jit = (C.c_float * 2)()
        ptr = lambda k: out[k].data_ptr() if k in out else None
        _lib.check(_lib.load().fhp_vbuffer_execute(self.handle, self.scene.handle, width, height, ptr("vbuffer"),
                                                   ptr("depth"), ptr("viewW"), ptr("mask"), jit, _lib.stream_ptr()))
        self.jitter = (jit[0], jit[1])
        self.frame_count += 1
        return out


class WARDiffPathTracer(_NativePass):
    """WARDiffPathTracer (WARDiffPathTracer.cpp). Outputs color and dColor."""

    kind = "WARDiffPathTracer"
    outputs = ("color", "dColor")

    def __init__(self, props=None, name=None):
        super().__init__(props, name)
        self.scene_gradients: SceneGradients | None = None
        self.run_backward = 0
        self.dL_dI = None
        self.mesh_to_optimize = 0

    DIFF_VAR_NONE, DIFF_VAR_MATERIAL, DIFF_VAR_GEOMETRY_TRANSLATION = 0, 1, 2

    def set_diff_debug_params(self, var_type: int, id: int = 0, offset: int = 0, grad=(0.0, 0.0, 0.0, 0.0)):
        """WARDiffPathTracer::setDiffDebugParams (WARDiffPathTracer.cpp:689-695): the variable of the
        gradient-image debug modes (DIFF_VAR_MATERIAL: the parameter entry at `offset` of material
        `id`'s 20-float layout along `grad`; DIFF_VAR_GEOMETRY_TRANSLATION: mesh `id` moved along
        grad[:3])."""
        g = np.zeros(4, np.float32)
        g[:len(grad)] = grad
        _lib.check(_lib.load().fhp_pass_set_diff_debug(self.handle, int(var_type), int(id), int(offset), g.ctypes.data))

    def render(self, width: int, height: int, spp: int = 1, frame: int | None = None, out=None, stats: bool = False,
               dcolor=None):
        """Renders color ([H, W, 4] float32); with `dcolor` (a [H, W, 4] float32 tensor) bound as the
        dColor output, the gradient-image debug modes write it as well."""
        L = _lib.load()
        _lib.check(L.fhp_pass_set_output(self.handle, b"dColor", dcolor.data_ptr() if dcolor is not None else None))
        try:
            return super().render(width, height, spp, frame, out, stats)
        finally:
            _lib.check(L.fhp_pass_set_output(self.handle, b"dColor", None))

    def set_mesh_to_optimize(self, mesh_id: int):
        """WARDiffPathTracer::setMeshToOptimize (WARDiffPathTracer.cpp:697-715): vertex gradients
        (useWAR) flow into this mesh only."""
        self.mesh_to_optimize = int(mesh_id)

    def backward(self, width: int, height: int, spp: int = 1, frame: int | None = None, stats: bool = False):
        """Accumulates d(sum dL_dI . color)/d(params) into scene_gradients (raw sum over samples).
        With useWAR, vertex position/normal gradients of the mesh set by set_mesh_to_optimize
        are accumulated too (maxBounces up to 8, reflective materials)."""
        debug = self.get_property("diffMode") == 3  # BackwardDiffDebug: dL/dI = (1, 0, 0) everywhere
        if self.scene_gradients is None or (self.dL_dI is None and not debug):
            raise _lib.FhpError("backward needs scene_gradients and dL_dI")
        if self.get_property("useWAR") != 0 and self.scene_gradients.mesh_id != self.mesh_to_optimize:
            self.scene_gradients.enable_mesh(self.mesh_to_optimize)
        torch = _torch()
        d = None
        if not debug:
            d = self.dL_dI.detach().to(device=_device(), dtype=torch.float32).contiguous()
            if d.numel() != width * height * 3:
                raise ValueError("dL_dI must hold width * height * 3 floats")
        f = self.frame_count if frame is None else frame
        st = np.zeros(6, np.uint64) if stats else None
        _lib.check(_lib.load().fhp_pass_backward(self.handle, self.scene.handle, width, height, spp, f,
                                                 d.data_ptr() if d is not None else None,
                                                 self.scene_gradients.handle, st.ctypes.data if stats else None,
                                                 _lib.stream_ptr()))
        if frame is None:
            self.frame_count += 1
        return st

    def backward_batch(self, views: list[dict], width: int, height: int, spp: int = 1, stats: bool = False):
        """backward() for several views of the scene in one launch sequence (fhp_pass_backward_batch):
        each view is a dict with dL_dI ([H, W, 3] float32), optional frame (0) and camera (a
        scene.Camera, default the scene's). All views accumulate into scene_gradients."""
        debug = self.get_property("diffMode") == 3
        if self.scene_gradients is None:
            raise _lib.FhpError("backward needs scene_gradients")
        if self.get_property("useWAR") != 0 and self.scene_gradients.mesh_id != self.mesh_to_optimize:
            self.scene_gradients.enable_mesh(self.mesh_to_optimize)
        torch = _torch()
        arr = (_lib.DiffView * len(views))()
        keep, cams = [], []
        for i, v in enumerate(views):
            d = v.get("dL_dI")
            if d is None and not debug:
                raise _lib.FhpError("backward needs dL_dI for every view")
            if d is not None:
                d = d.detach().to(device=_device(), dtype=torch.float32).contiguous()
                if d.numel() != width * height * 3:
                    raise ValueError("dL_dI must hold width * height * 3 floats")
                keep.append(d)
            cam = v.get("camera")
            if cam is not None:
                cams.append(np.ascontiguousarray(cam.pack(), np.float32))
            arr[i].h_camera = cams[-1].ctypes.data if cam is not None else None
            arr[i].frame = int(v.get("frame", 0))
            arr[i].d_dLdI = d.data_ptr() if d is not None else None
        st = np.zeros(6, np.uint64) if stats else None
        _lib.check(_lib.load().fhp_pass_backward_batch(self.handle, self.scene.handle, len(views), arr, width, height, spp,
                                                       self.scene_gradients.handle, st.ctypes.data if stats else None,
                                                       _lib.stream_ptr()))
        return st


class AccumulatePass(_NativePass):
    """AccumulatePass (Source/RenderPasses/AccumulatePass/AccumulatePass.cpp) on the GPU.

    Properties: enabled, autoReset, precisionMode ('Double' | 'Single' | 'SingleCompensated'),
    maxFrameCount (0 = no limit), overflowMode ('Stop' | 'Reset' | 'EMA'); the deprecated
    'enableAccumulation' is accepted as the reference does. `accumulate(frame)` adds one
    [H, W, 4] frame (float32, int32 or uint32) and returns the accumulated [H, W, 4] float32
    image, a buffer the pass keeps (with overflowMode 'Stop' it keeps its last contents once
    maxFrameCount frames are in). With autoReset, a change of the scene's parameters (materials,
    textures, mesh vertices) since the last frame restarts the accumulation, as scene updates do
    in AccumulatePass::execute (:129-155); `refresh=True` forces the same.
    """

    kind = "AccumulatePass"
    outputs = ("output",)
    _FORMATS = {"float32": 0, "uint32": 1, "int32": 2}

    def __init__(self, props=None, name=None):
        super().__init__(props, name)
        self._out = None
        self._scene_version = None

    @staticmethod
    def _encode(key, value):
        if key == "precisionMode" and isinstance(value, str):
            return {"Double": 0, "Single": 1, "SingleCompensated": 2}[value]
        if key == "overflowMode" and isinstance(value, str):
            return {"Stop": 0, "Reset": 1, "EMA": 2}[value]
        return float(value)

    def set_scene(self, scene):
        """A new scene restarts accumulation (AccumulatePass::setScene, :351-357)."""
        super().set_scene(scene)
        self._scene_version = scene.version if scene is not None else None
        self.reset()

    def reset(self):
        _lib.check(_lib.load().fhp_accumulate_reset(self.handle))

    @property
    def enabled(self) -> bool:
        return self.get_property("enabled") != 0

    @enabled.setter
    def enabled(self, value: bool):
        self.set_property("enabled", int(bool(value)))

    @property
    def frame_count(self) -> int:
        n = C.c_uint32()
        _lib.check(_lib.load().fhp_accumulate_frame_count(self.handle, C.byref(n)))
        return n.value

    @frame_count.setter
    def frame_count(self, value):
        pass  # the count lives in the native pass; RenderPass.__init__ assigns it

    def accumulate(self, cur, refresh: bool = False):
        torch = _torch()
        fmt = self._FORMATS.get(str(cur.dtype).replace("torch.", ""))
        if fmt is None:
            raise ValueError("AccumulatePass input must be float32, uint32 or int32")
        if cur.dim() != 3 or cur.shape[2] != 4:
            raise ValueError("AccumulatePass input must be [H, W, 4]")
        cur = cur.contiguous()
        h, w = cur.shape[0], cur.shape[1]
        if self._out is None or self._out.shape[:2] != (h, w):
            self._out = torch.zeros((h, w, 4), dtype=torch.float32, device=cur.device)
        if self.scene is not None and self._scene_version != self.scene.version:
            self._scene_version = self.scene.version
            refresh = True
        _lib.check(_lib.load().fhp_accumulate(self.handle, cur.data_ptr(), fmt, w, h, self._out.data_ptr(), int(bool(refresh)),
                                              _lib.stream_ptr()))
        return self._out


class SceneGradients:
    """SceneGradients (Source/Falcor/DiffRendering/SceneGradients.cpp:42-206): material, per-texel
    baseColor, and the MeshPosition / MeshNormal gradients of one mesh."""

    MATERIAL = "Material"
    TEXTURE = "Texture"
    MESH_POSITION = "MeshPosition"
    MESH_NORMAL = "MeshNormal"

    def __init__(self, scene: Scene, hash_size: int = 256, texture_grads: bool = True, mesh_id: int | None = None):
        h = C.c_void_p()
        _lib.check(_lib.load().fhp_grads_create(scene.handle, hash_size, int(texture_grads), C.byref(h)))
        self.handle = h
        self.scene = scene
        self.hash_size = hash_size
        self.mesh_id = None
        if mesh_id is not None:
            self.enable_mesh(mesh_id)

    def enable_mesh(self, mesh_id: int):
        """Allocates the MeshPosition / MeshNormal gradients of mesh_id (3 floats per vertex each)."""
        _lib.check(_lib.load().fhp_grads_set_mesh(self.handle, int(mesh_id)))
        self.mesh_id = int(mesh_id)

    def __del__(self):
        if getattr(self, "handle", None):
            try:
                _lib.load().fhp_grads_destroy(self.handle)
            except Exception:
                pass
            self.handle = None

    def clear_all_grads(self):
        _lib.check(_lib.load().fhp_grads_clear(self.handle, _lib.stream_ptr()))

    def aggregate_all_grads(self):
        _lib.check(_lib.load().fhp_grads_aggregate(self.handle, _lib.stream_ptr()))

    def get_grad_types(self):
        types = [self.MATERIAL, self.TEXTURE]
        if self.mesh_id is not None:
            types += [self.MESH_POSITION, self.MESH_NORMAL]
        return types

    def get_grads_buffer(self, grad_type: str):
        """Device tensor of the aggregated gradients (copied to a new tensor)."""
        torch = _torch()
        if grad_type in (self.MESH_POSITION, self.MESH_NORMAL):
            if self.mesh_id is None:
                raise _lib.FhpError("mesh gradients are not enabled (enable_mesh / set_mesh_to_optimize)")
            v0, v1 = self.scene.mesh_vertex_range(self.mesh_id)
            n = 3 * (v1 - v0)
            which = 2 if grad_type == self.MESH_POSITION else 3
        else:
            dm, nm, dt, nt = C.c_void_p(), C.c_uint64(), C.c_void_p(), C.c_uint64()
            _lib.check(_lib.load().fhp_grads_get(self.handle, C.byref(dm), C.byref(nm), C.byref(dt), C.byref(nt)))
            n = nm.value if grad_type == self.MATERIAL else nt.value
            which = 0 if grad_type == self.MATERIAL else 1
        out = torch.empty(n, dtype=torch.float32, device=_device())
        if n:
            _lib.check(_lib.load().fhp_grads_copy(self.handle, which, out.data_ptr(), n, _lib.stream_ptr()))
        return out


# --------------------------------------------------------------------------- graph / testbed
_PASS_TYPES = {"MinimalPathTracer": MinimalPathTracer, "WARDiffPathTracer": WARDiffPathTracer,
               "AccumulatePass": AccumulatePass, "VBufferRT": VBufferRT}


class RenderGraph:
    def __init__(self, name: str = "RenderGraph"):
        self.name = name
        self.passes: dict[str, RenderPass] = {}
        self.edges: list[tuple[str, str]] = []
        self.marked: list[str] = []
        self._outputs: dict[str, object] = {}

    def create_pass(self, name: str, pass_type: str, props: dict | None = None) -> RenderPass:
        if pass_type not in _PASS_TYPES:
            raise ValueError(f"unknown pass type '{pass_type}'")
        p = _PASS_TYPES[pass_type](props, name)
        self.passes[name] = p
        return p

    def add_edge(self, src: str, dst: str):
        self.edges.append((src, dst))

    def mark_output(self, name: str):
        self.marked.append(name)

    def get_output(self, name: str):
        return self._outputs[name]

    def _order(self):
        deps = {n: set() for n in self.passes}
        for s, d in self.edges:
            deps[d.split(".")[0]].add(s.split(".")[0])
        done, order = set(), []
        while len(order) < len(self.passes):
            ready = [n for n in self.passes if n not in done and deps[n] <= done]
            if not ready:
                raise RuntimeError("render graph has a cycle")
            for n in ready:
                done.add(n)
                order.append(n)
        return order

    def execute(self, scene: Scene | None, width: int, height: int):
        torch = _torch()
        vals: dict[str, object] = {}
        for n in self._order():
            p = self.passes[n]
            if isinstance(p, AccumulatePass):
                if p.scene is not scene:
                    p.set_scene(scene)
                srcs = [s for s, d in self.edges if d == f"{n}.input"]
                if srcs:
                    vals[f"{n}.output"] = p.accumulate(vals[srcs[0]])
            elif isinstance(p, VBufferRT):
                if p.scene is not scene:
                    p.set_scene(scene)
                for k, t in p.execute_frame(width, height).items():
                    vals[f"{n}.{k}"] = t
            elif isinstance(p, _NativePass):
                if p.scene is not scene:
                    p.set_scene(scene)
                if isinstance(p, MinimalPathTracer):
                    # Input channels (MinimalPathTracer.cpp:48-53) from the graph's edges.
                    for ch in ("vbuffer", "viewW"):
                        srcs = [s for s, d in self.edges if d == f"{n}.{ch}"]
                        p.set_input(ch, vals[srcs[0]] if srcs else None)
                if isinstance(p, WARDiffPathTracer) and p.run_backward:
                    p.backward(width, height, 1)
                    vals[f"{n}.dColor"] = torch.zeros((height, width, 4), device=_device())
                    vals[f"{n}.color"] = torch.zeros((height, width, 4), device=_device())
                elif isinstance(p, WARDiffPathTracer):
                    dcolor = torch.zeros((height, width, 4), device=_device())
                    vals[f"{n}.color"] = p.render(width, height, 1, dcolor=dcolor)
                    vals[f"{n}.dColor"] = dcolor
                else:
                    vals[f"{n}.color"] = p.render(width, height, 1)
        self._outputs = vals
        return vals


class Testbed:
    """Headless frame driver (Testbed.cpp:72-112)."""

    def __init__(self, width: int = 512, height: int = 512, create_window: bool = False, gpu: int = 0):
        self.width, self.height = width, height
        self.scene: Scene | None = None
        self.render_graph: RenderGraph | None = None
        self.frame_count = 0

    def load_scene(self, data, use_compressed_hit_info: bool = False) -> Scene:
        """data: SceneData, or the path of a .pyscene file (Testbed::loadScene; the camera's aspect
        ratio follows the frame size, as the reference sets it from the frame dimensions)."""
        if isinstance(data, (str, Path)):
            from .pyscene import load_pyscene
            warns = []
            data = load_pyscene(data, aspect=self.width / self.height, warnings=warns)
            for w in warns:
                warnings.warn(w)
        self.scene = Scene(data, use_compressed_hit_info=use_compressed_hit_info)
        return self.scene

    def create_render_graph(self, name: str = "RenderGraph") -> RenderGraph:
        return RenderGraph(name)

    def frame(self):
        if self.render_graph is None:
            raise RuntimeError("no render graph")
        self.render_graph.execute(self.scene, self.width, self.height)
        self.frame_count += 1
