"""ctypes binding of libfhp.so (the C ABI declared in include/fhp.h).

The library is loaded from this package directory. There is no fallback: if libfhp.so is
missing or fails to load, every entry point raises, so nothing silently runs elsewhere.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

_LIB_PATH = Path(__file__).resolve().parent / "libfhp.so"
_lib = None

FHP_OK = 0


class FhpError(RuntimeError):
    pass


class SceneDesc(C.Structure):
    _fields_ = [
        ("num_vertices", C.c_uint32),
        ("h_positions", C.c_void_p),
        ("h_normals", C.c_void_p),
        ("h_texcoords", C.c_void_p),
        ("num_triangles", C.c_uint32),
        ("h_indices", C.c_void_p),
        ("h_tri_material", C.c_void_p),
        ("num_materials", C.c_uint32),
        ("h_materials", C.c_void_p),
        ("num_textures", C.c_uint32),
        ("h_tex_desc", C.c_void_p),
        ("num_texels_floats", C.c_uint64),
        ("h_texels", C.c_void_p),
        ("h_camera", C.c_void_p),
        ("num_meshes", C.c_uint32),
        ("h_mesh_vertex_offsets", C.c_void_p),
        ("flags", C.c_uint32),
    ]


SCENE_USE_COMPRESSED_HIT_INFO = 0x8000


class SceneInfo(C.Structure):
    _fields_ = [
        ("num_triangles", C.c_uint32), ("num_vertices", C.c_uint32), ("num_materials", C.c_uint32),
        ("num_textures", C.c_uint32), ("bvh_nodes", C.c_uint32), ("bvh_leaves", C.c_uint32),
        ("bvh_depth", C.c_uint32), ("stack_need", C.c_uint32), ("num_lights", C.c_uint32),
        ("num_active_lights", C.c_uint32), ("light_bvh_nodes", C.c_uint32), ("num_meshes", C.c_uint32),
        ("bvh_sah_cost", C.c_double),
        ("device_bytes", C.c_uint64), ("camera_basis", C.c_float * 9),
        ("hit_info_compressed", C.c_uint32), ("hit_info_type_bits", C.c_uint32),
        ("hit_info_instance_bits", C.c_uint32), ("hit_info_primitive_bits", C.c_uint32),
        ("meshes_contiguous", C.c_uint32),
    ]


class View(C.Structure):
    _fields_ = [
        ("scene", C.c_void_p),
        ("h_camera", C.c_void_p),
        ("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32), ("frame", C.c_uint32),
        ("d_color", C.c_void_p),
    ]


class DiffView(C.Structure):
    _fields_ = [("h_camera", C.c_void_p), ("frame", C.c_uint32), ("d_dLdI", C.c_void_p)]


# name -> (restype, argtypes)
_SIGS = {
    "fhp_get_version": (C.c_int, []),
    "fhp_last_error": (C.c_char_p, []),
    "fhp_last_warning": (C.c_char_p, []),
    "fhp_scene_create": (C.c_int, [C.POINTER(SceneDesc), C.POINTER(C.c_void_p)]),
    "fhp_scene_destroy": (C.c_int, [C.c_void_p]),
    "fhp_scene_info": (C.c_int, [C.c_void_p, C.POINTER(SceneInfo)]),
    "fhp_scene_trace_errors": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.c_int, C.c_void_p]),
    "fhp_scene_set_camera": (C.c_int, [C.c_void_p, C.c_void_p]),
    "fhp_scene_update_vertices": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]),
    "fhp_scene_set_materials": (C.c_int, [C.c_void_p, C.c_void_p]),
    "fhp_scene_set_texels": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]),
    "fhp_scene_freeze_sampling": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "fhp_scene_light_bvh": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "fhp_trace_closest": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]),
    "fhp_trace_any": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]),
    "fhp_pass_create": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.c_char_p), C.POINTER(C.c_double),
                                  C.POINTER(C.c_void_p)]),
    "fhp_pass_destroy": (C.c_int, [C.c_void_p]),
    "fhp_pass_set_property": (C.c_int, [C.c_void_p, C.c_char_p, C.c_double]),
    "fhp_pass_get_property": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_double)]),
    "fhp_pass_execute": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p,
                                   C.c_void_p, C.c_void_p]),
    "fhp_pass_execute_batch": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(View), C.c_void_p, C.c_void_p]),
    "fhp_pass_execute_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                           C.c_void_p, C.c_void_p]),
    "fhp_pass_kernel_times": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]),
    "fhp_grads_create": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int, C.POINTER(C.c_void_p)]),
    "fhp_grads_destroy": (C.c_int, [C.c_void_p]),
    "fhp_grads_clear": (C.c_int, [C.c_void_p, C.c_void_p]),
    "fhp_grads_aggregate": (C.c_int, [C.c_void_p, C.c_void_p]),
    "fhp_grads_get": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_uint64), C.POINTER(C.c_void_p),
                                C.POINTER(C.c_uint64)]),
    "fhp_grads_set_mesh": (C.c_int, [C.c_void_p, C.c_uint32]),
    "fhp_grads_copy": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_uint64, C.c_void_p]),
    "fhp_pass_backward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p,
                                    C.c_void_p, C.c_void_p, C.c_void_p]),
    "fhp_pass_backward_batch": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(DiffView), C.c_uint32, C.c_uint32,
                                          C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "fhp_accumulate": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_uint32, C.c_uint32, C.c_void_p, C.c_uint32,
                                 C.c_void_p]),
    "fhp_accumulate_reset": (C.c_int, [C.c_void_p]),
    "fhp_vbuffer_execute": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_void_p]),
    "fhp_pass_set_input": (C.c_int, [C.c_void_p, C.c_char_p, C.c_void_p]),
    "fhp_pass_set_output": (C.c_int, [C.c_void_p, C.c_char_p, C.c_void_p]),
    "fhp_pass_set_diff_debug": (C.c_int, [C.c_void_p, C.c_int, C.c_uint32, C.c_uint32, C.c_void_p]),
    "fhp_scene_lights": (C.c_int, [C.c_void_p, C.c_void_p]),
    "fhp_scene_set_env_map": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_uint32]),
    "fhp_scene_get_env_importance": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "fhp_env_sample": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]),
    "fhp_env_eval_pdf": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]),
    "fhp_scene_update_vertices_device": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]),
    "fhp_scene_update_node_transform": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]),
    "fhp_session_create": (C.c_int, [C.c_int, C.c_void_p]),
    "fhp_session_stream": (C.c_int, [C.c_void_p, C.c_void_p]),
    "fhp_session_destroy": (C.c_int, [C.c_void_p]),
    "fhp_bvh_build": (C.c_int, [C.c_void_p, C.c_uint32]),
    "fhp_scene_set_analytic_lights": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p]),
    "fhp_scene_get_materials": (C.c_int, [C.c_void_p, C.c_void_p]),
    "fhp_scene_get_material_params": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]),
    "fhp_scene_set_material_params": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]),
    "fhp_scene_get_mesh_vertices": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]),
    "fhp_accumulate_frame_count": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32)]),
    "fhp_image_compare": (C.c_int, [C.c_int, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_int, C.c_void_p,
                                    C.c_void_p, C.POINTER(C.c_double), C.c_void_p]),
}

EXPORTED = tuple(_SIGS)


def lib_path() -> Path:
    return _LIB_PATH


def load():
    """Loads libfhp.so (raises FhpError if it is missing)."""
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            raise FhpError(f"{_LIB_PATH} is missing: build it with `python -m falcor_amd.build_lib`")
        # Load torch's HIP runtime first so libfhp binds to the same libamdhip64 instance as
        # the tensors and streams it is handed (a second runtime copy would not see them).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = C.CDLL(str(_LIB_PATH))
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(rc: int):
    if rc != FHP_OK:
        msg = load().fhp_last_error().decode(errors="replace")
        raise FhpError(f"fhp error {rc}: {msg}")


def last_warning() -> str:
    return load().fhp_last_warning().decode(errors="replace")


def ptr(a) -> int:
    """Address of a contiguous numpy array or torch tensor."""
    if isinstance(a, np.ndarray):
        assert a.flags["C_CONTIGUOUS"]
        return a.ctypes.data
    return int(a.data_ptr())


def stream_ptr(stream=None) -> int | None:
    if stream is None:
        try:
            import torch
            if torch.cuda.is_available():
                return int(torch.cuda.current_stream().cuda_stream) or None
        except Exception:
            pass
        return None
    return int(getattr(stream, "cuda_stream", stream)) or None
