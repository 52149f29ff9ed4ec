"""C-ABI checks that need no GPU: the library loads, exports every entry point declared in
include/fhp.h, and reports argument errors with codes and messages (no device calls)."""
import ctypes as C

import numpy as np
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def _declared():
    text = (ROOT / "include" / "fhp.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(fhp_\w+)\(", text, re.M)))


def test_header_declares_expected_entry_points():
    names = _declared()
    for n in ("fhp_scene_create", "fhp_trace_closest", "fhp_trace_any", "fhp_pass_execute", "fhp_pass_backward",
              "fhp_grads_create", "fhp_grads_aggregate"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from falcor_amd import _lib
    lib = _lib.load()
    for n in _declared():
        assert hasattr(lib, n), n
    assert set(_declared()) == set(_lib.EXPORTED)


def test_argument_errors_without_device():
    from falcor_amd import _lib
    lib = _lib.load()
    h = C.c_void_p()
    assert lib.fhp_scene_create(None, C.byref(h)) == -1
    assert "null" in lib.fhp_last_error().decode()
    keys = (C.c_char_p * 1)(b"maxBounces")
    vals = (C.c_double * 1)(3.0)
    assert lib.fhp_pass_create(b"NoSuchPass", 1, keys, vals, C.byref(h)) == -1
    assert "unknown pass kind" in lib.fhp_last_error().decode()


def test_pass_properties_defaults_and_unknown_keys():
    """Unknown properties warn (MinimalPathTracer.cpp:86); defaults follow MinimalPathTracer.h:79-83."""
    from falcor_amd import _lib
    lib = _lib.load()
    keys = (C.c_char_p * 2)(b"maxBounces", b"bogus")
    vals = (C.c_double * 2)(5.0, 1.0)
    h = C.c_void_p()
    assert lib.fhp_pass_create(b"MinimalPathTracer", 2, keys, vals, C.byref(h)) == 0
    assert "bogus" in lib.fhp_last_warning().decode()
    v = C.c_double()
    assert lib.fhp_pass_get_property(h, b"maxBounces", C.byref(v)) == 0 and v.value == 5.0
    assert lib.fhp_pass_get_property(h, b"computeDirect", C.byref(v)) == 0 and v.value == 1.0
    assert lib.fhp_pass_set_property(h, b"maxBounces", 99.0) == -1
    lib.fhp_pass_destroy(h)
    keys = (C.c_char_p * 1)(b"maxBounces")
    vals = (C.c_double * 1)(3.0)
    assert lib.fhp_pass_create(b"WARDiffPathTracer", 1, keys, vals, C.byref(h)) == 0
    assert lib.fhp_pass_get_property(h, b"Log10vMFConcentration", C.byref(v)) == 0 and v.value == 5.0
    lib.fhp_pass_destroy(h)


def test_traversal_launch_properties_are_validated():
    """traceBlockSize / traceStackDepth / traceBlocksPerCU accept only compiled variants that fit a CU."""
    from falcor_amd import _lib
    from falcor_amd import autotune
    lib = _lib.load()
    h = C.c_void_p()
    assert lib.fhp_pass_create(b"MinimalPathTracer", 0, None, None, C.byref(h)) == 0
    v = C.c_double()
    for k, d in ((b"traceBlockSize", 128.0), (b"traceStackDepth", 24.0), (b"traceBlocksPerCU", 8.0)):
        assert lib.fhp_pass_get_property(h, k, C.byref(v)) == 0 and v.value == d
    assert lib.fhp_pass_set_property(h, b"traceBlockSize", 96.0) == -1
    assert lib.fhp_pass_set_property(h, b"traceStackDepth", 20.0) == -1
    assert lib.fhp_pass_set_property(h, b"traceBlocksPerCU", 0.0) == -1
    assert lib.fhp_pass_set_property(h, b"traceBlocksPerCU", 17.0) == -1   # 17 * 128 > 2048 threads per CU
    assert lib.fhp_pass_set_property(h, b"traceBlocksPerCU", 16.0) == 0
    assert lib.fhp_pass_set_property(h, b"traceBlockSize", 256.0) == -1    # 16 * 256 > 2048
    for cfg in autotune.candidates():
        assert lib.fhp_pass_set_property(h, b"traceBlocksPerCU", 1.0) == 0
        assert lib.fhp_pass_set_property(h, b"traceBlockSize", float(cfg["traceBlockSize"])) == 0
        assert lib.fhp_pass_set_property(h, b"traceStackDepth", float(cfg["traceStackDepth"])) == 0
        assert lib.fhp_pass_set_property(h, b"traceBlocksPerCU", float(cfg["traceBlocksPerCU"])) == 0
    lib.fhp_pass_destroy(h)
    keys = (C.c_char_p * 2)(b"traceBlockSize", b"traceBlocksPerCU")
    vals = (C.c_double * 2)(64.0, 40.0)
    assert lib.fhp_pass_create(b"WARDiffPathTracer", 2, keys, vals, C.byref(h)) == -1
    assert "traceBlocksPerCU" in lib.fhp_last_error().decode()


def test_scene_mesh_offsets_are_validated():
    """Mesh vertex ranges must partition the vertices (checked before any device work)."""
    from falcor_amd import _lib
    lib = _lib.load()
    pos = np.zeros((3, 3), np.float32)
    nrm = np.tile(np.array([0, 0, 1], np.float32), (3, 1))
    uv = np.zeros((3, 2), np.float32)
    idx = np.array([0, 1, 2], np.uint32)
    tm = np.zeros(1, np.uint32)
    mats = np.zeros((1, 16), np.float32)
    mats[0, 12] = -1
    cam = np.zeros(14, np.float32)
    d = _lib.SceneDesc()
    d.num_vertices, d.num_triangles, d.num_materials = 3, 1, 1
    d.h_positions, d.h_normals, d.h_texcoords = pos.ctypes.data, nrm.ctypes.data, uv.ctypes.data
    d.h_indices, d.h_tri_material, d.h_materials, d.h_camera = idx.ctypes.data, tm.ctypes.data, mats.ctypes.data, cam.ctypes.data
    h = C.c_void_p()
    for off in ([0, 2], [1, 3], [0, 2, 1, 3]):
        o = np.array(off, np.uint32)
        d.num_meshes = len(off) - 1
        d.h_mesh_vertex_offsets = o.ctypes.data
        assert lib.fhp_scene_create(C.byref(d), C.byref(h)) == -1
        assert "mesh vertex offsets" in lib.fhp_last_error().decode()
    assert lib.fhp_scene_update_vertices(None, 0, pos.ctypes.data, None) == -1
    assert lib.fhp_grads_set_mesh(None, 0) == -1


def test_product_never_imports_the_oracle():
    for p in (ROOT / "falcor_amd").rglob("*.py"):
        src = p.read_text()
        assert not re.search(r"^\s*(import|from)\s+oracle\b", src, re.M), p


def _acc_pass(lib, props):
    keys = (C.c_char_p * max(len(props), 1))(*[k.encode() for k in props])
    vals = (C.c_double * max(len(props), 1))(*[float(v) for v in props.values()])
    h = C.c_void_p()
    rc = lib.fhp_pass_create(b"AccumulatePass", len(props), keys, vals, C.byref(h))
    return rc, h


def test_accumulate_pass_properties():
    """Defaults of AccumulatePass.h:135-145; enum ranges checked; enableAccumulation is the
    deprecated alias of enabled (AccumulatePass.cpp:87-92)."""
    from falcor_amd import _lib
    lib = _lib.load()
    rc, h = _acc_pass(lib, {})
    assert rc == 0
    v = C.c_double()
    want = {b"enabled": 1, b"autoReset": 1, b"precisionMode": 1, b"maxFrameCount": 0, b"overflowMode": 0}
    for k, x in want.items():
        assert lib.fhp_pass_get_property(h, k, C.byref(v)) == 0 and v.value == x, k
    n = C.c_uint32(7)
    assert lib.fhp_accumulate_frame_count(h, C.byref(n)) == 0 and n.value == 0
    assert lib.fhp_pass_set_property(h, b"precisionMode", 3.0) == -1
    assert "precisionMode" in lib.fhp_last_error().decode()
    assert lib.fhp_pass_set_property(h, b"overflowMode", -1.0) == -1
    assert lib.fhp_pass_set_property(h, b"maxFrameCount", 2.5) == -1
    assert lib.fhp_pass_set_property(h, b"precisionMode", 2.0) == 0
    # An AccumulatePass does not render and a path tracer does not accumulate.
    assert lib.fhp_pass_execute(h, None, 4, 4, 1, 0, C.c_void_p(16), None, None) == -1
    assert "fhp_accumulate" in lib.fhp_last_error().decode()
    lib.fhp_pass_destroy(h)

    rc, h = _acc_pass(lib, {"enableAccumulation": 0})
    assert rc == 0 and "deprecated" in lib.fhp_last_warning().decode()
    assert lib.fhp_pass_get_property(h, b"enabled", C.byref(v)) == 0 and v.value == 0
    lib.fhp_pass_destroy(h)
    rc, h = _acc_pass(lib, {"enableAccumulation": 0, "enabled": 1})
    assert lib.fhp_pass_get_property(h, b"enabled", C.byref(v)) == 0 and v.value == 1
    lib.fhp_pass_destroy(h)
    rc, h = _acc_pass(lib, {"overflowMode": 7})
    assert rc == -1 and "overflowMode" in lib.fhp_last_error().decode()

    keys = (C.c_char_p * 1)(b"maxBounces")
    vals = (C.c_double * 1)(3.0)
    assert lib.fhp_pass_create(b"MinimalPathTracer", 1, keys, vals, C.byref(h)) == 0
    assert lib.fhp_accumulate(h, C.c_void_p(16), 0, 4, 4, C.c_void_p(16), 0, None) == -1
    assert "AccumulatePass" in lib.fhp_last_error().decode()
    lib.fhp_pass_destroy(h)


def test_image_compare_argument_errors():
    from falcor_amd import _lib
    lib = _lib.load()
    e = C.c_double()
    p = C.c_void_p(16)
    assert lib.fhp_image_compare(4, p, p, 4, 4, 0, None, None, C.byref(e), None) == -1
    assert "metric" in lib.fhp_last_error().decode()
    assert lib.fhp_image_compare(0, p, p, 0, 4, 0, None, None, C.byref(e), None) == -1
    assert lib.fhp_image_compare(0, p, p, 4, 4, 0, None, p, C.byref(e), None) == -1
    assert "heat map" in lib.fhp_last_error().decode()
