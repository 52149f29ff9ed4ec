"""Reference-shaped scene calls of the boundary: serialized material parameters
(Scene.get/set_material_params with StandardMaterialParamLayout), the explicit BVH build and the
session (device + stream) handles.

* get_material_params returns each StandardMaterial's values at the layout's offsets
  (StandardMaterialParamLayout.slang:32-41), zeros elsewhere; set_material_params clamps to the
  layout ranges (MaterialParamLayout.slangh:40-45) and rounds the float16-stored fields
  (BasicMaterialData.slang:50-66), as StandardMaterial::deserializeParams does; non-standard
  materials are refused; images after set_material_params equal images after setting the
  resulting records directly.
* fhp_bvh_build with another leaf size gives the same closest hits, bit for bit.
* A session's stream runs the ray queries like the default stream.
* Host-array mutators wait for work queued on other streams: frames queued on a session stream
  before a material and vertex update equal renders of the old scene, the next one a render of
  a fresh scene built with the new materials and vertices.
"""
import ctypes as C

import numpy as np
import pytest

from falcor_amd import scene as S

pytestmark = pytest.mark.gpu


def _scene_with_known_material():
    sb = S.SceneBuilder()
    known = sb.add_material(S.standard_material(base=(0.2, 0.3, 0.4), roughness=0.5, metallic=0.25, ior=1.4,
                                                emissive=(0.5, 0.25, 0.125), emissive_factor=3.0, specular_transmission=0.2,
                                                diffuse_transmission=0.1, transmission=(0.9, 0.8, 0.7)))
    matte = sb.add_material(S.pbrt_diffuse_material((0.5, 0.5, 0.5)))
    shiny = sb.add_material(S.standard_material(base=(0.7, 0.6, 0.5), roughness=0.3, metallic=0.0))
    quad = S.TriangleMesh.create_quad()
    sb.add_mesh_instance(quad, matte, S.Transform((4, 1, 4)), "floor")
    sb.add_mesh_instance(S.TriangleMesh.create_sphere(0.5, 24, 12), known, S.Transform((1, 1, 1), (-0.6, 0.5, 0)), "a")
    sb.add_mesh_instance(S.TriangleMesh.create_sphere(0.5, 24, 12), shiny, S.Transform((1, 1, 1), (0.6, 0.5, 0)), "b")
    sb.add_mesh_instance(quad, sb.add_material(S.standard_material(base=(0, 0, 0), emissive=(8, 8, 8))),
                         S.Transform.deg((2, 1, 2), (0, 2.5, 0), (180, 0, 0)), "light")
    sb.camera = S.Camera((0, 1.2, 3.5), (0, 0.4, 0), (0, 1, 0), np.radians(45.0), 1.0)
    return sb.build()


def test_material_params_layout_clamp_and_float16(gpu):
    import torch
    from falcor_amd import _lib
    from falcor_amd.render_graph import Material, Scene, get_material_param_layout
    sc = Scene(_scene_with_known_material())
    assert Material.PARAM_COUNT == 20
    lay = get_material_param_layout("Standard")
    assert lay["roughness"] == {"offset": 4, "size": 1} and lay["emissive_factor"] == {"offset": 14, "size": 1}
    p = sc.get_material_params([0, 2])
    want = np.zeros(20, np.float32)
    want[0:3], want[3], want[4], want[5] = (0.2, 0.3, 0.4), 0.25, 0.5, 1.4
    want[6:9], want[9], want[10], want[11:14], want[14] = (0.9, 0.8, 0.7), 0.1, 0.2, (0.5, 0.25, 0.125), 3.0
    np.testing.assert_array_equal(p[0], want)
    # Fill a caller's tensor, as the reference fills its parameter buffer.
    t = torch.zeros(40)
    sc.get_material_params(torch.tensor([0, 2], dtype=torch.int32), t)
    np.testing.assert_array_equal(t.numpy().reshape(2, 20), p)
    # Deserialize: clamps and float16 storage.
    q = p.copy()
    q[0, 0:3] = (0.123456, -1.0, 7.0)  # base colour: float16, clamped to [1e-4, 1 - 1e-4]
    q[0, 4] = 0.01                      # roughness >= 0.05
    q[0, 5] = 5000.0                    # ior <= 1e3
    q[0, 10] = -0.5                     # specular transmission >= 0
    q[0, 11:14] = (3.0, 0.5, 0.3)       # emissive colour: float32, clamped below 1
    q[0, 14] = 1e9                      # emissive factor <= 1e8
    q[1, 3] = 0.0                       # metallic >= 1e-4
    sc.set_material_params([0, 2], q)
    r = sc.get_material_params([0, 2])
    h = lambda v: float(np.float16(np.float32(v)))
    np.testing.assert_array_equal(r[0, 0:3], np.float32([h(0.123456), h(1e-4), h(1 - 1e-4)]))
    assert r[0, 4] == h(0.05) and r[0, 5] == h(1e3) and r[0, 10] == 0.0
    np.testing.assert_array_equal(r[0, 11:14], np.float32([1 - 1e-4, 0.5, 0.3]))
    assert r[0, 14] == np.float32(1e8) and r[1, 3] == h(1e-4)
    rec = sc.get_material_records()
    assert rec[0, S.M_TYPE] == S.MAT_STANDARD and rec[0, S.M_DOUBLE_SIDED] == 0 and rec[0, S.M_TEX] == -1
    with pytest.raises(_lib.FhpError, match="serializ"):
        sc.get_material_params([1])  # PBRTDiffuseMaterial
    with pytest.raises(_lib.FhpError, match="out of range"):
        sc.get_material_params([99])


def test_set_material_params_renders_like_records(gpu):
    from falcor_amd.render_graph import MinimalPathTracer, Scene
    data = _scene_with_known_material()
    a, b = Scene(data), Scene(data)
    p = a.get_material_params([2])
    p[0, 0:3] = (0.25, 0.5, 0.75)
    p[0, 4] = 0.625
    a.set_material_params([2], p)
    rec = a.get_material_records()
    # The set values (exact in float16), and the clamped defaults: metallic 0 and the emissive
    # colour 0 become 1e-4 (in float16 and float32 respectively).
    np.testing.assert_array_equal(rec[2, S.M_BASE:S.M_BASE + 3], [0.25, 0.5, 0.75])
    assert rec[2, S.M_ROUGH] == 0.625 and rec[2, S.M_METAL] == np.float16(1e-4)
    np.testing.assert_array_equal(rec[2, S.M_EMIT:S.M_EMIT + 3], np.float32(1e-4))
    b.set_material_records(rec)
    imgs = []
    for sc in (a, b):
        t = MinimalPathTracer({"maxBounces": 3})
        t.set_scene(sc)
        imgs.append(t.render(48, 48, 2, frame=0).cpu().numpy())
    np.testing.assert_array_equal(imgs[0], imgs[1])


def test_bvh_build_and_session_stream(gpu):
    import torch
    from falcor_amd import _lib
    from falcor_amd.render_graph import Scene
    sc = Scene(S.sphere_field(4, num_spheres=48, num_emissive=64, box=12.0))
    rng = np.random.default_rng(2)
    n = 20000
    o = rng.uniform(-2, 14, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = torch.from_numpy(np.concatenate([o, np.zeros((n, 1), np.float32), d, np.full((n, 1), 1e30, np.float32)], 1)).cuda()
    h0 = sc.trace_closest(rays).cpu().numpy()
    nodes0 = sc.info()["bvh_nodes"]
    L = _lib.load()
    _lib.check(L.fhp_bvh_build(sc.handle, 2))
    assert sc.info()["bvh_nodes"] != nodes0
    np.testing.assert_array_equal(sc.trace_closest(rays).cpu().numpy(), h0)
    assert L.fhp_bvh_build(sc.handle, 9) == -1
    _lib.check(L.fhp_bvh_build(sc.handle, 0))
    assert sc.info()["bvh_nodes"] == nodes0
    # A session's stream.
    sess, st = C.c_void_p(), C.c_void_p()
    _lib.check(L.fhp_session_create(0, C.byref(sess)))
    _lib.check(L.fhp_session_stream(sess, C.byref(st)))
    hits = torch.empty((n, 4), dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    _lib.check(L.fhp_trace_closest(sc.handle, rays.data_ptr(), n, hits.data_ptr(), st))
    torch.cuda.synchronize()
    _lib.check(L.fhp_session_destroy(sess))
    np.testing.assert_array_equal(hits.cpu().numpy(), h0)
    assert L.fhp_session_create(1 << 20, C.byref(sess)) == -1


def test_host_mutators_wait_for_session_stream_work(gpu):
    """Host-array scene mutators (materials, vertices) are synchronous with every stream: frames
    queued on a session's non-blocking stream before an update render the old scene, the frame
    queued after it renders the new one, each bit-identical to a fresh scene's render."""
    import dataclasses
    import torch
    from falcor_amd import _lib
    from falcor_amd.render_graph import MinimalPathTracer, Scene
    data = S.displaced_sphere_in_box(1, radius=0.12, segments=(48, 24), sigma=0.05)
    mesh = data.mesh_names.index("Sphere")
    W, H, spp, frames = 192, 192, 16, 4
    props = {"maxBounces": 3, "useEmissiveNEE": True, "emissiveSampler": "LightBVH"}
    L = _lib.load()
    sess, st = C.c_void_p(), C.c_void_p()
    _lib.check(L.fhp_session_create(0, C.byref(sess)))
    _lib.check(L.fhp_session_stream(sess, C.byref(st)))
    try:
        sc = Scene(data)
        p = MinimalPathTracer(props)
        outs = [torch.empty((H, W, 4), device="cuda") for _ in range(frames + 1)]
        torch.cuda.synchronize()
        for f in range(frames):  # queued, not waited for
            _lib.check(L.fhp_pass_execute(p.handle, sc.handle, W, H, spp, f, outs[f].data_ptr(), None, st))
        mats = np.array(data.materials, np.float32)
        mats[:, S.M_BASE:S.M_BASE + 3] = mats[:, S.M_BASE:S.M_BASE + 3][::-1]  # permute the base colours
        _lib.check(L.fhp_scene_set_materials(sc.handle, mats.ctypes.data))
        v0, v1 = sc.mesh_vertex_range(mesh)
        pos = np.array(data.positions[v0:v1] + np.array([0.03, -0.02, 0.01], np.float32), np.float32)
        nrm = np.ascontiguousarray(data.normals[v0:v1], np.float32)
        _lib.check(L.fhp_scene_update_vertices(sc.handle, mesh, pos.ctypes.data, nrm.ctypes.data))
        _lib.check(L.fhp_pass_execute(p.handle, sc.handle, W, H, spp, frames, outs[frames].data_ptr(), None, st))
        # The trace error query synchronises the stream it is given.
        bits = C.c_uint32()
        _lib.check(L.fhp_scene_trace_errors(sc.handle, C.byref(bits), 1, st))
        assert bits.value == 0
        got = [o.cpu().numpy() for o in outs]
    finally:
        _lib.check(L.fhp_session_destroy(sess))
    positions = np.array(data.positions, np.float32)
    positions[v0:v1] = pos
    moved = dataclasses.replace(data, materials=mats, positions=positions)
    ref_old, ref_new = MinimalPathTracer(props), MinimalPathTracer(props)
    ref_old.set_scene(Scene(data))
    ref_new.set_scene(Scene(moved))
    for f in range(frames):
        np.testing.assert_array_equal(got[f], ref_old.render(W, H, spp, frame=f).cpu().numpy())
    np.testing.assert_array_equal(got[frames], ref_new.render(W, H, spp, frame=frames).cpu().numpy())
    assert not np.array_equal(got[frames - 1], got[frames])
