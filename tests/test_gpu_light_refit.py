"""Node transforms on the GPU (Scene::updateNodeTransform and the animation update it starts): the
BVH refit, LightCollection::updateTrianglePositions and LightBVH::refit, against the oracle's
restatement of the same steps on the same vertices.

* Light records: positions bit-exact (both read the same transformed vertices), face normals
  and areas within 2e-6, flux unchanged.
* Light BVH: topology, flux and leaf lists unchanged; boxes within 1e-6 of the scene extent;
  cones within 1e-4 where both have one, and the same nodes without a cone (at most 1 % may
  differ where the sine test sits on its threshold).
* Images after the transform (MinimalPathTracer with LightBVH NEE, WARDiffPathTracer primal)
  match the oracle with the forward tests' tolerance: at most 0.5 % of the pixels off by more
  than 1e-4 (1 + |ref|) and a relative L1 difference at most 1e-3.
* Transform semantics: not cumulative, identity returns the rest pose, set_mesh_vertices resets
  the rest pose; mirroring transforms and bad mesh ids are refused.
"""
import math

import numpy as np
import pytest

from falcor_amd import scene as S

pytestmark = pytest.mark.gpu

FRAC_TOL = 0.005
L1_TOL = 1e-3


def _compare(img, ref, name):
    diff = np.abs(img - ref)
    bad = (diff > 1e-4 * (1 + np.abs(ref))).any(-1).mean()
    l1 = diff.sum() / max(np.abs(ref).sum(), 1e-30)
    print(f"{name}: mean {ref.mean():.5f}, mismatched pixels {bad:.5f}, rel L1 {l1:.2e}", flush=True)
    assert np.isfinite(img).all(), name
    assert bad <= FRAC_TOL, name
    assert l1 <= L1_TOL, name


def _scene():
    return S.sphere_field(7, num_spheres=64, num_emissive=600, box=24.0, aspect=1.0)


def _emissive_mesh(data, lights):
    light_mesh = data.tri_mesh[lights[:, 11].view(np.uint32)]
    return int(np.bincount(light_mesh, minlength=len(data.mesh_names)).argmax())


def _move(sc, data, k, angle_deg=35.0, shift=(0.0, 4.0, -2.0)):
    """New local-to-world transform of mesh k: its build transform, then a rotation about its
    centre and a shift."""
    T0 = data.mesh_transforms[k]
    c = T0[:3, 3]
    a = math.radians(angle_deg)
    R = np.array([[math.cos(a), 0, math.sin(a)], [0, 1, 0], [-math.sin(a), 0, math.cos(a)]])
    A = np.eye(4)
    A[:3, :3] = R
    A[:3, 3] = c - R @ c + np.asarray(shift)
    sc.update_node_transform(k, A @ T0)


def _product_vs_oracle_lights(sc, osc):
    P = sc.lights()
    O = osc.lights()
    assert P.shape[0] == O.shape[0]
    pos_p = P[:, [0, 1, 2, 4, 5, 6, 8, 9, 10]]
    np.testing.assert_array_equal(pos_p, O[:, :9])
    np.testing.assert_allclose(P[:, 12:15], O[:, 9:12], atol=2e-6)
    np.testing.assert_allclose(P[:, 3], O[:, 12], rtol=2e-6)
    np.testing.assert_array_equal(P[:, 7], O[:, 16])


def _product_vs_oracle_lbvh(sc, osc):
    pn, ptri, _ = sc.light_bvh()
    on, otri, _ = osc.light_bvh()
    assert pn.shape[0] == on.shape[0]
    np.testing.assert_array_equal(ptri, otri)
    leaf_p = (pn[:, 12].view(np.uint32) & 0x80000000) != 0
    np.testing.assert_array_equal(leaf_p, on[:, 11] == 1)
    np.testing.assert_array_equal(pn[:, 3], on[:, 6])  # flux
    scale = float(np.abs(on[:, 0:3]).max())
    np.testing.assert_allclose(pn[:, 0:3], on[:, 0:3], atol=1e-6 * scale)
    np.testing.assert_allclose(pn[:, 4:7], on[:, 3:6], atol=1e-6 * scale)
    vp, vo = pn[:, 7] != -1, on[:, 10] != -1
    assert (vp != vo).mean() <= 0.01
    both = vp & vo
    assert both.mean() > 0.3
    np.testing.assert_allclose(pn[both, 7], on[both, 10], atol=1e-4)
    np.testing.assert_allclose(pn[both, 8:11], on[both, 7:10], atol=1e-4)


def test_node_transform_updates_lights_like_the_oracle(gpu, oracle_mod):
    from falcor_amd.render_graph import Scene
    data = _scene()
    sc = Scene(data)
    osc = oracle_mod.OracleScene(data)
    _product_vs_oracle_lbvh(sc, osc)  # the same tree before the motion
    lights0 = sc.lights()
    k = _emissive_mesh(data, lights0)
    _move(sc, data, k)
    _move(sc, data, 5, angle_deg=-20.0, shift=(1.0, 0.0, 0.0))  # a second mesh
    osc.animate(sc._arrays["pos"], sc._arrays["nrm"])
    _product_vs_oracle_lights(sc, osc)
    _product_vs_oracle_lbvh(sc, osc)
    moved = np.any(sc.lights()[:, :12] != lights0[:, :12], axis=1)
    assert 0 < moved.sum() < moved.size


def test_images_after_node_transform_match_oracle(gpu, oracle_mod):
    from falcor_amd.render_graph import MinimalPathTracer, Scene, WARDiffPathTracer
    data = _scene()
    sc = Scene(data)
    k = _emissive_mesh(data, sc.lights())
    _move(sc, data, k, angle_deg=60.0, shift=(-3.0, 2.0, 0.0))
    osc = oracle_mod.OracleScene(data)
    osc.animate(sc._arrays["pos"], sc._arrays["nrm"])
    p = MinimalPathTracer({"maxBounces": 3, "useEmissiveNEE": True, "emissiveSampler": "LightBVH"})
    p.set_scene(sc)
    img = p.render(96, 96, spp=2, frame=3).cpu().numpy()[..., :3]
    ref, _ = osc.render_mpt(96, 96, spp=2, frame=3, max_bounces=3, nee=True, light_sampler=1)
    _compare(img, ref, "MPT LightBVH NEE after node transform")
    w = WARDiffPathTracer({"maxBounces": 2, "useWAR": False})
    w.set_scene(sc)
    img = w.render(96, 96, spp=2, frame=0).cpu().numpy()[..., :3]
    ref, _ = osc.render_war(96, 96, spp=2, frame=0, max_bounces=2)
    _compare(img, ref, "WAR primal after node transform")


def test_node_transform_semantics(gpu):
    from falcor_amd import _lib
    from falcor_amd.render_graph import Scene
    data = _scene()
    sc = Scene(data)
    k = _emissive_mesh(data, sc.lights())
    v0, v1 = sc.mesh_vertex_range(k)
    rest = data.positions[v0:v1].copy()
    lights0, nodes0 = sc.lights(), sc.light_bvh()[0]
    _move(sc, data, k)
    once = sc.get_mesh_vertices_and_indices(k)[0]
    assert np.abs(once - rest).max() > 1.0
    _move(sc, data, k)  # the same node transform again: not cumulative
    np.testing.assert_array_equal(sc.get_mesh_vertices_and_indices(k)[0], once)
    sc.update_node_transform(k, data.mesh_transforms[k])  # back to the build transform
    np.testing.assert_allclose(sc.get_mesh_vertices_and_indices(k)[0], rest, atol=1e-5)
    np.testing.assert_allclose(sc.lights()[:, :12], lights0[:, :12], atol=1e-5)
    np.testing.assert_array_equal(sc.light_bvh()[0][:, [3, 12, 13, 14]], nodes0[:, [3, 12, 13, 14]])
    # After set_mesh_vertices the given vertices are the mesh's local vertices (identity pose).
    sc.set_mesh_vertices(k, rest + np.float32([0, 1, 0]))
    sc.update_node_transform(k, np.eye(4))
    np.testing.assert_array_equal(sc.get_mesh_vertices_and_indices(k)[0], rest + np.float32([0, 1, 0]))
    shift = np.eye(4)
    shift[:3, 3] = (0, -1, 0)
    sc.update_node_transform(k, shift)
    np.testing.assert_allclose(sc.get_mesh_vertices_and_indices(k)[0], rest, atol=1e-5)
    mirror = np.diag([-1.0, 1.0, 1.0, 1.0])
    with pytest.raises(_lib.FhpError, match="determinant"):
        sc.update_node_transform(k, mirror)
    with pytest.raises(ValueError):
        sc.update_node_transform(len(data.mesh_names), np.eye(4))
    m = np.eye(4, dtype=np.float32)[:3].copy()
    assert _lib.load().fhp_scene_update_node_transform(sc.handle, 10 ** 6, m.ctypes.data, None) == -1
