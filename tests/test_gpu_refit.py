"""GPU BVH refit after set_mesh_vertices (update_mode "refit", the reference's default BLAS update
mode): the refitted tree must return exactly the hits of a freshly built tree over the same
geometry (closest-hit triangle, t and barycentrics are properties of the ray and the triangles,
not of the tree) and of the traversal oracle; images rendered from the same seeds are identical."""
import numpy as np
import pytest

from falcor_amd import scene as S

pytestmark = pytest.mark.gpu


def _rays(n, seed=0):
    rng = np.random.default_rng(seed)
    o = rng.uniform([-0.25, 0.02, -0.25], [0.25, 0.53, 0.25], (n, 3))
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3], r[:, 3], r[:, 4:7], r[:, 7] = o, 0.0, d, 1e30
    return r


def _hits(sc, rays):
    import torch
    out = sc.trace_closest(torch.from_numpy(rays).cuda())
    torch.cuda.synchronize()
    return out.cpu().numpy().view(np.uint32)


def _deform(pos, k, rng):
    c = pos.mean(0)
    return (c + (pos - c) * (1 + 0.1 * k) + np.array([0.01 * k, -0.005 * k, 0.02 * k])
            + rng.normal(0, 0.002, pos.shape)).astype(np.float32)


def test_refit_matches_rebuild_and_oracle(gpu, oracle_mod):
    import torch
    from dataclasses import replace
    from falcor_amd.render_graph import MinimalPathTracer, Scene
    data = S.displaced_sphere_in_box(1, radius=0.1, segments=(40, 20), sigma=0.05)
    mesh = data.mesh_names.index("Sphere")
    a, b = Scene(data), Scene(data)
    pos0, nrm0, _ = a.get_mesh_vertices_and_indices(mesh)
    rng = np.random.default_rng(3)
    rays = _rays(20000)
    for k in range(1, 4):
        pos = _deform(pos0, k, rng)
        a.set_mesh_vertices(mesh, torch.from_numpy(pos).cuda(), torch.from_numpy(nrm0).cuda(), update_mode="refit")
        b.set_mesh_vertices(mesh, pos, nrm0)  # host arrays: full rebuild
        ha, hb = _hits(a, rays), _hits(b, rays)
        assert (ha[:, 0] != 0xFFFFFFFF).mean() > 0.5  # the box is open at the front
        assert np.array_equal(ha, hb), k
    v0, v1 = a.mesh_vertex_range(mesh)
    p_all = np.array(data.positions, copy=True)
    p_all[v0:v1] = pos
    exp = oracle_mod.OracleScene(replace(data, positions=p_all)).trace(rays)
    assert np.array_equal(ha, exp)
    # Rendering through the refitted tree.
    pa, pb = MinimalPathTracer({"maxBounces": 3}), MinimalPathTracer({"maxBounces": 3})
    pa.set_scene(a)
    pb.set_scene(b)
    assert torch.equal(pa.render(64, 64, 4, frame=2), pb.render(64, 64, 4, frame=2))


def test_refit_of_emissive_mesh_rebuilds_lights(gpu):
    import torch
    from falcor_amd.render_graph import MinimalPathTracer, Scene
    data = S.closed_box(0)
    mesh = data.mesh_names.index("Light")
    a, b = Scene(data), Scene(data)
    pos, nrm, _ = a.get_mesh_vertices_and_indices(mesh)
    moved = (pos + np.array([0.05, -0.01, 0.02], np.float32)).astype(np.float32)
    a.set_mesh_vertices(mesh, torch.from_numpy(moved).cuda(), update_mode="refit")
    b.set_mesh_vertices(mesh, moved)
    props = {"maxBounces": 2, "useEmissiveNEE": 1}
    pa, pb = MinimalPathTracer(props), MinimalPathTracer(props)
    pa.set_scene(a)
    pb.set_scene(b)
    assert torch.equal(pa.render(48, 48, 4, frame=1), pb.render(48, 48, 4, frame=1))


def test_update_vertices_device_errors(gpu):
    import torch
    from falcor_amd import _lib
    from falcor_amd.render_graph import Scene
    data = S.closed_box(0)
    sc = Scene(data)
    pos, _, _ = sc.get_mesh_vertices_and_indices(0)
    with pytest.raises(ValueError, match="vertices"):
        sc.set_mesh_vertices(0, torch.zeros((3, 3), device="cuda"))
    bad = torch.from_numpy(pos).cuda()
    bad[0, 0] = float("nan")
    with pytest.raises(_lib.FhpError, match="non-finite"):
        sc.set_mesh_vertices(0, bad)
    with pytest.raises(ValueError, match="update_mode"):
        sc.set_mesh_vertices(0, torch.from_numpy(pos).cuda(), update_mode="fast")
