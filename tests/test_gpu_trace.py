"""GPU parity of the BVH4 ray queries against the oracle: nearest-hit triangle ids, t and
barycentrics are bit-exact for identical ray sets (closed box: brute force over all triangles;
sphere fields: oracle BVH, itself pinned to brute force on CPU)."""
import numpy as np
import pytest

from falcor_amd import scene as S

pytestmark = pytest.mark.gpu


def _gpu_trace(scene, rays, any_hit=False):
    import torch
    r = torch.from_numpy(np.ascontiguousarray(rays, np.float32)).cuda()
    out = scene.trace_any(r) if any_hit else scene.trace_closest(r)
    torch.cuda.synchronize()
    return out.cpu().numpy().view(np.uint32)


def _record_rays(osc, W, H, nee, max_bounces=3):
    rec = osc.record_mpt(W, H, 0, W * H, sample_number=0, max_bounces=max_bounces, nee=nee, brute=osc.num_triangles < 100)
    return rec


def test_closed_box_all_rays_bit_exact(gpu, oracle_mod):
    from falcor_amd.render_graph import Scene
    data = S.closed_box(0)
    sc, osc = Scene(data), oracle_mod.OracleScene(data)
    rec = _record_rays(osc, 256, 256, nee=True)
    closest = rec[rec[:, 12].view(np.uint32) == 0]
    print("closed box rays:", closest.shape[0], flush=True)
    got = _gpu_trace(sc, closest[:, :8])
    exp = np.ascontiguousarray(closest[:, 8:12]).view(np.uint32)
    # The oracle recorded these hits by brute force over all 36 triangles.
    assert np.array_equal(got[:, 0], exp[:, 0])
    assert np.array_equal(got, exp)
    vis = rec[rec[:, 12].view(np.uint32) == 1]
    occ = _gpu_trace(sc, vis[:, :8], any_hit=True)
    assert np.array_equal(occ != 0, vis[:, 8].view(np.uint32) != 0xFFFFFFFF)


def test_sphere_field_random_rays_bit_exact(gpu, oracle_mod):
    from falcor_amd.render_graph import Scene
    data = S.sphere_field(11, num_spheres=256, num_emissive=512, box=40.0)
    sc, osc = Scene(data), oracle_mod.OracleScene(data)
    info = sc.info()
    print("bvh", info["bvh_nodes"], "depth", info["bvh_depth"], "stack", info["stack_need"], flush=True)
    rng = np.random.default_rng(0)
    n = 200000
    o = rng.uniform(-5, 45, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d[: n // 10, 1] = 0.0  # axis-parallel components exercise the slab test's zero-direction path
    rays = np.concatenate([o, np.zeros((n, 1), np.float32), d, np.full((n, 1), 3.0e38, np.float32)], 1)
    got = _gpu_trace(sc, rays)
    exp = osc.trace(rays)
    assert (exp[:, 0] != 0xFFFFFFFF).mean() > 0.2
    assert np.array_equal(got, exp)
    # Finite tmax / tmin windows and the any-hit query.
    rays[:, 3] = rng.uniform(0, 5, n)
    rays[:, 7] = rng.uniform(5, 30, n)
    got = _gpu_trace(sc, rays)
    exp = osc.trace(rays)
    assert np.array_equal(got, exp)
    occ = _gpu_trace(sc, rays, any_hit=True)
    exp_any = osc.trace(rays, any_hit=True)
    assert np.array_equal(occ != 0, exp_any[:, 0] != 0xFFFFFFFF)


def test_scatter_rays_of_large_scene_bit_exact(gpu, oracle_mod):
    """Secondary rays recorded by the oracle integrator (NEE on, LightBVH) on a 1024-sphere field."""
    from falcor_amd.render_graph import Scene
    data = S.sphere_field(1, num_spheres=1024, num_emissive=4096, box=100.0)
    sc, osc = Scene(data), oracle_mod.OracleScene(data)
    W, H = 192, 108
    rec = osc.record_mpt(W, H, 0, W * H, sample_number=3, max_bounces=6, nee=True)
    print("recorded rays:", rec.shape[0], flush=True)
    kind = rec[:, 12].view(np.uint32)
    c = rec[kind == 0]
    got = _gpu_trace(sc, c[:, :8])
    assert np.array_equal(got, np.ascontiguousarray(c[:, 8:12]).view(np.uint32))
    v = rec[kind == 1]
    occ = _gpu_trace(sc, v[:, :8], any_hit=True)
    assert np.array_equal(occ != 0, v[:, 8].view(np.uint32) != 0xFFFFFFFF)


def test_empty_ray_batch_and_errors(gpu):
    import torch
    from falcor_amd import _lib
    from falcor_amd.render_graph import Scene
    sc = Scene(S.closed_box(0))
    out = sc.trace_closest(torch.zeros((0, 8), device="cuda"))
    assert out.shape == (0, 4)
    with pytest.raises(_lib.FhpError):
        _lib.check(_lib.load().fhp_trace_closest(sc.handle, None, 4, None, None))
