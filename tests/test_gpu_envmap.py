"""Environment map in MinimalPathTracer (EnvMap.eval on primary misses and on scatter misses)
against oracle/envmap.py and a furnace closed form.

* Background pixels equal the oracle's float64 lookup of the exact primary directions (taken from
  VBufferRT's viewW) within 2e-5 relative (float atan2 / acos and bilinear weights).
* A convex Lambertian object (a cube) under a constant environment L reflects exactly albedo * L:
  every cosine-sampled scatter ray escapes and carries weight albedo. Tolerance 1e-5 relative.
"""
import numpy as np
import pytest

from falcor_amd import scene as S
from oracle import envmap as O

pytestmark = pytest.mark.gpu

W, H = 64, 48


def _cube_scene(albedo=(0.6, 0.6, 0.6)):
    sb = S.SceneBuilder()
    m = sb.add_material(S.pbrt_diffuse_material(albedo))
    sb.add_mesh_instance(S.TriangleMesh.create_cube((0.6, 0.6, 0.6)), m, S.Transform.deg((1, 1, 1), (0, 0, 0), (25, 40, 0)), "Cube")
    sb.camera = S.Camera((0.3, 0.4, 3.0), (0, 0, 0), (0, 1, 0), np.radians(45.0), W / H)
    return sb.build()


def _primary(sc):
    from falcor_amd.render_graph import VBufferRT
    vb = VBufferRT({})
    vb.set_scene(sc)
    out = vb.execute_frame(W, H)
    return -out["viewW"].cpu().numpy()[..., :3].astype(np.float64), out["mask"].cpu().numpy() > 0


@pytest.mark.parametrize("rotation", [(0, 0, 0), (10, 30, -20)])
def test_background_matches_oracle(gpu, rotation):
    from falcor_amd.render_graph import MinimalPathTracer, Scene
    sc = Scene(_cube_scene())
    tex = np.random.default_rng(1).uniform(0.1, 4.0, (16, 32, 3)).astype(np.float32)
    sc.set_env_map(tex, intensity=2.0, tint=(1.0, 0.5, 0.25), rotation=rotation)
    d, hit = _primary(sc)
    assert 0.03 < hit.mean() < 0.9
    p = MinimalPathTracer({"maxBounces": 0})
    p.set_scene(sc)
    img = p.render(W, H, 1, frame=0).cpu().numpy()[..., :3]
    want = O.env_eval(tex, d, 2.0, (1.0, 0.5, 0.25), rotation)
    np.testing.assert_allclose(img[~hit], want[~hit], rtol=2e-5, atol=1e-6)


def test_furnace_cube_in_constant_environment(gpu):
    from falcor_amd.render_graph import MinimalPathTracer, Scene
    albedo = np.array([0.6, 0.4, 0.2])
    L = np.array([1.5, 1.0, 0.5])
    sc = Scene(_cube_scene(tuple(albedo)))
    sc.set_env_map(np.broadcast_to(L, (4, 8, 3)).astype(np.float32))
    _, hit = _primary(sc)
    for bounces in (1, 3):
        p = MinimalPathTracer({"maxBounces": bounces})
        p.set_scene(sc)
        img = p.render(W, H, 4, frame=3).cpu().numpy()[..., :3]
        np.testing.assert_allclose(img[hit], np.broadcast_to(albedo * L, img[hit].shape), rtol=1e-5)
        np.testing.assert_allclose(img[~hit], np.broadcast_to(L, img[~hit].shape), rtol=1e-6)
    # Light only (no background); computeDirect off removes the first scatter's environment term.
    sc.set_env_map(np.broadcast_to(L, (4, 8, 3)).astype(np.float32), use_background=False)
    p = MinimalPathTracer({"maxBounces": 1})
    p.set_scene(sc)
    img = p.render(W, H, 1, frame=0).cpu().numpy()[..., :3]
    assert np.all(img[~hit] == 0)
    np.testing.assert_allclose(img[hit], np.broadcast_to(albedo * L, img[hit].shape), rtol=1e-5)
    p.set_property("computeDirect", 0)
    img = p.render(W, H, 1, frame=0).cpu().numpy()[..., :3]
    assert np.all(img == 0)
    sc.set_env_map(None)
    p.set_property("computeDirect", 1)
    assert np.all(p.render(W, H, 1, frame=0).cpu().numpy()[..., :3] == 0)


def test_batched_view_with_environment_matches_single(gpu):
    import torch
    from falcor_amd.render_graph import MinimalPathTracer, Scene
    sc = Scene(_cube_scene())
    sc.set_env_map(np.random.default_rng(2).uniform(0, 2, (8, 16, 3)).astype(np.float32), rotation=(0, 45, 0))
    other = Scene(S.closed_box(0))
    p = MinimalPathTracer({"maxBounces": 3})
    p.set_scene(sc)
    single = p.render(W, H, 2, frame=4).clone()
    outs = p.render_batch([{"scene": other, "width": 32, "height": 32, "spp": 1},
                           {"scene": sc, "width": W, "height": H, "spp": 2, "frame": 4}])
    assert torch.equal(outs[1], single)
