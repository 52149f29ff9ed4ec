"""Env map importance sampling (EnvMapSampler) and env map next-event estimation in
MinimalPathTracer on the GPU, against the oracle.

* The importance map built on the GPU equals the oracle's within 1e-4 relative per texel (device
  transcendental functions differ from the host's in the last bits).
* EnvMapSampler::sample on the same uniforms: at least 99.9 % of the directions agree within 1e-4
  and their pdfs within 1e-4 relative (a texel choice can flip where two sums differ in the last
  bit); evalPdf likewise.
* Images with useEnvNEE (alone, and with emissive NEE and analytic lights chosen per vertex) match
  the oracle's path for path with the image tolerance of the other forward tests: at most 0.5 %
  of the pixels off by more than 1e-4 (1 + |ref|), relative L1 difference at most 1e-3.
* Batched views with env NEE equal their single-view renders bit for bit; useEnvNEE without
  computeDirect is refused.
* A zero-luminance map has no importance map: env NEE is off (same image as without it) and the
  sampler entry points refuse it; removing the map removes the importance map.
"""
import math

import numpy as np
import pytest

from falcor_amd import scene as S

pytestmark = pytest.mark.gpu


def _sky(h=64, w=128, seed=0):
    rng = np.random.default_rng(seed)
    v = (np.arange(h) + 0.5) / h
    u = (np.arange(w) + 0.5) / w
    sky = np.zeros((h, w, 3), np.float32)
    sky[...] = (0.2 + 0.8 * (1 - v))[:, None, None] * np.float32([0.5, 0.7, 1.0])
    du, dv = (u[None, :] - 0.3), (v[:, None] - 0.25)
    sky += (40.0 * np.exp(-(du ** 2 + dv ** 2) / (2 * 0.03 ** 2)))[..., None] * np.float32([1.0, 0.9, 0.7])
    sky *= rng.uniform(0.8, 1.2, (h, w, 1)).astype(np.float32)
    return sky


def _compare(img, ref, name):
    diff = np.abs(img - ref)
    bad = (diff > 1e-4 * (1 + np.abs(ref))).any(-1).mean()
    l1 = diff.sum() / max(np.abs(ref).sum(), 1e-30)
    print(f"{name}: mean {ref.mean():.5f}, mismatched pixels {bad:.5f}, rel L1 {l1:.2e}", flush=True)
    assert np.isfinite(img).all(), name
    assert bad <= 0.005, name
    assert l1 <= 1e-3, name


def _spheres_under_sky(seed, lights=(), rotation=(0, 30, 0)):
    data = S.sphere_field(seed, num_spheres=60, num_emissive=200, box=12.0, aspect=1.0)
    data.env_map = {"texels": _sky(seed=seed), "intensity": 0.7, "tint": (1.0, 0.9, 0.8), "rotation": rotation}
    data.analytic_lights = list(lights)
    return data


def test_importance_map_and_sampler_match_oracle(gpu, oracle_mod):
    import torch
    from falcor_amd.render_graph import Scene
    data = _spheres_under_sky(41, rotation=(15, 60, -20))
    sc = Scene(data)
    ref = oracle_mod.OracleScene(data)
    lv, lr = sc.get_env_importance(), ref.env_importance()
    assert len(lv) == len(lr) == 10 and lv[-1].shape == (512, 512)
    for a, b in zip(lv, lr):
        np.testing.assert_allclose(a, b, rtol=1e-4, atol=1e-7)
    u = np.random.default_rng(5).random((200000, 2), np.float32)
    g = sc.env_sample(torch.from_numpy(u).cuda()).cpu().numpy()
    o = ref.env_sample(u)
    assert (g[:, 7] == 1).all()
    close = (np.abs(g[:, :3] - o[:, :3]).max(1) <= 1e-4) & (np.abs(g[:, 3] - o[:, 3]) <= 1e-4 * o[:, 3])
    print("env samples agreeing with the oracle:", close.mean(), flush=True)
    assert close.mean() >= 0.999
    np.testing.assert_allclose(g[close, 4:7], o[close, 4:7], rtol=1e-3, atol=1e-6)
    d = o[:, :3]
    pg = sc.env_eval_pdf(torch.from_numpy(np.ascontiguousarray(d)).cuda()).cpu().numpy()
    po = ref.env_pdf(d)
    assert (np.abs(pg - po) <= 1e-4 * po).mean() >= 0.999


@pytest.mark.parametrize("importance", [True, False])
def test_env_nee_images_match_oracle(gpu, oracle_mod, importance):
    from falcor_amd.render_graph import MinimalPathTracer, Scene
    data = _spheres_under_sky(42)
    p = MinimalPathTracer({"maxBounces": 3, "useEnvNEE": True, "useImportanceSampling": importance})
    p.set_scene(Scene(data))
    img = p.render(96, 96, 2, frame=3).cpu().numpy()[..., :3]
    ref, _ = oracle_mod.OracleScene(data).render_mpt(96, 96, spp=2, frame=3, max_bounces=3, importance=importance, env_nee=True)
    _compare(img, ref, f"env NEE importance={importance}")


def test_all_light_types_match_oracle(gpu, oracle_mod):
    from falcor_amd.render_graph import MinimalPathTracer, Scene
    lights = [S.directional_light((0.6, 0.6, 0.5), (0.2, -1.0, 0.3)),
              S.point_light((200.0, 150.0, 100.0), (6.0, 14.0, 6.0), (0, -1, 0), 0.8, 0.3)]
    data = _spheres_under_sky(43, lights)
    for sampler in (0, 1):
        p = MinimalPathTracer({"maxBounces": 4, "useEnvNEE": True, "useEmissiveNEE": True,
                               "emissiveSampler": ("Uniform", "LightBVH")[sampler]})
        p.set_scene(Scene(data))
        img = p.render(96, 96, 2, frame=1).cpu().numpy()[..., :3]
        ref, _ = oracle_mod.OracleScene(data).render_mpt(96, 96, spp=2, frame=1, max_bounces=4, nee=True, env_nee=True,
                                                        light_sampler=sampler)
        _compare(img, ref, f"env + emissive + analytic, sampler {sampler}")


def test_batch_with_env_nee_and_refusal(gpu):
    import torch
    from falcor_amd import _lib
    from falcor_amd.render_graph import MinimalPathTracer, Scene
    a = Scene(_spheres_under_sky(44))
    b = Scene(S.closed_box(2))  # no env map: env NEE is off for its views
    views = [dict(scene=a, width=40, height=32, spp=2, frame=1), dict(scene=b, width=32, height=32, spp=1, frame=2),
             dict(scene=a, width=17, height=23, spp=3, frame=4)]
    props = {"maxBounces": 3, "useEnvNEE": True}
    outs = MinimalPathTracer(props).render_batch(views)
    torch.cuda.synchronize()
    for v, o in zip(views, outs):
        q = MinimalPathTracer(props)
        q.set_scene(v["scene"])
        np.testing.assert_array_equal(o.cpu().numpy(), q.render(v["width"], v["height"], v["spp"], frame=v["frame"]).cpu().numpy())
    p = MinimalPathTracer({"useEnvNEE": True, "computeDirect": False})
    p.set_scene(a)
    with pytest.raises(_lib.FhpError, match="computeDirect"):
        p.render(8, 8, 1, frame=0)


def test_unsampleable_and_removed_env_maps(gpu):
    # A map whose luminance integrates to zero has no importance map: env NEE is off for the scene
    # (the image equals the one without useEnvNEE) and the sampler entry points refuse it; removing
    # the map removes the importance map.
    import torch
    from falcor_amd import _lib
    from falcor_amd.render_graph import MinimalPathTracer, Scene
    data = S.sphere_field(45, num_spheres=30, num_emissive=60, box=8.0, aspect=1.0)
    data.env_map = {"texels": np.zeros((8, 16, 3), np.float32)}
    sc = Scene(data)
    assert sc.get_env_importance() == []
    with pytest.raises(_lib.FhpError, match="importance sampled"):
        sc.env_sample(torch.zeros((4, 2), device="cuda"))
    imgs = []
    for env_nee in (False, True):
        p = MinimalPathTracer({"maxBounces": 2, "useEnvNEE": env_nee})
        p.set_scene(sc)
        imgs.append(p.render(32, 32, 1, frame=0).cpu().numpy())
    np.testing.assert_array_equal(imgs[0], imgs[1])
    sc.set_env_map(_sky())
    assert len(sc.get_env_importance()) == 10
    sc.set_env_map(None)
    assert sc.get_env_importance() == []
