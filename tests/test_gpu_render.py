"""GPU parity of the forward passes (MinimalPathTracer, WARDiffPathTracer primal) and of the
path-replay adjoint against the oracle.

Image tolerance (FP32): the GPU and the oracle run the same operation sequence, but cos/sin/
log/pow come from different math libraries (device ocml vs host libm, both within a few ULP),
so a path can take a different branch near a triangle edge. The test bounds
  * the fraction of pixels differing by more than 1e-4 * (1 + |ref|):  <= 0.5 %
  * the relative L1 difference of the whole image:                     <= 1e-3
"""
import math

import numpy as np
import pytest

from falcor_amd import scene as S

pytestmark = pytest.mark.gpu

FRAC_TOL = 0.005
L1_TOL = 1e-3


def _compare(gpu_img, ref, name):
    diff = np.abs(gpu_img - ref)
    bad = (diff > 1e-4 * (1 + np.abs(ref))).any(-1).mean()
    l1 = diff.sum() / max(np.abs(ref).sum(), 1e-30)
    print(f"{name}: mean {ref.mean():.5f}, mismatched pixels {bad:.5f}, rel L1 {l1:.2e}", flush=True)
    assert bad <= FRAC_TOL, name
    assert l1 <= L1_TOL, name
    return bad, l1


def _mpt(props):
    from falcor_amd.render_graph import MinimalPathTracer
    return MinimalPathTracer(props)


@pytest.mark.parametrize("max_bounces", [0, 3])
def test_mpt_closed_box_matches_oracle(gpu, oracle_mod, max_bounces):
    from falcor_amd.render_graph import Scene
    data = S.closed_box(0)
    sc, osc = Scene(data), oracle_mod.OracleScene(data)
    p = _mpt({"maxBounces": max_bounces})
    p.set_scene(sc)
    img, st = p.render(256, 256, spp=1, frame=0, stats=True)
    ref, rst = osc.render_mpt(256, 256, spp=1, frame=0, max_bounces=max_bounces, brute=True)
    _compare(img.cpu().numpy()[..., :3], ref, f"closed box B={max_bounces}")
    assert np.all(img.cpu().numpy()[..., 3] == 1.0)
    assert abs(int(st[0]) - int(rst[0])) <= 0.001 * int(rst[0])


def test_mpt_multi_spp_frames(gpu, oracle_mod):
    from falcor_amd.render_graph import Scene
    data = S.closed_box(2)
    sc, osc = Scene(data), oracle_mod.OracleScene(data)
    p = _mpt({"maxBounces": 2})
    p.set_scene(sc)
    img = p.render(64, 48, spp=5, frame=3).cpu().numpy()[..., :3]
    ref, _ = osc.render_mpt(64, 48, spp=5, frame=3, max_bounces=2, brute=True)
    _compare(img, ref, "closed box 5 spp")


@pytest.mark.parametrize("sampler", ["Uniform", "LightBVH"])
def test_mpt_emissive_nee_matches_oracle(gpu, oracle_mod, sampler):
    from falcor_amd.render_graph import Scene
    data = S.sphere_field(5, num_spheres=200, num_emissive=800, box=30.0, aspect=1.5)
    sc, osc = Scene(data), oracle_mod.OracleScene(data)
    p = _mpt({"maxBounces": 4, "useEmissiveNEE": True, "emissiveSampler": sampler})
    p.set_scene(sc)
    img, st = p.render(120, 80, spp=2, frame=1, stats=True)
    ref, rst = osc.render_mpt(120, 80, spp=2, frame=1, max_bounces=4, nee=True, light_sampler=0 if sampler == "Uniform" else 1)
    _compare(img.cpu().numpy()[..., :3], ref, f"sphere field NEE {sampler}")
    assert st[1] > 0 and abs(int(st[1]) - int(rst[1])) <= 0.01 * int(rst[1])


def test_mpt_standard_material_scene(gpu, oracle_mod):
    """Standard materials (Frostbite + GGX, delta lobes for roughness < 0.08, metals)."""
    from falcor_amd.render_graph import Scene
    data = S.sphere_field(9, num_spheres=120, num_emissive=400, box=20.0, aspect=1.0)
    sc, osc = Scene(data), oracle_mod.OracleScene(data)
    p = _mpt({"maxBounces": 6})
    p.set_scene(sc)
    img = p.render(96, 96, spp=4, frame=0).cpu().numpy()[..., :3]
    ref, _ = osc.render_mpt(96, 96, spp=4, frame=0, max_bounces=6)
    _compare(img, ref, "standard materials")


def test_war_primal_matches_oracle(gpu, oracle_mod):
    from falcor_amd.render_graph import Scene, WARDiffPathTracer
    data = S.closed_box(4)
    sc, osc = Scene(data), oracle_mod.OracleScene(data)
    p = WARDiffPathTracer({"maxBounces": 2, "useWAR": False})
    p.set_scene(sc)
    img, st = p.render(128, 128, spp=4, frame=0, stats=True)
    ref, rst = osc.render_war(128, 128, spp=4, frame=0, max_bounces=2, brute=True)
    _compare(img.cpu().numpy()[..., :3], ref, "WAR primal")
    assert abs(int(st[1]) - int(rst[1])) <= 0.01 * int(rst[1])


def test_graph_replays_match_eager(gpu):
    """The captured bounce loop replays bit-identically to eager launches, frame after frame, with
    unrelated eager work and synchronisations in between."""
    import torch
    from falcor_amd.render_graph import Scene
    data = S.sphere_field(5, num_spheres=100, num_emissive=300, box=25.0)
    sc = Scene(data)
    pg = _mpt({"maxBounces": 3, "useEmissiveNEE": True, "useGraph": 1})
    pe = _mpt({"maxBounces": 3, "useEmissiveNEE": True, "useGraph": 0})
    pg.set_scene(sc)
    pe.set_scene(sc)
    for f in range(4):
        a = pg.render(80, 60, spp=2, frame=f).clone()
        junk = torch.randn(1 << 16, device="cuda").sum()  # eager work between replays
        torch.cuda.synchronize()
        b = pe.render(80, 60, spp=2, frame=f)
        torch.cuda.synchronize()
        assert torch.equal(a, b), f"frame {f}"
        assert float(junk) == float(junk)


def test_launch_configurations_give_identical_results(gpu):
    """Every compiled traversal configuration (block size x LDS stack depth x grid size) renders
    the same image: traversal results do not depend on the launch, and with 16 LDS entries the
    HBM overflow stack is exercised (the tree needs more)."""
    import torch
    from falcor_amd import autotune
    from falcor_amd.render_graph import Scene, WARDiffPathTracer
    data = S.sphere_field(5, num_spheres=150, num_emissive=400, box=25.0)
    sc = Scene(data)
    need = sc.info()["stack_need"]
    print(f"stack need {need}", flush=True)
    assert need > 16
    pm = _mpt({"maxBounces": 3, "useEmissiveNEE": True})
    pw = WARDiffPathTracer({"maxBounces": 2})
    pm.set_scene(sc)
    pw.set_scene(sc)
    ref_m = pm.render(64, 48, spp=2, frame=3).clone()
    ref_w = pw.render(64, 48, spp=2, frame=3).clone()
    configs = autotune.candidates(threads_per_cu=(256, 2048))
    for cfg in configs:
        print(f"launch {cfg}", flush=True)
        autotune.apply(pm, cfg)
        autotune.apply(pw, cfg)
        assert torch.equal(pm.render(64, 48, spp=2, frame=3), ref_m), cfg
        assert torch.equal(pw.render(64, 48, spp=2, frame=3), ref_w), cfg
    # Visibility rays in launches of their own instead of sharing the closest-hit launch.
    pm.set_property("splitVisibilityLaunch", 1)
    assert torch.equal(pm.render(64, 48, spp=2, frame=3), ref_m)


def test_no_scene_clears_output(gpu):
    import torch
    p = _mpt({})
    out = torch.full((8, 8, 4), 5.0, device="cuda")
    p.render(8, 8, out=out)
    torch.cuda.synchronize()
    assert float(out.abs().max()) == 0.0


# --------------------------------------------------------------------------- adjoint
def _loss_weights(W, H, seed=0):
    return np.random.default_rng(seed).uniform(-1, 1, (H, W, 3)).astype(np.float32)


def _gpu_grad(sc, props, W, H, spp, frame, dLdI, texture=False):
    import torch
    from falcor_amd.render_graph import SceneGradients, WARDiffPathTracer
    p = WARDiffPathTracer(dict(props, diffMode="BackwardDiff"))
    p.set_scene(sc)
    g = SceneGradients(sc, hash_size=16, texture_grads=texture)
    p.scene_gradients = g
    p.dL_dI = torch.from_numpy(dLdI).cuda()
    g.clear_all_grads()
    p.backward(W, H, spp=spp, frame=frame)
    g.aggregate_all_grads()
    mat = g.get_grads_buffer(g.MATERIAL).cpu().numpy().reshape(-1, 20) / spp
    tex = g.get_grads_buffer(g.TEXTURE).cpu().numpy() / spp if texture else None
    return mat, tex


def _fd(osc, base_mats, mats_fn, W, H, spp, frame, dLdI, eps, bounces, texels=None, tex_fn=None):
    """Central difference of sum(dLdI * I) with frozen sampling (sampling/pdfs at the base point)."""
    osc.freeze_sampling(base_mats, texels)
    vals = []
    for sgn in (+1, -1):
        if tex_fn is not None:
            osc.set_texels(tex_fn(sgn * eps))
            osc.set_materials(base_mats)
        else:
            osc.set_materials(mats_fn(sgn * eps))
        img, _ = osc.render_war(W, H, spp=spp, frame=frame * spp, max_bounces=bounces)
        vals.append(float((img.astype(np.float64) * dLdI).sum()))
    osc.freeze_sampling(None)
    osc.set_materials(base_mats)
    if texels is not None:
        osc.set_texels(texels)
    return (vals[0] - vals[1]) / (2 * eps)


def test_adjoint_lambert_albedo_matches_fd(gpu, oracle_mod):
    """Closed box (Lambertian walls): d(sum dLdI * I)/d albedo vs central differences (eps 1e-3)."""
    from falcor_amd.render_graph import Scene
    data = S.closed_box(6)
    W = H = 48
    spp, frame, bounces = 4, 2, 2
    sc, osc = Scene(data), oracle_mod.OracleScene(data)
    dLdI = _loss_weights(W, H)
    mat, _ = _gpu_grad(sc, {"maxBounces": bounces}, W, H, spp, frame, dLdI)
    base = data.materials.copy()
    errs = []
    for m in range(7):
        for c in range(3):
            def mf(e, m=m, c=c):
                mm = base.copy()
                mm[m, 1 + c] += e
                return mm
            fd = _fd(osc, base, mf, W, H, spp, frame, dLdI, 1e-3, bounces)
            errs.append((mat[m, c], fd))
    a = np.array(errs)
    rel = np.linalg.norm(a[:, 0] - a[:, 1]) / np.linalg.norm(a[:, 1])
    print("lambert albedo grad rel err", rel, flush=True)
    assert rel < 1e-2


def test_adjoint_standard_params_and_texels_match_fd(gpu, oracle_mod):
    """Textured StandardMaterial spheres: per-texel baseColor and per-material roughness/metallic
    gradients vs central differences with frozen sampling; texel sums equal the material slot."""
    import torch
    from falcor_amd.render_graph import Scene
    data = S.textured_spheres(3, num_objects=4, tex_res=8, subdivisions=2, target=True)
    data.materials[:4, 5] = [0.0, 0.3, 0.6, 1.0]
    W = H = 32
    spp, frame, bounces = 4, 0, 1
    sc, osc = Scene(data), oracle_mod.OracleScene(data)
    dLdI = _loss_weights(W, H, 1)
    mat, tex = _gpu_grad(sc, {"maxBounces": bounces}, W, H, spp, frame, dLdI, texture=True)
    base, texels = data.materials.copy(), data.texels.copy()
    pairs = []
    for m in range(4):
        for slot, field in ((4, 4), (3, 5)):  # roughness, metallic
            def mf(e, m=m, f=field):
                mm = base.copy()
                mm[m, f] += e
                return mm
            pairs.append((mat[m, slot], _fd(osc, base, mf, W, H, spp, frame, dLdI, 1e-3, bounces)))
    rng = np.random.default_rng(2)
    for ti in rng.choice(texels.size, 24, replace=False):
        def tf(e, ti=ti):
            t = texels.copy()
            t[ti] += e
            return t
        pairs.append((tex[ti], _fd(osc, base, None, W, H, spp, frame, dLdI, 1e-3, bounces, texels=texels, tex_fn=tf)))
    a = np.array(pairs)
    rel = np.linalg.norm(a[:, 0] - a[:, 1]) / np.linalg.norm(a[:, 1])
    print("standard params grad rel err", rel, flush=True)
    assert rel < 1e-2
    # Sum of per-texel gradients of a texture equals its material's baseColor gradient.
    td = data.tex_desc
    for t in range(td.shape[0]):
        off, w, h = td[t, 0], td[t, 1], td[t, 2]
        s = tex[off:off + 3 * w * h].reshape(-1, 3).sum(0)
        np.testing.assert_allclose(s, mat[t, 0:3], rtol=1e-3, atol=1e-5)


@pytest.mark.parametrize("bounces", [1, 4])
def test_adjoint_hit_log_replay_equals_retrace(gpu, bounces):
    """The adjoint replays its first pass's traversal answers from the hit log (useHitLog, the
    default) instead of tracing again: same gradients as the replay that traces (up to the order
    of the float atomics), on a StandardMaterial sphere field with textures, and half the rays."""
    from falcor_amd.render_graph import Scene
    data = S.textured_spheres(3, num_objects=16, tex_res=16, subdivisions=2, target=True)
    W = H = 64
    sc = Scene(data)
    dLdI = _loss_weights(W, H, 1)
    got = {}
    for log in (0, 1):
        got[log] = _gpu_grad(sc, {"maxBounces": bounces, "useHitLog": log}, W, H, 4, 3, dLdI, texture=True)
    for a, b in zip(got[0], got[1]):
        assert np.abs(a).max() > 0
        np.testing.assert_allclose(b, a, rtol=1e-4, atol=1e-6 * np.abs(a).max())
