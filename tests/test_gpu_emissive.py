"""Textured emitters on the GPU against the oracle: the light collection's pre-integrated flux
(coverage-weighted texel average), and images of MinimalPathTracer (BSDF-hit emission and NEE
with both samplers) and of the WARDiffPathTracer primal, whose light samples evaluate the
emissive texture at the sample's uv. Image tolerances as in test_gpu_render.py."""
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


def _textured_light_box(seed=0):
    """The closed box with its ceiling light and an emissive sphere carrying emissive textures."""
    rng = np.random.default_rng(seed)
    data = S.closed_box(seed)
    sb = S.SceneBuilder()
    # Rebuild the closed box's meshes through a builder so new materials can be added.
    offs = data.mesh_vertex_offsets
    light_tex = sb.add_texture(rng.uniform(0.0, 30.0, (8, 8, 3)).astype(np.float32))
    ball_tex = sb.add_texture(rng.uniform(0.0, 3.0, (16, 32, 3)).astype(np.float32))
    mats = [sb.add_material(m) for m in data.materials]
    light_mat = sb.add_material(S.standard_material(base=(0, 0, 0), emissive_factor=1.5, emissive_texture=light_tex))
    ball_mat = sb.add_material(S.standard_material(base=(0.3, 0.3, 0.3), roughness=0.6, emissive_factor=0.8,
                                                   emissive_texture=ball_tex))
    for k, name in enumerate(data.mesh_names):
        v0, v1 = offs[k], offs[k + 1]
        sel = np.all((data.indices >= v0) & (data.indices < v1), axis=1)
        mesh = S.TriangleMesh(data.positions[v0:v1], data.normals[v0:v1], data.texcoords[v0:v1],
                              (data.indices[sel] - v0).astype(np.uint32))
        mat = light_mat if name == "Light" else mats[int(data.tri_material[sel][0])]
        sb.add_mesh_instance(mesh, mat, None, name)
    sb.add_mesh_instance(S.TriangleMesh.create_sphere(0.06, 16, 8), ball_mat, S.Transform((1, 1, 1), (0.12, 0.35, 0.1)), "Ball")
    sb.camera = data.camera
    return sb.build()


def test_flux_matches_oracle(gpu, oracle_mod):
    from falcor_amd.render_graph import Scene
    data = _textured_light_box()
    sc, osc = Scene(data), oracle_mod.OracleScene(data)
    got, ref = sc.lights(), osc.lights()
    assert got.shape[0] == ref.shape[0] > 100
    tri = got[:, 11].view(np.uint32)
    assert np.array_equal(tri, ref[:, 17].view(np.uint32))
    np.testing.assert_allclose(got[:, 7], ref[:, 16], rtol=1e-6)
    assert np.all(got[:, 19] > 0)  # every emitter here is textured


@pytest.mark.parametrize("nee,sampler", [(False, "Uniform"), (True, "Uniform"), (True, "LightBVH")])
def test_mpt_textured_emitters_match_oracle(gpu, oracle_mod, nee, sampler):
    from falcor_amd.render_graph import MinimalPathTracer, Scene
    data = _textured_light_box(1)
    sc, osc = Scene(data), oracle_mod.OracleScene(data)
    p = MinimalPathTracer({"maxBounces": 3, "useEmissiveNEE": nee, "emissiveSampler": sampler})
    p.set_scene(sc)
    img = p.render(96, 96, spp=4, frame=2).cpu().numpy()[..., :3]
    ref, _ = osc.render_mpt(96, 96, spp=4, frame=2, max_bounces=3, nee=nee, light_sampler=0 if sampler == "Uniform" else 1)
    _compare(img, ref, f"textured emitters nee={nee} {sampler}")


def test_war_primal_textured_emitters_match_oracle(gpu, oracle_mod):
    from falcor_amd.render_graph import Scene, WARDiffPathTracer
    data = _textured_light_box(2)
    sc, osc = Scene(data), oracle_mod.OracleScene(data)
    p = WARDiffPathTracer({"maxBounces": 2, "useWAR": False})
    p.set_scene(sc)
    img = p.render(96, 96, spp=4, frame=0).cpu().numpy()[..., :3]
    ref, _ = osc.render_war(96, 96, spp=4, frame=0, max_bounces=2)
    _compare(img, ref, "WAR primal, textured emitters")


def test_emission_updates_reach_light_samples(gpu, oracle_mod):
    """Light samples evaluate the emitter's current material (evalEmissive at the sample,
    EmissiveLightSamplerHelpers.slang:79): after a material update NEE sees the new emission,
    while the light collection (flux, selection) stays as built, as in the reference."""
    from falcor_amd.render_graph import MinimalPathTracer, Scene
    data = S.closed_box(3)
    sc, osc = Scene(data), oracle_mod.OracleScene(data)
    p = MinimalPathTracer({"maxBounces": 2, "useEmissiveNEE": True, "emissiveSampler": "Uniform"})
    p.set_scene(sc)
    before = p.render(64, 64, 2, frame=1).cpu().numpy()[..., :3]
    m = sc.get_material_records()
    light = int(np.nonzero(m[:, S.M_EMIT:S.M_EMIT + 3].max(1) > 0)[0][0])
    m[light, S.M_EMIT:S.M_EMIT + 3] = (5.0, 20.0, 40.0)
    m[light, S.M_EMIT_FACTOR] = 0.5
    sc.set_material_records(m)
    osc.set_materials(m)
    img = p.render(64, 64, 2, frame=1).cpu().numpy()[..., :3]
    ref, _ = osc.render_mpt(64, 64, spp=2, frame=1, max_bounces=2, nee=True, light_sampler=0)
    _compare(img, ref, "emission update, uniform NEE")
    assert not np.allclose(img, before)
