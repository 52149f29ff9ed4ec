"""Pins the oracle's analytic lights in MinimalPathTracer (evalDirectAnalytic,
MinimalPathTracer.rt.slang:253-286; samplePointLight / sampleDirectionalLight,
LightHelpers.slang:190-258) with closed forms:

* a point light over a Lambertian floor, direct light only: every floor pixel is
  albedo / pi * I * cos(theta) / d^2 at its hit point;
* a spot light: zero outside the opening angle, the full value inside the opening minus the
  penumbra, the smoothstep of the angle margin in between;
* a directional light: albedo / pi * E * cos(theta) everywhere;
* several lights picked uniformly (contribution times the light count) average to the sum of
  the single-light images;
* the .pyscene loader turns the reference CornellBox's PointLight into a light record.
"""
import math
from pathlib import Path

import numpy as np
import pytest

from falcor_amd import scene as S
from oracle import hitinfo as HI

W = H = 32
ALBEDO = 0.6


def _floor_scene(*lights):
    sb = S.SceneBuilder()
    m = sb.add_material(S.pbrt_diffuse_material((ALBEDO, ALBEDO, ALBEDO)))
    sb.add_mesh_instance(S.TriangleMesh.create_quad(), m, S.Transform((4, 1, 4)), "floor")
    for l in lights:
        sb.add_light(l)
    sb.camera = S.Camera((0, 2, 0), (0, 0, 0), (0, 0, -1), math.radians(60.0), 1.0)
    return sb.build()


def _floor_hits(data):
    cam = data.camera_array().astype(np.float64)
    d = HI.camera_dirs(cam, W, H)
    t = -cam[1] / d[..., 1]
    return cam[:3] + d * t[..., None]


def test_point_light_closed_form(oracle_mod):
    I, lp = np.array([2.0, 1.0, 0.5]), np.array([0.3, 1.0, -0.2])
    data = _floor_scene(S.point_light(I, lp))
    img, st = oracle_mod.OracleScene(data).render_mpt(W, H, spp=1, max_bounces=0)
    x = _floor_hits(data)
    v = lp - x
    d2 = (v ** 2).sum(-1)
    cos = v[..., 1] / np.sqrt(d2)
    want = ALBEDO / math.pi * I * (cos / d2)[..., None]
    np.testing.assert_allclose(img, want, rtol=2e-5)
    assert st[1] == W * H  # one shadow ray per pixel


def test_spot_light_cone_and_penumbra(oracle_mod):
    I, lp = np.array([3.0, 3.0, 3.0]), np.array([0.0, 1.5, 0.0])
    opening, penumbra = math.radians(25.0), math.radians(10.0)
    data = _floor_scene(S.point_light(I, lp, (0, -1, 0), opening, penumbra))
    img, _ = oracle_mod.OracleScene(data).render_mpt(W, H, spp=1, max_bounces=0)
    x = _floor_hits(data)
    v = lp - x
    d2 = (v ** 2).sum(-1)
    cos = v[..., 1] / np.sqrt(d2)
    ang = np.arccos(cos)  # angle from the spot axis (straight down)
    t = np.clip((opening - ang) / penumbra, 0, 1)
    fall = np.where(ang > opening, 0.0, t * t * (3 - 2 * t))
    want = ALBEDO / math.pi * I * (fall * cos / d2)[..., None]
    np.testing.assert_allclose(img, want, rtol=1e-4, atol=1e-6)
    assert (fall == 0).mean() > 0.2 and (fall == 1).mean() > 0.05 and ((fall > 0) & (fall < 1)).mean() > 0.05


def test_directional_light_closed_form(oracle_mod):
    E, dirn = np.array([1.5, 1.0, 0.25]), np.array([0.3, -1.0, 0.2])
    data = _floor_scene(S.directional_light(E, dirn))
    img, _ = oracle_mod.OracleScene(data).render_mpt(W, H, spp=1, max_bounces=0)
    cos = 1.0 / np.linalg.norm(dirn)
    np.testing.assert_allclose(img, np.broadcast_to(ALBEDO / math.pi * E * cos, img.shape), rtol=1e-5)


def test_uniform_light_pick_averages_to_the_sum(oracle_mod):
    # Two point lights inside the closed box (its walls and ceiling block anything outside).
    a = S.point_light((0.3, 0.2, 0.1), (0.1, 0.4, 0.05))
    b = S.point_light((0.05, 0.1, 0.2), (-0.12, 0.3, 0.1), (0, -1, 0), 1.0, 0.2)
    data = S.closed_box(1)
    imgs = []
    for lights in ([a, b], [a], [b], []):
        data.analytic_lights = lights
        img, _ = oracle_mod.OracleScene(data).render_mpt(32, 32, spp=256, max_bounces=2)
        imgs.append(img.mean())
    # The box's emissive ceiling lights every image; subtract it.
    both, only_a, only_b, none = imgs
    assert abs((both - none) - (only_a - none) - (only_b - none)) < 0.02 * (both - none)
    assert only_a - none > 0.05 and only_b - none > 0.01


def test_pyscene_point_light(tmp_path):
    from falcor_amd.pyscene import load_pyscene
    (tmp_path / "l.pyscene").write_text(
        "m = StandardMaterial('m')\nq = TriangleMesh.createQuad()\n"
        "sceneBuilder.addMeshInstance(sceneBuilder.addNode('q'), sceneBuilder.addTriangleMesh(q, m))\n"
        "p = PointLight('p')\np.intensity = float3(1, 2, 3)\np.position = float3(0, 1, 0)\n"
        "p.openingAngle = 0.5\np.penumbraAngle = 0.1\np.direction = float3(0, -2, 0)\nsceneBuilder.addLight(p)\n"
        "d = DirectionalLight('d')\nd.direction = float3(1, -1, 0)\nd.intensity = float3(0.5, 0.5, 0.5)\n"
        "sceneBuilder.addLight(d)\n")
    d = load_pyscene(tmp_path / "l.pyscene")
    np.testing.assert_array_equal(d.analytic_lights[0],
                                  S.point_light((1, 2, 3), (0, 1, 0), (0, -2, 0), 0.5, 0.1))
    np.testing.assert_array_equal(d.analytic_lights[1], S.directional_light((0.5, 0.5, 0.5), (1, -1, 0)))
    ref = Path("/root/reference/Scene/Box/CornellBox.pyscene")
    if ref.exists():
        c = load_pyscene(ref)
        assert len(c.analytic_lights) == 1
        np.testing.assert_allclose(c.analytic_lights[0][1:7], [0.2, 0.2, 0.2, 0, 0.4, 0.05], rtol=1e-6)
