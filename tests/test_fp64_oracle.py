"""Pins the double-precision MinimalPathTracer of the oracle (oracle/fo_f64.h), the reference that
converged GPU images are compared against.

* Same seeds as the FP32 oracle (itself pinned in test_oracle_pins.py): the two images agree
  except where precision changes a branch (at most 1 % of pixels off by more than
  1e-4 (1 + |ref|), relative L1 difference at most 1e-3).
* Furnace closed form (a Lambertian floor under a huge emitter, maxBounces = 0): every floor
  pixel is exactly albedo * Le in double.
* Smooth glass slab in front of an emitter: Le (1 - F) / (1 + F) T^2 (statistical, 0.6 %).
* Importance and reference (cosine) sampling converge to the same closed-box image.
"""
import math

import numpy as np
import pytest

from falcor_amd import scene as S


def _furnace(albedo=0.6, le=17.0):
    sb = S.SceneBuilder()
    floor = sb.add_material(S.pbrt_diffuse_material((albedo, albedo, albedo)))
    light = sb.add_material(S.standard_material(base=(0, 0, 0), roughness=1.0, emissive=(le, le, le)))
    quad = S.TriangleMesh.create_quad()
    sb.add_mesh_instance(quad, floor, S.Transform((2.0, 1.0, 2.0)), "floor")
    sb.add_mesh_instance(quad, light, S.Transform.deg((1e4, 1.0, 1e4), (0, 1.0, 0), (180, 0, 0)), "sky")
    sb.camera = S.Camera((0, 0.5, 0.0), (0, 0, 0.0), (0, 0, -1), math.radians(40), 1.0)
    return sb.build()


def test_fp64_matches_fp32_oracle_on_same_seeds(oracle_mod):
    data = S.closed_box(1)
    sc = oracle_mod.OracleScene(data)
    for bounces in (0, 3):
        a = sc.render_mpt_f64(64, 64, spp=4, frame=2, max_bounces=bounces)
        b, _ = sc.render_mpt(64, 64, spp=4, frame=2, max_bounces=bounces)
        diff = np.abs(a - b)
        bad = (diff > 1e-4 * (1 + np.abs(a))).any(-1).mean()
        l1 = diff.sum() / np.abs(a).sum()
        assert bad <= 0.01 and l1 <= 1e-3, (bounces, bad, l1)
        assert a.mean() > 0.01


def test_fp64_furnace_closed_form(oracle_mod):
    sc = oracle_mod.OracleScene(_furnace())
    img = sc.render_mpt_f64(16, 16, spp=8, max_bounces=0)
    want = float(np.float32(0.6)) * 17.0  # albedo as stored (float32) times Le, in double
    np.testing.assert_allclose(img, want, rtol=1e-14)


def test_fp64_smooth_slab(oracle_mod):
    c = np.array([0.64, 0.36, 0.81])
    sc = oracle_mod.OracleScene(S.glass_slab(tuple(c)))
    img = sc.render_mpt_f64(32, 32, spp=64, max_bounces=20)
    np.testing.assert_allclose(img.reshape(-1, 3).mean(0), 0.96 / 1.04 * c, rtol=6e-3)


def test_fp64_importance_and_reference_sampling_agree(oracle_mod):
    sc = oracle_mod.OracleScene(S.closed_box(2))
    a = sc.render_mpt_f64(24, 24, spp=256, max_bounces=2)
    b = sc.render_mpt_f64(24, 24, spp=256, max_bounces=2, importance=False)
    assert abs(a.mean() - b.mean()) < 0.03 * a.mean()
