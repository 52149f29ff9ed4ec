"""The bounce loop's tail kernel (MinimalPathTracer property `tailPaths`).

Once a depth holds at most `tailPaths` hits, one launch runs those paths to their end in one thread
each instead of launching traversal and shading per depth. Its steps and its order of radiance
additions are those of the per-depth launches, so every image must equal the render with the tail
off (`tailPaths` 0) bit for bit, whichever depth the tail starts at (a threshold of 1 starts it at
the last depths only, 2^24 at the first), and the traced ray counts must be the same. The scenes
cover each shading specialisation: emissive NEE only (LightBVH and uniform), any light mix
(analytic lights, environment map NEE with missed rays shaded), no light sampling, delta and
transmission lobes, and computeDirect off.
"""
import numpy as np
import pytest

from falcor_amd import scene as S

pytestmark = pytest.mark.gpu

# (tailPaths, tailLanes): thresholds from the last depths only to the first, with 32, 8 and 1
# paths per wave of the tail launch.
THRESHOLDS = ((1, 32), (300, 8), (4000, 1), (4000, 32), (1 << 24, 8))


def _transmissive(data, seed):
    rng = np.random.default_rng(seed)
    m = data.materials
    for i in range(0, len(m), 3):
        m[i, S.M_ROUGH], m[i, S.M_METAL], m[i, S.M_SPEC_TRANS] = float(rng.uniform(0.0, 0.5)), 0.0, 1.0
    for i in range(1, len(m), 4):
        m[i, S.M_DIFF_TRANS] = float(rng.uniform(0.3, 1.0))
        m[i, S.M_THIN] = float(i % 2)
    return data


def _cases():
    from falcor_amd.render_graph import Scene
    field = S.sphere_field(11, num_spheres=200, num_emissive=600, box=25.0)
    yield "field NEE LightBVH", Scene(field), {"maxBounces": 6, "useEmissiveNEE": True, "emissiveSampler": "LightBVH"}
    yield "field NEE uniform", Scene(field), {"maxBounces": 4, "useEmissiveNEE": True, "emissiveSampler": "Uniform"}
    yield "field no NEE", Scene(field), {"maxBounces": 5}
    yield "field computeDirect off", Scene(field), {"maxBounces": 3, "computeDirect": False}
    glass = _transmissive(S.sphere_field(12, num_spheres=150, num_emissive=400, box=20.0), 12)
    yield "transmissive NEE", Scene(glass), {"maxBounces": 6, "useEmissiveNEE": True}
    lit = S.sphere_field(13, num_spheres=150, num_emissive=300, box=20.0)
    lit.analytic_lights = [S.directional_light((1.2, 1.0, 0.8), (0.3, -1.0, 0.2)),
                           S.point_light((400.0, 300.0, 200.0), (10.0, 24.0, 10.0))]
    yield "analytic + emissive NEE", Scene(lit), {"maxBounces": 5, "useEmissiveNEE": True}
    env = Scene(S.sphere_field(14, num_spheres=120, num_emissive=200, box=20.0))
    env.set_env_map(np.random.default_rng(3).uniform(0, 2, (16, 32, 3)).astype(np.float32), rotation=(0, 30, 0))
    yield "env NEE + emissive NEE", env, {"maxBounces": 4, "useEmissiveNEE": True, "useEnvNEE": True}
    yield "env light without NEE", env, {"maxBounces": 3}


def test_tail_equals_per_depth_launches(gpu):
    import torch
    from falcor_amd.render_graph import MinimalPathTracer
    W, H, spp = 96, 72, 3
    for name, sc, props in _cases():
        p = MinimalPathTracer(dict(props, tailPaths=0))
        p.set_scene(sc)
        ref, rs = p.render(W, H, spp, frame=5, stats=True)
        ref = ref.clone()
        rays = (int(rs[0]), int(rs[1]))
        print(f"{name}: mean {float(ref[..., :3].mean()):.5f}, rays {rays}", flush=True)
        for t, lanes in THRESHOLDS:
            p.set_property("tailPaths", t)
            p.set_property("tailLanes", lanes)
            img, st = p.render(W, H, spp, frame=5, stats=True)
            assert torch.equal(img, ref), f"{name}, tailPaths {t}, tailLanes {lanes}"
            assert (int(st[0]), int(st[1])) == rays, f"{name}, tailPaths {t}, tailLanes {lanes}"
            assert int(st[6]) <= rays[0] and int(st[7]) <= rays[1]
            if t == 1 << 24 and props["maxBounces"] > 0:
                assert int(st[6]) > 0, name  # the tail started at depth 0 and traced every bounce ray
        assert int(rs[6]) == 0 and int(rs[7]) == 0


def test_tail_in_batches_and_graph_replays(gpu):
    """A batch whose depths cross the threshold at different points than its single views still
    equals them, and the captured loop with the tail replays like eager frames."""
    import torch
    from falcor_amd.render_graph import MinimalPathTracer, Scene
    a = Scene(S.sphere_field(15, num_spheres=150, num_emissive=300, box=20.0))
    b = Scene(S.closed_box(2))
    views = [dict(scene=a, width=80, height=60, spp=2, frame=3), dict(scene=b, width=64, height=48, spp=4, frame=1)]
    props = {"maxBounces": 5, "useEmissiveNEE": True, "tailPaths": 2000}
    pb = MinimalPathTracer(props)
    outs = [o.clone() for o in pb.render_batch(views)]
    for i, v in enumerate(views):
        q = MinimalPathTracer(dict(props, tailPaths=0))
        q.set_scene(v["scene"])
        assert torch.equal(outs[i], q.render(v["width"], v["height"], v["spp"], frame=v["frame"])), f"view {i}"
    pg = MinimalPathTracer(dict(props, useGraph=1))
    pe = MinimalPathTracer(dict(props, useGraph=0, tailPaths=0))
    pg.set_scene(a)
    pe.set_scene(a)
    for f in range(3):
        g = pg.render(80, 60, spp=2, frame=f).clone()
        junk = torch.randn(1 << 16, device="cuda").sum()
        torch.cuda.synchronize()
        assert torch.equal(g, pe.render(80, 60, spp=2, frame=f)), f"frame {f}"
        assert float(junk) == float(junk)


def test_tail_property_checks(gpu):
    from falcor_amd.render_graph import MinimalPathTracer
    p = MinimalPathTracer({})
    assert p.get_property("tailPaths") == 65536
    assert p.get_property("tailLanes") == 8
    for bad in (-1, 0.5, float(1 << 25)):
        with pytest.raises(Exception):
            p.set_property("tailPaths", bad)
    for bad in (0, 33, 2.5):
        with pytest.raises(Exception):
            p.set_property("tailLanes", bad)
