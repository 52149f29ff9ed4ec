"""Batched MinimalPathTracer (fhp_pass_execute_batch): views of different scenes, frame shapes,
sample counts and cameras in one launch sequence. Each view's path state, seeds and traversal are
those of a single-view frame, so every view must equal its single-view render bit for bit."""
import math

import numpy as np
import pytest

from falcor_amd import scene as S

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nee", [False, True])
def test_batch_views_equal_single_view_renders(gpu, nee):
    import torch
    from falcor_amd.render_graph import MinimalPathTracer, Scene
    box = Scene(S.closed_box(3))
    field = Scene(S.sphere_field(5, num_spheres=120, num_emissive=300, box=25.0))
    cam2 = S.Camera((0.05, 0.3, 1.1), (0.0, 0.25, 0.0), (0, 1, 0), math.radians(35.0), 48 / 40)
    views = [
        dict(scene=box, width=64, height=48, spp=2, frame=5),
        dict(scene=field, width=40, height=32, spp=1, frame=2),
        dict(scene=box, width=48, height=40, spp=3, frame=1, camera=cam2),
        dict(scene=field, width=33, height=17, spp=4, frame=9),
    ]
    props = {"maxBounces": 3, "useEmissiveNEE": nee}
    p = MinimalPathTracer(props)
    outs, st = p.render_batch(views, stats=True)
    torch.cuda.synchronize()
    rays = 0
    for i, v in enumerate(views):
        q = MinimalPathTracer(props)
        if "camera" in v:
            d = S.closed_box(3, aspect=48 / 40)
            d.camera = v["camera"]
            sc = Scene(d)
        else:
            sc = v["scene"]
        q.set_scene(sc)
        ref, rs = q.render(v["width"], v["height"], v["spp"], frame=v["frame"], stats=True)
        rays += int(rs[0])
        print(f"view {i}: mean {float(ref.mean()):.5f}", flush=True)
        assert torch.equal(outs[i], ref), f"view {i}"
    assert int(st[0]) == rays


def test_batch_rejects_bad_views(gpu):
    from falcor_amd import _lib
    from falcor_amd.render_graph import MinimalPathTracer, Scene, WARDiffPathTracer
    box = Scene(S.closed_box(0))
    p = MinimalPathTracer({})
    with pytest.raises(_lib.FhpError, match="view 0"):
        p.render_batch([dict(scene=box, width=0, height=8)])
    with pytest.raises(_lib.FhpError, match="frame shape"):
        p.render_batch([dict(scene=box, width=8, height=8, spp=0)])
    w = WARDiffPathTracer({})
    with pytest.raises(_lib.FhpError, match="MinimalPathTracer"):
        MinimalPathTracer.render_batch(w, [dict(scene=box, width=8, height=8)])
