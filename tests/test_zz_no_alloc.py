"""fhp_pass_execute_noalloc: the MinimalPathTracer frame of fhp_pass_execute with its wavefront
state in the library's static device scratch array (no allocation). Called through the C ABI,
it must equal fhp_pass_execute bit for bit on two calls in a row, the second with a smaller
frame (the array keeps the first call's contents), and refuse frames beyond its maximum."""
import numpy as np
import pytest

from falcor_amd import scene as S

pytestmark = pytest.mark.gpu

MAX_PATHS = 1 << 19


def _noalloc(L, _lib, p, sc, W, H, spp, frame):
    import torch
    out = torch.full((H, W, 4), -1.0, device="cuda")
    rc = L.fhp_pass_execute_noalloc(p.handle, sc.handle, W, H, spp, frame, out.data_ptr(), _lib.stream_ptr())
    _lib.check(rc)
    torch.cuda.synchronize()
    return out


def test_noalloc_frame_equals_pass_execute(gpu):
    import torch
    from falcor_amd import _lib
    from falcor_amd.render_graph import MinimalPathTracer, Scene
    L = _lib.load()
    # A deep tree (the traversal stack spills to the overflow part) at the maximum path count,
    # then a small closed box frame: the second call runs on the first call's leftover state.
    field = Scene(S.sphere_field(2, num_spheres=400, num_emissive=1000, box=40.0))
    box = Scene(S.closed_box(4))
    cases = [(field, {"maxBounces": 4, "useEmissiveNEE": True, "emissiveSampler": "LightBVH"}, 512, 512, 2, 7),
             (box, {"maxBounces": 3}, 96, 80, 3, 2)]
    assert cases[0][2] * cases[0][3] * cases[0][4] == MAX_PATHS
    for sc, props, W, H, spp, frame in cases:
        p = MinimalPathTracer(props)
        p.set_scene(sc)
        got = _noalloc(L, _lib, p, sc, W, H, spp, frame)
        want = p.render(W, H, spp, frame=frame)
        torch.cuda.synchronize()
        assert torch.equal(got, want), (W, H, spp)
        assert float(got[..., :3].abs().sum()) > 0
        assert sc.trace_errors() == 0


def test_noalloc_refuses_frames_beyond_its_maximum(gpu):
    import torch
    from falcor_amd import _lib
    from falcor_amd.render_graph import MinimalPathTracer, Scene
    L = _lib.load()
    sc = Scene(S.closed_box(0))
    p = MinimalPathTracer({"maxBounces": 1})
    p.set_scene(sc)
    out = torch.empty((1024, 1024, 4), device="cuda")
    for W, H, spp in ((1024, 1024, 1), (512, 512, 3), (2048, 257, 1)):
        rc = L.fhp_pass_execute_noalloc(p.handle, sc.handle, W, H, spp, 0, out.data_ptr(), _lib.stream_ptr())
        assert rc != 0 and "too large" in L.fhp_last_error().decode()
    wt = MinimalPathTracer({"maxBounces": 1})
    from falcor_amd.render_graph import WARDiffPathTracer
    w = WARDiffPathTracer({})
    rc = L.fhp_pass_execute_noalloc(w.handle, sc.handle, 8, 8, 1, 0, out.data_ptr(), _lib.stream_ptr())
    assert rc != 0 and "MinimalPathTracer" in L.fhp_last_error().decode()
    del wt
