"""AccumulatePass and ImageCompare on the GPU (fhp_accumulate, fhp_image_compare) against the
CPU restatement in oracle/image_ops.py.

Accumulated frames, per-pixel error maps and heat maps are compared bit for bit: the kernels
perform the reference's float / double operations in its order, without fused multiply-adds.
The image-comparison total is a sum over pixels in another order than the reference's
sequential loop, so it is compared with the exactly rounded sum (math.fsum) at relative
tolerance 1e-12.
"""
import numpy as np
import pytest

from oracle import image_ops as O

pytestmark = pytest.mark.gpu

MODES = {"Double": O.DOUBLE, "Single": O.SINGLE, "SingleCompensated": O.SINGLE_COMPENSATED}
OVERFLOW = {"Stop": O.STOP, "Reset": O.RESET, "EMA": O.EMA}


def _frames(seed, n, h, w, dtype="float32"):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        if dtype == "float32":
            x = (rng.standard_normal((h, w, 4)) * 3.0).astype(np.float32)
            x[0, 0] = [0.1, -0.0, 1e-30, 3e30]  # a fraction, a negative zero, a denormal-scale and a large value
        elif dtype == "int32":
            x = rng.integers(-2**31, 2**31 - 1, (h, w, 4), dtype=np.int64).astype(np.int32)
        else:
            x = rng.integers(0, 2**32 - 1, (h, w, 4), dtype=np.uint64).astype(np.uint32)
        out.append(x)
    return out


def _to_dev(x):
    import torch
    if x.dtype == np.uint32:
        return torch.from_numpy(x.view(np.int32)).cuda().view(torch.uint32)
    return torch.from_numpy(x).cuda()


def _run(props, frames, oracle, refresh=()):
    from falcor_amd.render_graph import AccumulatePass
    p = AccumulatePass(props)
    for k, f in enumerate(frames):
        r = k in refresh
        got = p.accumulate(_to_dev(f), refresh=r).cpu().numpy()
        want = oracle.execute(f, refresh=r)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (props, k)
        assert p.frame_count == oracle.count, (props, k)
    return p


@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("max_frames", [0, 3])
def test_accumulate_modes_bit_exact(gpu, mode, max_frames):
    fr = _frames(1, 7, 9, 13)
    props = {"precisionMode": mode, "maxFrameCount": max_frames, "overflowMode": "EMA"}
    _run(props, fr, O.AccumulateOracle(precision=MODES[mode], max_frames=max_frames, overflow=O.EMA))


@pytest.mark.parametrize("dtype", ["int32", "uint32"])
@pytest.mark.parametrize("mode", list(MODES))
def test_accumulate_integer_inputs(gpu, dtype, mode):
    fr = _frames(2, 4, 5, 6, dtype)
    _run({"precisionMode": mode}, fr, O.AccumulateOracle(precision=MODES[mode]))


@pytest.mark.parametrize("overflow", list(OVERFLOW))
def test_accumulate_overflow_modes(gpu, overflow):
    fr = _frames(3, 8, 4, 4)
    _run({"maxFrameCount": 3, "overflowMode": overflow}, fr,
         O.AccumulateOracle(max_frames=3, overflow=OVERFLOW[overflow]))


def test_accumulate_disabled_and_resets(gpu):
    fr = _frames(4, 6, 6, 5)
    ints = _frames(5, 2, 6, 5, "int32")
    _run({"enabled": False}, fr[:2] + ints, O.AccumulateOracle(enabled=False))
    # refresh with autoReset restarts; without autoReset it is ignored.
    _run({}, fr, O.AccumulateOracle(), refresh=(2, 4))
    _run({"autoReset": False}, fr, O.AccumulateOracle(auto_reset=False), refresh=(2, 4))
    # Resolution change.
    from falcor_amd.render_graph import AccumulatePass
    p = AccumulatePass({})
    o = O.AccumulateOracle()
    for f in fr[:2] + _frames(6, 2, 3, 7) + fr[2:3]:
        got = p.accumulate(_to_dev(f)).cpu().numpy()
        assert np.array_equal(got.view(np.uint32), o.execute(f).view(np.uint32))
        assert p.frame_count == o.count


def test_accumulate_property_changes_reset(gpu):
    from falcor_amd.render_graph import AccumulatePass
    fr = _frames(7, 6, 5, 5)
    p = AccumulatePass({})
    o = O.AccumulateOracle()
    steps = [None, None, ("enabled", False), None, ("enabled", True), None, ("precisionMode", "Double"), None,
             ("precisionMode", "SingleCompensated"), None]
    k = 0
    for s in steps:
        if s is None:
            got = p.accumulate(_to_dev(fr[k % len(fr)])).cpu().numpy()
            want = o.execute(fr[k % len(fr)])
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), k
            k += 1
        elif s[0] == "enabled":
            p.enabled = s[1]
            o.set_enabled(s[1])
        else:
            p.set_property("precisionMode", s[1])
            o.set_precision(MODES[s[1]])
        assert p.frame_count == o.count


def test_accumulate_full_hd_frames(gpu):
    """1920 x 1080, 4 frames of each precision mode."""
    fr = _frames(8, 4, 1080, 1920)
    for mode in MODES:
        _run({"precisionMode": mode}, fr, O.AccumulateOracle(precision=MODES[mode]))


def test_accumulate_after_path_tracer_in_graph(gpu):
    """MinimalPathTracer -> AccumulatePass: frames average; a material change restarts (autoReset)."""
    import torch
    from falcor_amd import scene as S
    from falcor_amd.render_graph import RenderGraph, Scene
    sc = Scene(S.closed_box())
    g = RenderGraph("MinimalPathTracer")
    pt = g.create_pass("MinimalPathTracer", "MinimalPathTracer", {"maxBounces": 2})
    acc = g.create_pass("AccumulatePass", "AccumulatePass", {"enabled": True, "precisionMode": "Single"})
    g.add_edge("MinimalPathTracer.color", "AccumulatePass.input")
    g.mark_output("AccumulatePass.output")
    frames = []
    for _ in range(3):
        v = g.execute(sc, 32, 24)
        frames.append(v["MinimalPathTracer.color"].clone())
    out = g.get_output("AccumulatePass.output")
    mean = (frames[0] + frames[1] + frames[2]) * np.float32(1.0 / 3.0)
    assert torch.allclose(out, mean, rtol=1e-6, atol=1e-7)
    assert acc.frame_count == 3
    m = sc.get_material_records()
    sc.set_material_records(m)
    v = g.execute(sc, 32, 24)
    assert acc.frame_count == 1
    assert torch.equal(v["AccumulatePass.output"], v["MinimalPathTracer.color"])


def test_accumulate_argument_errors(gpu):
    import torch
    from falcor_amd.render_graph import AccumulatePass
    p = AccumulatePass({})
    with pytest.raises(ValueError):
        p.accumulate(torch.zeros((4, 4, 4), dtype=torch.float64, device="cuda"))
    with pytest.raises(ValueError):
        p.accumulate(torch.zeros((4, 4, 3), dtype=torch.float32, device="cuda"))


# --------------------------------------------------------------------------- ImageCompare
def _pair(seed, h, w):
    rng = np.random.default_rng(seed)
    a = rng.random((h, w, 4)).astype(np.float32) * np.float32(2.0)
    b = a + (rng.standard_normal((h, w, 4)) * 0.1).astype(np.float32)
    return a, b


@pytest.mark.parametrize("metric", O.METRICS)
@pytest.mark.parametrize("alpha", [False, True])
def test_image_compare_matches_oracle(gpu, metric, alpha):
    from falcor_amd.image_compare import compare_tensors
    for h, w in ((1, 1), (7, 13), (1080, 1920)):
        a, b = _pair(h * 31 + w, h, w)
        err, emap, heat = compare_tensors(_to_dev(a), _to_dev(b), metric, alpha=alpha, heat_map=True)
        want, wmap = O.compare(a, b, metric, alpha)
        assert err == pytest.approx(want, rel=1e-12), (metric, h, w)
        assert np.array_equal(emap.cpu().numpy().view(np.uint32), wmap.view(np.uint32)), (metric, h, w)
        assert np.array_equal(heat.cpu().numpy().view(np.uint32), O.heat_map(wmap).view(np.uint32)), (metric, h, w)


def test_image_compare_identical_images(gpu):
    from falcor_amd.image_compare import compare_tensors
    a, _ = _pair(5, 16, 16)
    for metric in O.METRICS:
        err, _, _ = compare_tensors(_to_dev(a), _to_dev(a), metric)
        assert err == 0.0


def test_image_compare_files_threshold_and_heat_map(gpu, tmp_path):
    from falcor_amd import image_io as I
    from falcor_amd.image_compare import compare_files, main
    a, b = _pair(9, 12, 10)
    I.write_exr(tmp_path / "a.exr", a)
    I.write_exr(tmp_path / "b.exr", b)
    want, wmap = O.compare(a, b, "mse")
    ok, err = compare_files(tmp_path / "a.exr", tmp_path / "b.exr", "mse", threshold=want * 1.01,
                            heat_map_path=tmp_path / "h.exr")
    assert ok and err == pytest.approx(want, rel=1e-12)
    assert np.array_equal(I.read_rgba(tmp_path / "h.exr"), O.heat_map(wmap))
    assert main([str(tmp_path / "a.exr"), str(tmp_path / "b.exr"), "-t", repr(want * 0.5)]) == 1
    assert main([str(tmp_path / "a.exr"), str(tmp_path / "a.exr")]) == 0
    I.write_exr(tmp_path / "c.exr", a[:5])
    assert main([str(tmp_path / "a.exr"), str(tmp_path / "c.exr")]) == 1
