"""Pins the AccumulatePass and ImageCompare oracle (oracle/image_ops.py) with closed forms and
properties a plausible mistake breaks: a wrong weight, a dropped or sign-flipped Kahan term, a
frame-count update in the wrong place, an overflow mode mixed up, a metric with a dropped term.
"""
import math

import numpy as np
import pytest

from oracle import image_ops as O


def _frames(seed, n, h=4, w=5, scale=1.0):
    rng = np.random.default_rng(seed)
    return [(rng.random((h, w, 4)) * scale).astype(np.float32) for _ in range(n)]


def test_single_mean_of_integers():
    # Frames 1, 2, ..., n: the sum n (n + 1) / 2 is exact and the mean is sum * fl(1 / n): exact
    # when n is a power of two, within one ulp of (n + 1) / 2 otherwise.
    acc = O.AccumulateOracle(precision=O.SINGLE)
    for n in range(1, 20):
        out = acc.execute(np.full((2, 3, 4), n, np.float32))
        want = np.float32((n + 1) / 2)
        if n & (n - 1) == 0:
            assert np.all(out == want), n
        assert np.all(np.abs(out - want) <= np.spacing(want)), n
    assert acc.count == 19


@pytest.mark.parametrize("precision", [O.SINGLE, O.SINGLE_COMPENSATED, O.DOUBLE])
def test_running_mean_matches_float64_mean(precision):
    fr = _frames(1, 50)
    acc = O.AccumulateOracle(precision=precision)
    for k, f in enumerate(fr):
        out = acc.execute(f)
        ref = np.mean(np.stack(fr[:k + 1]).astype(np.float64), axis=0)
        np.testing.assert_allclose(out, ref, rtol=2e-6, atol=1e-7)


def test_double_is_correctly_rounded_mean():
    fr = _frames(2, 30, scale=1e3)
    acc = O.AccumulateOracle(precision=O.DOUBLE)
    for f in fr:
        out = acc.execute(f)
    vals = np.stack(fr).astype(np.float64).reshape(len(fr), -1)
    exact = np.array([math.fsum(vals[:, i]) for i in range(vals.shape[1])]) / len(fr)
    # The double sum of 30 float32 values is exact to far below float32 resolution.
    assert np.array_equal(out.ravel(), exact.astype(np.float32))


def test_kahan_beats_naive_summation():
    # Accumulating many frames of a value that is not a float32 fraction of its running sum:
    # naive float32 summation drifts, compensated summation stays within an ulp of the mean.
    v = np.float32(0.1)
    naive = O.AccumulateOracle(precision=O.SINGLE)
    kahan = O.AccumulateOracle(precision=O.SINGLE_COMPENSATED)
    f = np.full((1, 1, 4), v, np.float32)
    for _ in range(20000):
        a = naive.execute(f)[0, 0, 0]
        b = kahan.execute(f)[0, 0, 0]
    assert abs(float(b) - float(v)) <= float(np.spacing(v))
    assert abs(float(a) - float(v)) > 50 * float(np.spacing(v))


def test_kahan_with_flipped_correction_fails():
    # A plausible mistake (y = cur + c) gives errors no better than naive summation.
    v = np.float32(0.1)
    s = c = np.float32(0)
    for _ in range(20000):
        y = v + c
        sn = s + y
        c = (sn - s) - y
        s = sn
    assert abs(float(s / np.float32(20000)) - float(v)) > 50 * float(np.spacing(v))


def test_moving_average_closed_form():
    # maxFrameCount N, overflow EMA: the first N frames average; then every frame blends in with
    # weight 1 / (N + 1): after k frames of b, value = b + (m - b) * (N / (N + 1))^k.
    N = 4
    for precision in (O.SINGLE, O.DOUBLE):
        acc = O.AccumulateOracle(precision=precision, max_frames=N, overflow=O.EMA)
        a, b = 2.0, 10.0
        for _ in range(N):
            out = acc.execute(np.full((1, 2, 4), a, np.float32))
        assert np.allclose(out, a)
        for k in range(1, 8):
            out = acc.execute(np.full((1, 2, 4), b, np.float32))
            want = b + (a - b) * (N / (N + 1)) ** k
            assert np.allclose(out, want, rtol=1e-6), (precision, k)
        assert acc.count == N


def test_moving_average_without_overflow_equals_mean():
    # Below the limit the moving average is the running mean.
    fr = _frames(3, 6)
    acc = O.AccumulateOracle(precision=O.SINGLE, max_frames=10)
    for k, f in enumerate(fr):
        out = acc.execute(f)
    np.testing.assert_allclose(out, np.mean(np.stack(fr), axis=0), rtol=2e-6)


def test_overflow_stop_keeps_output():
    acc = O.AccumulateOracle(max_frames=3, overflow=O.STOP)
    fr = _frames(4, 5)
    for f in fr[:3]:
        out = acc.execute(f).copy()
    for f in fr[3:]:
        assert np.array_equal(acc.execute(f), out)
    assert acc.count == 3


def test_overflow_reset_restarts():
    acc = O.AccumulateOracle(max_frames=2, overflow=O.RESET)
    fr = _frames(5, 3)
    for f in fr[:2]:
        acc.execute(f)
    out = acc.execute(fr[2])
    assert np.array_equal(out, fr[2])
    assert acc.count == 1


def test_compensated_ignores_frame_limit():
    acc = O.AccumulateOracle(precision=O.SINGLE_COMPENSATED, max_frames=2, overflow=O.EMA)
    fr = _frames(6, 5)
    for f in fr:
        out = acc.execute(f)
    assert acc.count == 5
    np.testing.assert_allclose(out, np.mean(np.stack(fr), axis=0), rtol=2e-6)


def test_disabled_passes_float_through_and_counts_integer_frames():
    acc = O.AccumulateOracle(enabled=False)
    f = _frames(7, 1)[0]
    assert np.array_equal(acc.execute(f), f) and acc.count == 0
    i = np.arange(2 * 3 * 4, dtype=np.int32).reshape(2, 3, 4) - 7
    assert np.array_equal(acc.execute(i), i.astype(np.float32)) and acc.count == 1


def test_auto_reset_and_resolution_reset():
    acc = O.AccumulateOracle()
    fr = _frames(8, 2)
    acc.execute(fr[0])
    assert np.array_equal(acc.execute(fr[1], refresh=True), fr[1])
    acc.execute(fr[0])
    g = _frames(9, 1, h=3, w=3)[0]
    assert np.array_equal(acc.execute(g), g) and acc.count == 1
    manual = O.AccumulateOracle(auto_reset=False)
    manual.execute(fr[0])
    out = manual.execute(fr[1], refresh=True)
    np.testing.assert_allclose(out, (fr[0].astype(np.float64) + fr[1]) / 2, rtol=1e-6)


def test_enable_toggle_and_precision_change_reset():
    acc = O.AccumulateOracle()
    fr = _frames(10, 3)
    acc.execute(fr[0])
    acc.set_enabled(False)
    acc.set_enabled(True)
    assert acc.count == 0
    acc.execute(fr[1])
    acc.set_precision(O.DOUBLE)
    assert np.array_equal(acc.execute(fr[2]), fr[2])


# --------------------------------------------------------------------------- ImageCompare
def _img(seed, h=6, w=7):
    return np.random.default_rng(seed).random((h, w, 4)).astype(np.float32) + np.float32(0.25)


def test_mse_constant_offset():
    a = _img(0)
    d = np.float32(0.5)
    e, m = O.compare(a, a + d, "mse")
    assert e == pytest.approx(0.25, rel=1e-6)
    assert np.allclose(m, 0.25, rtol=1e-6)


def test_rmse_relative_error_closed_form():
    a = np.full((3, 4, 4), 2.0, np.float32)
    b = np.full((3, 4, 4), 3.0, np.float32)
    e, _ = O.compare(a, b, "rmse")
    assert e == pytest.approx(1.0 / (4.0 + 1e-3), rel=1e-12)


def test_mae_is_mean_squared_difference_like_the_reference():
    a, b = _img(1), _img(2)
    assert O.compare(a, b, "mae")[0] == O.compare(a, b, "mse")[0]


def test_mape_closed_form():
    a = np.full((2, 2, 4), 1.0, np.float32)
    b = np.full((2, 2, 4), 1.5, np.float32)
    e, _ = O.compare(a, b, "mape")
    assert e == pytest.approx(100.0 * 0.5 / 1.001, rel=1e-12)


def test_alpha_channel_enters_only_with_alpha():
    a = _img(3)
    b = a.copy()
    b[..., 3] += 1
    assert O.compare(a, b, "mse")[0] == 0.0
    assert O.compare(a, b, "mse", alpha=True)[0] == pytest.approx(0.25, rel=1e-6)


def test_heat_map_stops():
    err = np.array([[0.0, 0.375, 0.5, 1.0]], np.float32)
    hm = O.heat_map(err)
    assert np.array_equal(hm[0, 0], [0, 0, 1, 1])       # min: blue
    assert np.array_equal(hm[0, 1], [0, 1, 0.5, 1])     # between teal and green
    assert np.array_equal(hm[0, 2], [0, 1, 0, 1])       # middle: green
    assert np.array_equal(hm[0, 3], [1, 0, 0, 1])       # max: red
    flat = O.heat_map(np.full((2, 2), 3.0, np.float32))
    assert np.array_equal(flat[..., :3], np.broadcast_to([0, 0, 1], (2, 2, 3)))
