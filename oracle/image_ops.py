"""CPU restatement of AccumulatePass and of ImageCompare's metrics (test infrastructure only).

numpy float32 / float64 element-wise operations are single IEEE operations with no fused
multiply-add, so the per-pixel arithmetic below follows the reference's operation order exactly
and is the bit-level oracle of the GPU kernels.

Sources:
  Accumulate.cs.slang:76-114 (Single), :118-153 (SingleCompensated), :156-190 (Double);
  AccumulatePass.cpp:127-218 (execute: auto reset, overflow modes, resolution reset, blit when
  disabled), :220-292 (accumulate: frame count update), :366-378 (setEnabled, reset),
  :380-420 (prepareAccumulation: buffers cleared when the count is 0).
  ImageCompare.cpp:191-233 (MSE, RMSE, MAE, MAPE), :235-252 (compare), :263-292 (heat map).
"""
from __future__ import annotations

import math

import numpy as np

DOUBLE, SINGLE, SINGLE_COMPENSATED = 0, 1, 2
STOP, RESET, EMA = 0, 1, 2


class AccumulateOracle:
    """AccumulatePass state machine over [H, W, 4] frames (numpy)."""

    def __init__(self, enabled=True, auto_reset=True, precision=SINGLE, max_frames=0, overflow=STOP):
        self.enabled, self.auto_reset = bool(enabled), bool(auto_reset)
        self.precision, self.max_frames, self.overflow = precision, int(max_frames), overflow
        self.count = 0
        self.dim = None
        self.mode = None
        self.sum = self.corr = None
        self.out = None

    def reset(self):
        self.count = 0

    def set_enabled(self, enabled):
        # AccumulatePass::setEnabled (:366-373)
        if bool(enabled) != self.enabled:
            self.enabled = bool(enabled)
            self.reset()

    def set_precision(self, precision):
        if precision != self.precision:
            self.precision = precision
            self.reset()

    def execute(self, frame: np.ndarray, refresh: bool = False) -> np.ndarray:
        """One AccumulatePass::execute; returns the output image (kept between calls)."""
        h, w = frame.shape[:2]
        if self.out is None or self.out.shape[:2] != (h, w):
            self.out = np.zeros((h, w, 4), np.float32)
        if self.auto_reset and refresh:
            self.reset()
        if self.max_frames > 0 and self.count == self.max_frames:
            if self.overflow == STOP:
                return self.out
            if self.overflow == RESET:
                self.reset()
        if self.dim != (h, w):
            self.dim = (h, w)
            self.reset()
        is_float = frame.dtype == np.float32
        if not self.enabled and is_float:
            self.out[...] = frame
            return self.out
        if self.mode != self.precision or self.sum is None or self.sum.shape[:2] != (h, w):
            # prepareAccumulation: the current mode's buffers are (re)created, which resets.
            self.mode = self.precision
            self.reset()
        if self.count == 0:
            dt = np.float64 if self.precision == DOUBLE else np.float32
            self.sum = np.zeros((h, w, 4), dt)
            self.corr = np.zeros((h, w, 4), np.float32)
        cur = frame.astype(np.float32)
        if self.enabled:
            self.out[...] = self._kernel(cur)
        else:
            self.out[...] = cur
        if self.max_frames == 0 or self.precision == SINGLE_COMPENSATED or self.count < self.max_frames:
            self.count += 1
        return self.out

    def _kernel(self, cur):
        n = self.count
        mavg = self.max_frames > 0
        if self.precision == SINGLE:
            wgt = np.float32(1.0) / np.float32(n + 1)
            if mavg:
                # lerp(x, y, s) = x * (1 - s) + y * s
                o = self.sum * (np.float32(1.0) - wgt) + cur * wgt
                self.sum = o
                return o
            self.sum = self.sum + cur
            return self.sum * wgt
        if self.precision == SINGLE_COMPENSATED:
            y = cur - self.corr
            sn = self.sum + y
            o = sn / np.float32(n + 1)
            self.corr = (sn - self.sum) - y
            self.sum = sn
            return o
        wgt = 1.0 / float(n + 1)
        c = cur.astype(np.float64)
        if mavg:
            self.sum = self.sum * (1.0 - wgt) + c * wgt
            return self.sum.astype(np.float32)
        self.sum = self.sum + c
        return (self.sum * wgt).astype(np.float32)


METRICS = ("mse", "rmse", "mae", "mape")


def pixel_errors(a: np.ndarray, b: np.ndarray, metric: str, alpha: bool = False) -> np.ndarray:
    """Per-pixel error (float64) of [H, W, 4] float32 images (ImageCompare.cpp:191-233): the mean
    over RGB (or RGBA) of each channel term, channels added in order."""
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    nch = 4 if alpha else 3
    e = np.zeros(a.shape[:2], np.float64)
    for k in range(nch):
        ak, bk = a[..., k], b[..., k]
        d = ak - bk  # float
        if metric == "mse":
            t = (d * d).astype(np.float64)
        elif metric == "rmse":
            t = (d * d).astype(np.float64) / ((ak * ak).astype(np.float64) + 1e-3)
        elif metric == "mae":
            t = np.abs(d * d).astype(np.float64)  # the reference's MAE squares the difference
        elif metric == "mape":
            t = np.abs(d.astype(np.float64) / (ak.astype(np.float64) + 1e-3))
        else:
            raise ValueError(metric)
        e = e + t
    if metric == "mape":
        return 100.0 * e / nch
    return e / nch


def compare(a, b, metric: str, alpha: bool = False):
    """(error, error map): the mean of the per-pixel errors, summed exactly (math.fsum), and the
    per-pixel errors as float32 (compare, ImageCompare.cpp:235-252)."""
    e = pixel_errors(a, b, metric, alpha)
    return math.fsum(e.ravel().tolist()) / e.size, e.astype(np.float32)


_RAMP = np.array([[0, 0, 1], [0, 1, 1], [0, 1, 0], [1, 1, 0], [1, 0, 0]], np.float32)


def heat_map(err: np.ndarray) -> np.ndarray:
    """generateHeatMap (ImageCompare.cpp:263-292) for an error map without NaNs."""
    err = np.asarray(err, np.float32)
    mn, mx = err.min(), err.max()
    rng = max(np.float32(1e-5), np.float32(mx - mn))
    x = (err - mn) / rng
    t = np.where(np.float32(0) < np.where(x < 1, x, np.float32(1)), np.where(x < 1, x, np.float32(1)), np.float32(0))
    t = t.astype(np.float32)
    c = np.clip(np.floor(t * np.float32(4)).astype(np.int32), 0, 3)
    u = t * np.float32(4) - c.astype(np.float32)
    c0, c1 = _RAMP[c], _RAMP[c + 1]
    out = np.ones(err.shape + (4,), np.float32)
    out[..., :3] = c0 + u[..., None] * (c1 - c0)
    return out
