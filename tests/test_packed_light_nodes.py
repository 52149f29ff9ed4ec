"""The precision of the light BVH's packed nodes (the reference's PackedNode,
LightBVHTypes.slang:97-198), restated in the oracle (oracle/fo_impl.cpp) and used by the product's
builder, refit and sampler (fhp_scene.h packLightNode / unpackLightNode).

* Half floats: the oracle's rounding (computed on the value in double precision) equals numpy's
  IEEE binary16 conversion, round to nearest even, bit for bit over normals, subnormals, ties and
  overflow; textbook values (1, 65504, 65520 -> inf, 2^-24, 2^-25 -> 0 by the tie rule).
* Cone cosines: (uint)((c + 1) * 32767) / 32767 - 1, never above c (the quantised cone is at
  least as wide), exact at -1 (no cone) and within 1 / 32767.
* Cone axes: octahedral 2 x 16-bit snorms (encodeNormal2x16): the axes and octant diagonals
  round trip to themselves, every unit vector within 1e-4 of itself and of unit length.
* The product's nodes (built on the host, refitted on the device) carry exactly these values: the
  decoded host mirror equals the oracle's nodes (tests/test_light_refit_oracle.py and
  tests/test_gpu_light_refit.py compare them after builds and refits).
"""
import numpy as np

import oracle


def test_half_precision_is_ieee_round_to_nearest_even():
    rng = np.random.default_rng(0)
    v = np.concatenate([
        rng.uniform(-1, 1, 20000) * np.exp2(rng.integers(-30, 17, 20000)),  # normals, subnormals, overflow
        np.float32([1.0, 65504.0, 65519.0, 65520.0, 2.0 ** -24, 2.0 ** -25, 3 * 2.0 ** -25, 2.0 ** -14,
                    1 + 2.0 ** -11, 1 + 3 * 2.0 ** -11, 0.1, -0.0, 0.0]),
        # exact ties between neighbouring halves in the normal and subnormal ranges
        (np.arange(1, 2049, dtype=np.float64) + 0.5).astype(np.float32) * np.float32(2.0 ** -24),
    ]).astype(np.float32)
    half, _, _ = oracle.packed_precision(v, np.tile(np.float32([0, 0, 1]), (v.size, 1)))
    with np.errstate(over="ignore"):
        ref = v.astype(np.float16).astype(np.float32)
    np.testing.assert_array_equal(half.view(np.uint32), ref.view(np.uint32))
    k = v.size - 2048 - 13
    assert half[k + 0] == 1.0 and half[k + 1] == 65504.0 and half[k + 2] == 65504.0 and np.isinf(half[k + 3])
    assert half[k + 4] == 2.0 ** -24 and half[k + 5] == 0.0 and half[k + 6] == 2.0 ** -23
    assert half[k + 8] == 1.0 and half[k + 9] == 1 + 2.0 ** -9


def test_cone_quantisation_widens():
    c = np.concatenate([np.linspace(-1, 1, 4001), np.float32([-1.0, 1.0, 0.5])]).astype(np.float32)
    _, q, _ = oracle.packed_precision(c, np.tile(np.float32([0, 0, 1]), (c.size, 1)))
    assert np.all(q <= c + 1e-7) and np.all(c - q <= 1.0 / 32767 + 1e-6)
    assert q[-3] == -1.0  # the invalid cone stays exactly -1
    assert q[-2] == np.float32(65534) / np.float32(32767) - 1


def test_octahedral_axis_round_trip():
    axes = np.float32([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]])
    _, _, o = oracle.packed_precision(np.zeros(6, np.float32), axes)
    np.testing.assert_allclose(o, axes, atol=1e-7)
    rng = np.random.default_rng(1)
    d = rng.normal(size=(50000, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    _, _, o = oracle.packed_precision(np.zeros(d.shape[0], np.float32), d)
    assert np.max(np.linalg.norm(o - d, axis=1)) < 1e-4
    np.testing.assert_allclose(np.linalg.norm(o, axis=1), 1.0, atol=1e-6)
    # a direction quantised twice stays where the first quantisation put it (up to normalisation)
    _, _, o2 = oracle.packed_precision(np.zeros(d.shape[0], np.float32), o)
    assert np.max(np.linalg.norm(o2 - o, axis=1)) < 2e-5
