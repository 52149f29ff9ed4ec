"""Pins the PackedHitInfo / jitter / depth oracle (oracle/hitinfo.py) with hand-built records,
closed forms and invariants, and checks VBufferRT's property handling (no GPU needed)."""
import numpy as np
import pytest

from oracle import hitinfo as O


def test_allocate_bits():
    # allocateBits (HitInfo.cpp:44-49): bitScanReverse(count - 1) + 1, 0 for count <= 1.
    assert [O.allocate_bits(c) for c in (0, 1, 2, 3, 4, 5, 7, 8, 9, 12, 16, 17)] == [0, 0, 1, 2, 2, 3, 3, 3, 4, 4, 4, 5]
    assert O.allocate_bits(1 << 20) == 20 and O.allocate_bits((1 << 20) + 1) == 21


def test_layouts():
    # A Cornell-box-sized scene (9 instances, <= 12 triangles each): 3 type bits uncompressed,
    # 2 + 4 + 4 = 10 header bits compressed.
    assert O.layout(9, 12, False) == {"compressed": False, "type_bits": 3, "inst_bits": 4, "prim_bits": 4}
    assert O.layout(9, 12, True) == {"compressed": True, "type_bits": 2, "inst_bits": 4, "prim_bits": 4}
    # Too large for 32 header bits: compression refused, two-word header.
    L = O.layout(1 << 12, 1 << 20, True)
    assert not L["compressed"] and L["type_bits"] + L["inst_bits"] + L["prim_bits"] == 3 + 12 + 20
    with pytest.raises(ValueError):
        O.layout(1 << 30, 4, False)


def test_compressed_record_by_hand():
    L = O.layout(9, 12, True)
    w = O.pack_triangle_hits(L, [5], [9], [0.25], [0.5], [True])
    # type 1 in bits 30-31, instance 5 above 4 primitive bits, primitive 9;
    # unorm16(0.25) = trunc(16384.25) = 0x4000, unorm16(0.5) = trunc(32768.0) = 0x8000.
    assert w.tolist() == [[0x40000059, 0x80004000]]
    typ, inst, prim, b1, b2 = O.unpack_triangle_hits(L, w)
    assert (typ[0], inst[0], prim[0]) == (1, 5, 9)
    assert b1[0] == np.float32(16384) * np.float32(1 / 65535) and b2[0] == np.float32(32768) * np.float32(1 / 65535)


def test_uncompressed_records_by_hand():
    L = O.layout(9, 12, False)
    w = O.pack_triangle_hits(L, [3, 0], [11, 0], [0.125, 0], [0.75, 0], [True, False])
    assert w[0].tolist() == [(1 << 29) | (3 << 4) | 11, 0, np.float32(0.125).view(np.uint32), np.float32(0.75).view(np.uint32)]
    assert w[1].tolist() == [0, 0, 0, 0]  # a miss is all zeros (HitType::None)
    L2 = O.layout(1 << 12, 1 << 20, False)
    w2 = O.pack_triangle_hits(L2, [4095], [1000000], [0.5], [0.25], [True])
    assert w2[0, 0] == (1 << 29) | 4095 and w2[0, 1] == 1000000
    typ, inst, prim, _, _ = O.unpack_triangle_hits(L2, w2)
    assert (typ[0], inst[0], prim[0]) == (1, 4095, 1000000)


def test_unorm16_round_trip_error():
    v = np.random.default_rng(0).random(10000).astype(np.float32)
    back = O.unpack_unorm16(O.pack_unorm16(v))
    assert np.max(np.abs(back - v)) <= 0.5 / 65535 + 1e-7
    assert O.pack_unorm16([0.0, 1.0]).tolist() == [0, 65535]


def test_halton_jitter_closed_forms():
    # Radical inverses: index 1 -> (1/2, 1/3), 2 -> (1/4, 2/3), 3 -> (3/4, 1/9), mapped by
    # frac(v + 0.5) - 0.5: 1/2 -> -1/2, 1/3 -> 1/3, 1/4 -> 1/4, 2/3 -> -1/3, 3/4 -> -1/4, 1/9 -> 1/9.
    assert O.halton_jitter(0) == (0.0, 0.0)
    want = {1: (-0.5, 1 / 3), 2: (0.25, -1 / 3), 3: (-0.25, 1 / 9)}
    for i, (wx, wy) in want.items():
        x, y = O.halton_jitter(i)
        assert x == pytest.approx(wx, abs=1e-7) and y == pytest.approx(wy, abs=1e-6), i


def test_depth_closed_form():
    cam = np.array([0, 0, 0, 0, 0, -1, 0, 1, 0, 1.0, 1.0, 1.0, 0.1, 100.0])
    d = O.ndc_depth(cam, [[0, 0, -0.1], [0, 0, -100], [3, 1, -10]])
    assert d[0] == pytest.approx(0.0, abs=1e-12) and d[1] == pytest.approx(1.0, abs=1e-12)
    assert d[2] == pytest.approx(100 * (10 - 0.1) / (99.9 * 10), rel=1e-12)


def test_camera_dirs_center_and_jitter():
    cam = np.array([0, 0, 0, 0, 0, -1, 0, 1, 0, np.pi / 2, 2.0, 1.0, 0.1, 100.0])
    d = O.camera_dirs(cam, 2, 2)
    # fovY 90 deg, aspect 2: pixel (0, 0) centre is at ndc (-0.5, 0.5) -> (-1, 0.5, -1) normalized.
    np.testing.assert_allclose(d[0, 0], np.array([-1, 0.5, -1]) / 1.5, atol=1e-12)
    # A jitter of +1/4 screen in x moves the sample left (p.x -= jitterX).
    dj = O.camera_dirs(cam, 2, 2, jitter=(0.25, 0.0))
    np.testing.assert_allclose(dj[0, 1], np.array([0, 0.5, -1]) / np.linalg.norm([0, 0.5, -1]), atol=1e-12)


def test_vbuffer_pass_properties():
    import ctypes as C
    from falcor_amd import _lib
    from falcor_amd.render_graph import VBufferRT
    p = VBufferRT({"samplePattern": "DirectX", "sampleCount": 3})
    assert p.get_property("sampleCount") == 8  # DxSamplePattern has 8 samples
    p.set_property("samplePattern", "Stratified")
    p.set_property("sampleCount", 5000)
    assert p.get_property("sampleCount") == 1024  # clamped (StratifiedSamplePattern.cpp:37-42)
    with pytest.raises(_lib.FhpError, match="samplePattern"):
        p.set_property("samplePattern", 7)
    with pytest.raises(_lib.FhpError, match="cull"):
        p.set_property("forceCullMode", True)  # with the default cull mode Back
    p.set_property("cull", "None")
    p.set_property("forceCullMode", True)
    with pytest.raises(_lib.FhpError, match="cull"):
        p.set_property("cull", "Back")
    lib = _lib.load()
    assert lib.fhp_vbuffer_execute(p.handle, None, 4, 4, C.c_void_p(16), None, None, None, None, None) == -1
    assert lib.fhp_pass_set_input(p.handle, b"vbuffer", None) == -1
    from falcor_amd.render_graph import MinimalPathTracer
    m = MinimalPathTracer({})
    assert lib.fhp_pass_set_input(m.handle, b"nope", None) == -1
    assert "unknown input channel" in lib.fhp_last_error().decode()
