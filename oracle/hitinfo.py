"""CPU restatement of PackedHitInfo, VBufferRT's camera jitter patterns and depth (test
infrastructure only).

Sources:
  Source/Falcor/Scene/HitInfo.cpp:37-106 (bit allocation, compression rule),
  Source/Falcor/Scene/HitInfo.slang:94-121 (TriangleHit pack / unpack), :300-312 (offsets),
  :457-487 (packHeader / unpackHeader), Source/Falcor/Scene/HitInfoType.slang:33-49 (HitType),
  Source/Falcor/Utils/Math/FormatConversion.slang:292-341 (unorm16),
  Source/Falcor/Utils/SampleGenerators/HaltonSamplePattern.cpp:33-68 (Halton jitter),
  Source/Falcor/Scene/Camera/Camera.slang:46-78 (jittered pinhole ray),
  Source/Falcor/Utils/Math/MatrixMath.h:553-566, :699-725 (projection and look-at: depth).
"""
from __future__ import annotations

import numpy as np

HIT_NONE, HIT_TRIANGLE = 0, 1
HIT_TYPE_COUNT = 7  # None, Triangle, Volume, DisplacedTriangle, Curve, (5 unused), SDFGrid
COMPRESSED_TYPE_BITS = 2


def allocate_bits(count: int) -> int:
    """Bits needed to store the values 0 .. count - 1."""
    return 0 if count <= 1 else int(count - 1).bit_length()


def layout(num_instances: int, max_primitives: int, use_compression: bool) -> dict:
    type_bits = allocate_bits(HIT_TYPE_COUNT)
    inst_bits = allocate_bits(num_instances)
    prim_bits = allocate_bits(max_primitives)
    if prim_bits > 32 or type_bits + inst_bits > 32:
        raise ValueError("hit info header needs more than 64 bits")
    compressed = use_compression and COMPRESSED_TYPE_BITS + inst_bits + prim_bits <= 32
    if compressed:
        type_bits = COMPRESSED_TYPE_BITS
    return {"compressed": compressed, "type_bits": type_bits, "inst_bits": inst_bits, "prim_bits": prim_bits}


def pack_unorm16(v) -> np.ndarray:
    """trunc(v * 65535 + 0.5) in float32."""
    v = np.asarray(v, np.float32)
    return np.trunc(v * np.float32(65535.0) + np.float32(0.5)).astype(np.uint32)


def unpack_unorm16(u) -> np.ndarray:
    return (np.asarray(u, np.uint32) & 0xFFFF).astype(np.float32) * np.float32(1.0 / 65535.0)


def pack_triangle_hits(L: dict, inst, prim, b1, b2, valid) -> np.ndarray:
    """[N, 2] (compressed) or [N, 4] uint32 records; rows with valid False are all zero (miss)."""
    inst = np.asarray(inst, np.uint64)
    prim = np.asarray(prim, np.uint64)
    n = inst.shape[0]
    words = 2 if L["compressed"] else 4
    out = np.zeros((n, 4), np.uint64)
    type_word = np.uint64(HIT_TRIANGLE) << np.uint64(32 - L["type_bits"])
    if L["type_bits"] + L["inst_bits"] + L["prim_bits"] <= 32:
        out[:, 0] = type_word | (inst << np.uint64(L["prim_bits"])) | prim
    else:
        out[:, 0] = type_word | inst
        out[:, 1] = prim
    if L["compressed"]:
        out[:, 1] = (pack_unorm16(b2).astype(np.uint64) << np.uint64(16)) | pack_unorm16(b1).astype(np.uint64)
    else:
        out[:, 2] = np.asarray(b1, np.float32).view(np.uint32)
        out[:, 3] = np.asarray(b2, np.float32).view(np.uint32)
    out[~np.asarray(valid, bool)] = 0
    return out[:, :words].astype(np.uint32)


def unpack_triangle_hits(L: dict, words: np.ndarray):
    """-> (type, inst, prim, b1, b2) arrays."""
    w = np.asarray(words, np.uint32).astype(np.uint64)
    typ = w[:, 0] >> np.uint64(32 - L["type_bits"])
    if L["type_bits"] + L["inst_bits"] + L["prim_bits"] <= 32:
        inst = (w[:, 0] >> np.uint64(L["prim_bits"])) & np.uint64((1 << L["inst_bits"]) - 1)
        prim = w[:, 0] & np.uint64((1 << L["prim_bits"]) - 1)
    else:
        inst = w[:, 0] & np.uint64((1 << L["inst_bits"]) - 1)
        prim = w[:, 1]
    if L["compressed"]:
        b1 = unpack_unorm16(w[:, 1].astype(np.uint32))
        b2 = unpack_unorm16((w[:, 1] >> np.uint64(16)).astype(np.uint32))
    else:
        b1 = w[:, 2].astype(np.uint32).view(np.float32)
        b2 = w[:, 3].astype(np.uint32).view(np.float32)
    return typ.astype(np.uint32), inst.astype(np.uint32), prim.astype(np.uint32), b1, b2


def halton_jitter(index: int) -> tuple[float, float]:
    """HaltonSamplePattern::next for sample `index`: radical inverses in bases 2 and 3 (float32),
    mapped by frac(v + 0.5) - 0.5."""
    def radical(i, base):
        r, f = np.float32(0), np.float32(1)
        while i > 0:
            f = np.float32(f / np.float32(base))
            r = np.float32(r + np.float32(f * np.float32(i % base)))
            i //= base
        return r
    out = []
    for base in (2, 3):
        v = np.float32(radical(index, base) + np.float32(0.5))
        out.append(float(np.float32(np.float32(v - np.floor(v)) - np.float32(0.5))))
    return out[0], out[1]


def camera_dirs(cam: np.ndarray, W: int, H: int, jitter=(0.0, 0.0)) -> np.ndarray:
    """Normalized pinhole ray directions [H, W, 3] in float64 (Camera.slang:46-78), with the
    camera jitter (jitterX, jitterY) in screen units."""
    cam = np.asarray(cam, np.float64)
    pos, target, up = cam[0:3], cam[3:6], cam[6:9]
    fovy, aspect, focal = cam[9], cam[10], cam[11]
    w = target - pos
    w = w / np.linalg.norm(w) * focal
    u = np.cross(w, up)
    u = u / np.linalg.norm(u)
    v = np.cross(u, w)
    v = v / np.linalg.norm(v)
    t = np.tan(fovy / 2)
    u = u * focal * t * aspect
    v = v * focal * t
    px = (np.arange(W) + 0.5) / W - jitter[0]
    py = (np.arange(H) + 0.5) / H + jitter[1]
    nx = 2 * px[None, :] - 1
    ny = 1 - 2 * py[:, None]
    d = nx[..., None] * u + ny[..., None] * v + w
    return d / np.linalg.norm(d, axis=-1, keepdims=True)


def ndc_depth(cam: np.ndarray, points: np.ndarray) -> np.ndarray:
    """z / w of viewProj * (p, 1) with the right-handed look-at and [0, 1]-depth perspective
    projection, in float64: far (d - near) / ((far - near) d), d the distance along the view axis."""
    cam = np.asarray(cam, np.float64)
    pos, target = cam[0:3], cam[3:6]
    near, far = cam[12], cam[13]
    f = (target - pos) / np.linalg.norm(target - pos)
    d = (np.asarray(points, np.float64) - pos) @ f
    return far * (d - near) / ((far - near) * d)
