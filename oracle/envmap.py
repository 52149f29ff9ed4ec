"""CPU restatement of the environment map lookup (test infrastructure only).

Sources: Source/Falcor/Scene/Lights/EnvMap.slang:58-112 (eval, worldToUv, toLocal, intensity *
tint), Source/Falcor/Utils/Math/MathHelpers.slang:111-118 (world_to_latlong_map),
Source/Falcor/Scene/Lights/EnvMap.cpp:62-72, :131-135 (rotation transform, bilinear sampler
wrapping in u and clamping in v), Source/Falcor/Utils/Math/MatrixMath.h:649-679
(matrixFromRotationXYZ). Evaluated in float64.
"""
from __future__ import annotations

import numpy as np


def rotation_xyz(deg) -> np.ndarray:
    """matrixFromRotationXYZ of the radian angles (which negates them), as a 3x3 array."""
    x, y, z = (-np.radians(np.asarray(deg, np.float64))).tolist()
    c1, c2, c3 = np.cos(x), np.cos(y), np.cos(z)
    s1, s2, s3 = np.sin(x), np.sin(y), np.sin(z)
    return np.array([[c2 * c3, c2 * s3, -s2],
                     [-c1 * s3 + s1 * s2 * c3, c1 * c3 + s1 * s2 * s3, s1 * c2],
                     [s1 * s3 + c1 * s2 * c3, -s1 * c3 + c1 * s2 * s3, c1 * c2]])


def latlong_uv(dirs) -> np.ndarray:
    p = np.asarray(dirs, np.float64)
    p = p / np.linalg.norm(p, axis=-1, keepdims=True)
    u = np.arctan2(p[..., 0], -p[..., 2]) / (2 * np.pi) + 0.5
    v = np.arccos(np.clip(p[..., 1], -1, 1)) / np.pi
    return np.stack([u, v], -1)


def bilinear(tex, uv) -> np.ndarray:
    """Bilinear lookup of an [H, W, 3] map at mip 0, wrapping in u and clamping in v."""
    tex = np.asarray(tex, np.float64)
    H, W = tex.shape[:2]
    x = uv[..., 0] * W - 0.5
    y = uv[..., 1] * H - 0.5
    x0, y0 = np.floor(x), np.floor(y)
    fx, fy = (x - x0)[..., None], (y - y0)[..., None]
    x0, y0 = x0.astype(np.int64), y0.astype(np.int64)
    xa, xb = np.mod(x0, W), np.mod(x0 + 1, W)
    ya, yb = np.clip(y0, 0, H - 1), np.clip(y0 + 1, 0, H - 1)
    top = tex[ya, xa] * (1 - fx) + tex[ya, xb] * fx
    bot = tex[yb, xa] * (1 - fx) + tex[yb, xb] * fx
    return top * (1 - fy) + bot * fy


def env_eval(tex, dirs, intensity=1.0, tint=(1.0, 1.0, 1.0), rotation=(0.0, 0.0, 0.0)) -> np.ndarray:
    """EnvMap::eval(dir) for world-space directions [..., 3]."""
    M = rotation_xyz(rotation)
    local = np.asarray(dirs, np.float64) @ M  # inverse(M) @ d = M^T d for a rotation
    return bilinear(tex, latlong_uv(local)) * intensity * np.asarray(tint, np.float64)
