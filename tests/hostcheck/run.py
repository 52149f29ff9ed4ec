"""Runs the host check program on a scene (falcor_amd.scene.SceneData) and a list of steps.

Writes the scene description's arrays and the dLdI image in the program's named-array format,
runs it under AddressSanitizer and returns (returncode, stdout, stderr).
"""
from __future__ import annotations

import os
import struct
import subprocess
import tempfile
from pathlib import Path

import numpy as np

from . import build as _build


def write_arrays(path: Path, arrays: dict):
    with open(path, "wb") as f:
        f.write(b"FHC1" + struct.pack("<I", len(arrays)))
        for name, a in arrays.items():
            b = np.ascontiguousarray(a).tobytes()
            f.write(name.encode().ljust(16, b"\0") + struct.pack("<Q", len(b)) + b)


def scene_arrays(data, dLdI=None) -> dict:
    off = data.mesh_vertex_offsets
    if off is None:
        off = np.array([0, data.positions.shape[0]], np.uint32)
    return {
        "positions": np.asarray(data.positions, np.float32), "normals": np.asarray(data.normals, np.float32),
        "texcoords": np.asarray(data.texcoords, np.float32), "indices": np.asarray(data.indices, np.uint32),
        "tri_material": np.asarray(data.tri_material, np.uint32), "materials": np.asarray(data.materials, np.float32),
        "tex_desc": np.asarray(data.tex_desc, np.int32), "texels": np.asarray(data.texels, np.float32),
        "camera": np.asarray(data.camera_array(), np.float32), "mesh_offsets": np.asarray(off, np.uint32),
        "dLdI": np.zeros(3, np.float32) if dLdI is None else np.asarray(dLdI, np.float32),
    }


def run(data, steps: list[str], dLdI=None, timeout: float = 600.0):
    exe = _build.build()
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    with tempfile.TemporaryDirectory() as d:
        scene = Path(d) / "scene.bin"
        write_arrays(scene, scene_arrays(data, dLdI))
        (Path(d) / "steps.txt").write_text("\n".join(steps) + "\n")
        res = subprocess.run([str(exe), str(scene), str(Path(d) / "steps.txt")], capture_output=True, text=True,
                             timeout=timeout, env=env)
    return res.returncode, res.stdout, res.stderr
