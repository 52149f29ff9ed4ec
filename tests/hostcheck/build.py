"""Builds the host check program (tests/hostcheck/hostcheck.cpp) with AddressSanitizer.

Test infrastructure: the kernels' per-path code compiled for the CPU (FHP_HOST_CHECK) together
with the library's host scene builder, so every array access of a path is bounds checked.
"""
from __future__ import annotations

import os
import shutil
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
CSRC = ROOT / "falcor_amd" / "csrc"
OUT_DIR = HERE / "_build"
EXE = OUT_DIR / "hostcheck"


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def build(force: bool = False) -> Path:
    srcs = [HERE / "hostcheck.cpp", CSRC / "scene_build.cpp"]
    deps = srcs + sorted(CSRC.glob("*.h")) + [ROOT / "include" / "fhp.h"]
    if EXE.exists() and not force and all(d.stat().st_mtime <= EXE.stat().st_mtime for d in deps):
        return EXE
    OUT_DIR.mkdir(exist_ok=True)
    # Host-only compilation: no device code is generated, and the sanitizer applies to the host
    # code (-Xarch_host), the only code in this program.
    cmd = [hipcc(), "-x", "hip", "--offload-host-only", "-DFHP_HOST_CHECK", "-std=c++17", "-O1", "-g",
           "-fno-omit-frame-pointer", "-ffp-contract=off", "-Xarch_host", "-fsanitize=address",
           "-Xarch_host", "-fsanitize=undefined", "-Wno-unused-command-line-argument",
           *map(str, srcs), "-o", str(EXE)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"host check build failed:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return EXE


if __name__ == "__main__":
    print(build(force=True))
