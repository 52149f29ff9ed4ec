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
BSDF_EXE = OUT_DIR / "bsdfcheck"


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def _compile(srcs, exe: Path, force: bool) -> Path:
    deps = list(srcs) + sorted(CSRC.glob("*.h")) + [ROOT / "include" / "fhp.h"]
    if exe.exists() and not force and all(d.stat().st_mtime <= exe.stat().st_mtime for d in deps):
        return exe
    OUT_DIR.mkdir(exist_ok=True)
    # Host-only compilation: no device code is generated, and the sanitizer applies to the host
    # code (-Xarch_host), the only code in this program.
    cmd = [hipcc(), "-x", "hip", "--offload-host-only", "-DFHP_HOST_CHECK", "-std=c++17", "-O1", "-g",
           "-fno-omit-frame-pointer", "-ffp-contract=off", "-Xarch_host", "-fsanitize=address",
           "-Xarch_host", "-fsanitize=undefined", "-Wno-unused-command-line-argument",
           *map(str, srcs), "-o", str(exe)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"host check build failed:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return exe


def build(force: bool = False) -> Path:
    return _compile([HERE / "hostcheck.cpp", CSRC / "scene_build.cpp"], EXE, force)


def build_bsdfcheck(force: bool = False) -> Path:
    """The BSDF adjoint check (bsdfcheck.cpp): adjoints against central differences of the evaluation."""
    return _compile([HERE / "bsdfcheck.cpp"], BSDF_EXE, force)


if __name__ == "__main__":
    print(build(force=True))
    print(build_bsdfcheck(force=True))
