"""BSDF adjoints of the WARDiffPathTracer kernels against central differences (CPU).

tests/hostcheck/bsdfcheck.cpp compiles the kernels' per-vertex material adjoint (matAdjoint: the
four-lobe StandardBSDF with transmission, and the reflection lobes) and BSDF evaluation for the
CPU and compares, over random StandardMaterial rows (reflection only, specular transmission,
diffuse transmission, both; thin and not), shading frames seen from either side and outgoing
directions on both sides, the gradient of every parameter of the 20-float layout the adjoint
writes (baseColor, metallic, roughness, IoR, transmissionColor, diffuseTransmission,
specularTransmission) with Richardson-extrapolated central differences of the evaluation, with
the lobe-selection probabilities held at the base point. Points where the evaluation is not
smooth (cosine cut-offs, total internal reflection, kinks of the transmission lobe) are detected
from the differences themselves and skipped. Tolerance: 2 % of the derivative plus an absolute
floor of 1e-4 (|f| + 1e-3).

A sign error in the relative-IoR chain rule, a missing square root derivative of the
transmission albedo, a wrong roughness-to-alpha factor, a dropped lobe in a mixing-weight
derivative or a wrong power of eta in the transmission term each fail thousands of the
comparisons (checked by planting them in a copy of the kernels' header).
"""
import re
import subprocess

import pytest

build = pytest.importorskip("tests.hostcheck.build")

_PARAMS = ["baseColor.r", "baseColor.g", "baseColor.b", "roughness", "metallic", "ior", "transmission.r",
           "transmission.g", "transmission.b", "specularTransmission", "diffuseTransmission"]


@pytest.mark.parametrize("seed", [1, 7])
def test_bsdf_adjoint_matches_central_differences(seed):
    exe = build.build_bsdfcheck()
    res = subprocess.run([str(exe), "20000", str(seed)], capture_output=True, text=True, timeout=300)
    assert "ERROR: AddressSanitizer" not in res.stderr and "runtime error" not in res.stderr, res.stderr[-4000:]
    assert res.returncode == 0, res.stdout[-4000:] + res.stderr[-2000:]
    m = re.search(r"compared=(\d+) nonzero=(\d+) skipped=(\d+) failed=(\d+)", res.stdout)
    assert m, res.stdout
    compared, nonzero, skipped, failed = map(int, m.groups())
    assert failed == 0 and skipped < 0.01 * compared, res.stdout
    # Every parameter is exercised away from zero on both the reflection and the transmission side.
    for name in _PARAMS:
        pm = re.search(rf"param {re.escape(name)} nonzero_reflection=(\d+) nonzero_transmission=(\d+)", res.stdout)
        assert pm, name
        assert int(pm.group(1)) >= 1000 and int(pm.group(2)) >= 1000, pm.group(0)
