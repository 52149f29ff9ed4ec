"""Chi-square goodness-of-fit of the oracle's BSDF sampling and the sample-weight check, as the
reference's BSDF tests run them (Source/Tools/FalcorTest/Tests/Scene/Material/BSDFTests.cpp:46-47,
:255-289, BSDFTests.cs.slang:140-185):

* sampled directions are binned over (cos theta, phi) (65 x 130 bins); the expected count of each
  bin is the sample count times the integral of evalPdf over the bin (8 x 8 midpoint rule);
  bins expecting fewer than 5 samples are pooled; the test passes when the p-value exceeds the
  Sidak-corrected significance level 0.01 (hypothesis::chi2_test);
* for single-lobe configurations (one lobe has all the selection probability) the weight
  returned by sample() equals eval() / evalPdf() within 1e-3 relative, and the returned pdf equals
  evalPdf() within 1e-3.
A lobe is isolated by zeroing the others' selection weight: Frostbite or Lambert transmission with
IoR 1 at normal incidence (no specular albedo), GGX reflection with metallic 1, the dielectric
BSDF with specular transmission 1 and a white base colour.
"""
import math
import zlib

import numpy as np
import pytest
from scipy.stats import chi2

N_COS, N_PHI = 64 + 1, 128 + 2
SAMPLES = 1 << 21


def _mat(base=(1, 1, 1), rough=0.5, metal=0.0, ior=1.5, eta=None, spec_t=0.0, diff_t=0.0, trans=(1, 1, 1)):
    eta = 1.0 / ior if eta is None else eta
    return np.float32([0, *base, rough, metal, ior, eta, spec_t, diff_t, *trans])


def _unit(v):
    v = np.asarray(v, np.float64)
    return (v / np.linalg.norm(v)).astype(np.float32)


CONFIGS = [
    ("frostbite", _mat(base=(0.8, 0.6, 0.4), rough=0.5, ior=1.0), _unit([0, 0, 1]), True),
    ("lambert_btdf", _mat(base=(0.8, 0.6, 0.4), ior=1.0, diff_t=1.0, trans=(0.9, 0.5, 0.7)), _unit([0, 0, 1]), True),
    ("ggx_reflection", _mat(base=(0.9, 0.7, 0.5), rough=0.5, metal=1.0), _unit([0.4, -0.3, 0.6]), True),
    ("dielectric_bsdf", _mat(rough=0.4, spec_t=1.0, trans=(0.9, 0.8, 0.7)), _unit([0.3, 0.2, 0.7]), True),
    ("dielectric_bsdf_inside", _mat(rough=0.3, eta=1.5, spec_t=1.0), _unit([0.1, -0.2, 0.9]), True),
    ("mixture", _mat(base=(0.6, 0.5, 0.4), rough=0.6, metal=0.2, spec_t=0.3, diff_t=0.4, trans=(0.8, 0.9, 1.0)),
     _unit([0.5, 0.1, 0.6]), False),
]


def _bin_index(w):
    cb = np.clip(np.floor((w[:, 2].astype(np.float64) * 0.5 + 0.5) * N_COS), 0, N_COS - 1).astype(np.int64)
    ph = np.arctan2(w[:, 1], w[:, 0]).astype(np.float64) / (2 * math.pi)
    ph = np.where(ph < 0, ph + 1, ph)
    pb = np.clip(np.floor(ph * N_PHI), 0, N_PHI - 1).astype(np.int64)
    return cb * N_PHI + pb


def _expected(O, mat, wi, sub=8):
    """Integral of evalPdf over each (cos theta, phi) bin: sub x sub midpoint rule per bin."""
    c = -1 + (np.arange(N_COS * sub) + 0.5) * (2.0 / (N_COS * sub))
    p = (np.arange(N_PHI * sub) + 0.5) * (2 * math.pi / (N_PHI * sub))
    C, P = np.meshgrid(c, p, indexing="ij")
    s = np.sqrt(np.maximum(0.0, 1 - C * C))
    dirs = np.stack([s * np.cos(P), s * np.sin(P), C], -1).reshape(-1, 3).astype(np.float32)
    pdf = O.bsdf_eval(mat, np.tile(wi, (dirs.shape[0], 1)), dirs)[:, 3].astype(np.float64)
    dA = (2.0 / (N_COS * sub)) * (2 * math.pi / (N_PHI * sub))
    return (pdf * dA).reshape(N_COS, sub, N_PHI, sub).sum(axis=(1, 3)).reshape(-1)


def _chi2_test(obs, exp, n, min_exp=5.0, alpha=0.01, tests=len(CONFIGS)):
    exp = exp * n
    order = np.argsort(exp)
    pooled_o = pooled_e = 0.0
    stat, dof = 0.0, -1
    for i in order:
        if exp[i] < min_exp:
            pooled_o += obs[i]
            pooled_e += exp[i]
            continue
        stat += (obs[i] - exp[i]) ** 2 / exp[i]
        dof += 1
    if pooled_e >= min_exp:
        stat += (pooled_o - pooled_e) ** 2 / pooled_e
        dof += 1
    p = chi2.sf(stat, dof)
    level = 1 - (1 - alpha) ** (1.0 / tests)  # Sidak correction
    return p, level, dof


@pytest.mark.parametrize("name,mat,wi,single", CONFIGS, ids=[c[0] for c in CONFIGS])
def test_bsdf_sampling_chi2_and_weights(oracle_mod, name, mat, wi, single):
    O = oracle_mod.Oracle
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    s = O.bsdf_sample(mat, np.tile(wi, (SAMPLES, 1)), rng.integers(0, 2 ** 31, (SAMPLES, 2), dtype=np.uint32))
    ok = s[:, 7] == 1
    assert ok.mean() > 0.5
    obs = np.bincount(_bin_index(s[ok, :3]), minlength=N_COS * N_PHI).astype(np.float64)
    exp = _expected(O, mat, wi)
    # The pdf mass over the sphere equals the fraction of valid samples.
    assert abs(exp.sum() - ok.mean()) < 5e-3, (exp.sum(), ok.mean())
    p, level, dof = _chi2_test(obs, exp, SAMPLES)
    assert dof > 100
    assert p > level, (name, p, level)
    # Returned pdf vs evalPdf, and for one lobe the weight vs eval / evalPdf.
    m = min(int(ok.sum()), 200000)
    wo = s[ok, :3][:m]
    e = O.bsdf_eval(mat, np.tile(wi, (m, 1)), wo)
    pdf_ref = e[:, 3].astype(np.float64)
    pdf_err = np.abs(s[ok, 6][:m] - pdf_ref) / np.where(pdf_ref > 0, pdf_ref, 1.0)
    assert pdf_err.max() <= 1e-3, pdf_err.max()
    if single:
        w_ref = np.where(pdf_ref[:, None] > 0, e[:, :3] / np.maximum(pdf_ref[:, None], 1e-30), 0.0)
        w_err = np.abs(s[ok, 3:6][:m] - w_ref) / np.where(w_ref > 0, w_ref, 1.0)
        assert w_err.max() <= 1e-3, w_err.max()
