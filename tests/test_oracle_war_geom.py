"""Pins of the oracle's restatement of WARDiffPathTracer's vertex-gradient estimator
(oracle/fo_war_geom.cpp: the path radiance as a function of vertex positions and normals with
every sampled decision held, differentiated by central differences in double precision).

* At theta = 0 the restatement reproduces the oracle's WAR primal: the loss it accumulates
  equals sum dLdI . (spp x image) of render_war to 1e-6 relative, so no contribution of the path
  is dropped or double counted.
* Vertex normals enter only the shading frame (no silhouette moves), so the estimator's normal
  gradients are unbiased: against central differences of the primal image (common random
  numbers, 1024 spp, eps 1e-2) on two normal perturbation fields, within 10 % (the Monte Carlo
  noise of both sides at these sample counts is a few percent).
* With the sampled NEE direction held (the useSolidAngleNEEWarp extension) the warped-area
  estimator is consistent: its position gradients along three translations and a uniform scale
  match central differences of the primal (1024 spp, eps 4e-3) to a relative L2 error of 0.2 at
  64 auxiliary rays and vMF concentration 10^4 (the estimator's bias at finite concentration
  plus noise; a wrong sign or a dropped warp term gives errors near or above 1).
* The reference's own NEE formulation (PTUtils.slang:211-249: the direction recomputed toward
  the held light point and the area factor a2s / detach(a2s)) differs from the held-direction
  one only in the NEE term: without NEE the two give the same derivatives; with NEE the
  reference's estimate of the vertical translation is about twice the finite difference in this
  scene (recorded here, so a change of either formulation shows up).
"""
import numpy as np
import pytest

from falcor_amd import scene as S

W = H = 32


@pytest.fixture(scope="module")
def box(oracle_mod):
    data = S.displaced_sphere_in_box(0, radius=0.12, segments=(24, 12), sigma=0.05)
    mesh = data.mesh_names.index("Sphere")
    off = data.mesh_vertex_offsets
    return data, int(off[mesh]), int(off[mesh + 1]), oracle_mod.OracleScene(data)


def _ramp():
    x = (np.arange(W) + 0.5) / W
    return np.repeat(np.broadcast_to(x[None, :], (H, W))[..., None], 3, axis=2).astype(np.float32)


def _position_probes(data, v0, v1, nV):
    pr = np.zeros((4, nV, 6), np.float32)
    for a in range(3):
        pr[a, v0:v1, a] = 1.0
    pr[3, v0:v1, 0:3] = data.positions[v0:v1] - data.positions[v0:v1].mean(0)
    return pr


def _fd(osc, data, v0, v1, field, kind, eps, spp, nee=True):
    p0, n0 = data.positions.copy(), data.normals.copy()
    vals = []
    for sg in (1, -1):
        p, n = p0.copy(), n0.copy()
        if kind == "pos":
            p[v0:v1] += sg * eps * field
        else:
            n[v0:v1] += sg * eps * field
        osc.animate(p, n)
        im, _ = osc.render_war(W, H, spp=spp, frame=0, max_bounces=0, nee=nee)
        vals.append(float((im.astype(np.float64) * _ramp()).sum()))
    osc.animate(p0, n0)
    return (vals[0] - vals[1]) / (2 * eps)


def test_restatement_reproduces_the_primal(box):
    data, v0, v1, osc = box
    spp = 8
    pr = _position_probes(data, v0, v1, osc.num_vertices)[:1]
    _, val = osc.war_geom_probe(W, H, spp, 40, _ramp(), pr, aux=16, log10_kappa=3, log10_kappa_screen=3)
    img, _ = osc.render_war(W, H, spp=spp, frame=40, max_bounces=0)
    ref = float((img.astype(np.float64) * _ramp()).sum() * spp)
    assert val == pytest.approx(ref, rel=1e-6)


def test_normal_gradients_match_fd(box):
    data, v0, v1, osc = box
    rng = np.random.default_rng(3)
    fields = [rng.normal(size=(v1 - v0, 3)).astype(np.float32), np.tile(np.float32([1, 0, 0]), (v1 - v0, 1))]
    pr = np.zeros((2, osc.num_vertices, 6), np.float32)
    for k, f in enumerate(fields):
        pr[k, v0:v1, 3:6] = f
    spp = 1024
    d, _ = osc.war_geom_probe(W, H, spp, 0, _ramp(), pr, aux=16, log10_kappa=3, log10_kappa_screen=3)
    for k, f in enumerate(fields):
        fd = _fd(osc, data, v0, v1, f, "nrm", 1e-2, 1024)
        est = d[k] / spp
        print(f"normal field {k}: estimator {est:.5f}  fd {fd:.5f}", flush=True)
        assert abs(est - fd) <= 0.1 * abs(fd)


def test_held_nee_position_gradients_match_fd(box):
    data, v0, v1, osc = box
    pr = _position_probes(data, v0, v1, osc.num_vertices)
    spp = 512
    d, _ = osc.war_geom_probe(W, H, spp, 100 * spp, _ramp(), pr, aux=64, log10_kappa=4, log10_kappa_screen=4,
                              nee_direction_held=True)
    est = d / spp
    fd = np.array([_fd(osc, data, v0, v1, pr[k, v0:v1, 0:3], "pos", 4e-3, 1024) for k in range(4)])
    rel = float(np.linalg.norm(est - fd) / np.linalg.norm(fd))
    print(f"held NEE: estimator {np.round(est, 3)}  fd {np.round(fd, 3)}  rel L2 {rel:.3f}", flush=True)
    assert rel <= 0.2


def test_reference_nee_differs_from_held_only_through_nee(box):
    data, v0, v1, osc = box
    pr = _position_probes(data, v0, v1, osc.num_vertices)
    spp = 64
    kw = dict(aux=16, log10_kappa=3, log10_kappa_screen=3)
    a, _ = osc.war_geom_probe(W, H, spp, 7 * spp, _ramp(), pr, nee=False, **kw)
    b, _ = osc.war_geom_probe(W, H, spp, 7 * spp, _ramp(), pr, nee=False, nee_direction_held=True, **kw)
    np.testing.assert_array_equal(a, b)
    spp = 256
    ref, _ = osc.war_geom_probe(W, H, spp, 100 * spp, _ramp(), pr, aux=64, log10_kappa=4, log10_kappa_screen=4)
    held, _ = osc.war_geom_probe(W, H, spp, 100 * spp, _ramp(), pr, aux=64, log10_kappa=4, log10_kappa_screen=4,
                                 nee_direction_held=True)
    print(f"vertical translation: reference NEE {ref[1] / spp:.2f}, held NEE {held[1] / spp:.2f}", flush=True)
    assert ref[1] / held[1] > 1.4


# ------------------------------------------------------------------ maxBounces > 0
# Every vertex of a longer path is attached to its incoming ray, whose origin follows the previous
# vertex and whose direction carries the previous vertex's warp, so a vertex's motion propagates
# to every later vertex. The same three pins hold at one and two bounces.

def _fd_b(osc, data, v0, v1, field, kind, eps, spp, bounces):
    p0, n0 = data.positions.copy(), data.normals.copy()
    vals = []
    for sg in (1, -1):
        p, n = p0.copy(), n0.copy()
        if kind == "pos":
            p[v0:v1] += sg * eps * field
        else:
            n[v0:v1] += sg * eps * field
        osc.animate(p, n)
        im, _ = osc.render_war(W, H, spp=spp, frame=0, max_bounces=bounces)
        vals.append(float((im.astype(np.float64) * _ramp()).sum()))
    osc.animate(p0, n0)
    return (vals[0] - vals[1]) / (2 * eps)


@pytest.mark.parametrize("bounces", [1, 2])
def test_restatement_reproduces_the_primal_multibounce(box, bounces):
    data, v0, v1, osc = box
    spp = 8
    pr = _position_probes(data, v0, v1, osc.num_vertices)[:1]
    _, val = osc.war_geom_probe(W, H, spp, 40, _ramp(), pr, aux=16, log10_kappa=3, log10_kappa_screen=3,
                                max_bounces=bounces)
    img, _ = osc.render_war(W, H, spp=spp, frame=40, max_bounces=bounces)
    ref = float((img.astype(np.float64) * _ramp()).sum() * spp)
    assert val == pytest.approx(ref, rel=1e-6)
    _, val0 = osc.war_geom_probe(W, H, spp, 40, _ramp(), pr, aux=16, log10_kappa=3, log10_kappa_screen=3)
    assert val > val0 * 1.05  # the indirect light is there


def test_normal_gradients_match_fd_two_bounces(box):
    data, v0, v1, osc = box
    rng = np.random.default_rng(4)
    fields = [rng.normal(size=(v1 - v0, 3)).astype(np.float32), np.tile(np.float32([0, 1, 0]), (v1 - v0, 1))]
    pr = np.zeros((2, osc.num_vertices, 6), np.float32)
    for k, f in enumerate(fields):
        pr[k, v0:v1, 3:6] = f
    spp = 1024
    d, _ = osc.war_geom_probe(W, H, spp, 0, _ramp(), pr, aux=16, log10_kappa=3, log10_kappa_screen=3, max_bounces=2)
    for k, f in enumerate(fields):
        fd = _fd_b(osc, data, v0, v1, f, "nrm", 1e-2, 1024, 2)
        est = d[k] / spp
        print(f"two bounces, normal field {k}: estimator {est:.5f}  fd {fd:.5f}", flush=True)
        assert abs(est - fd) <= 0.1 * abs(fd)


@pytest.mark.parametrize("bounces", [1, 2])
def test_held_nee_position_gradients_match_fd_multibounce(box, bounces):
    data, v0, v1, osc = box
    pr = _position_probes(data, v0, v1, osc.num_vertices)
    spp = 512
    d, _ = osc.war_geom_probe(W, H, spp, 100 * spp, _ramp(), pr, aux=64, log10_kappa=4, log10_kappa_screen=4,
                              nee_direction_held=True, max_bounces=bounces)
    est = d / spp
    fd = np.array([_fd_b(osc, data, v0, v1, pr[k, v0:v1, 0:3], "pos", 4e-3, 1024, bounces) for k in range(4)])
    rel = float(np.linalg.norm(est - fd) / np.linalg.norm(fd))
    print(f"{bounces} bounces, held NEE: estimator {np.round(est, 3)}  fd {np.round(fd, 3)}  rel L2 {rel:.3f}", flush=True)
    assert rel <= 0.2
