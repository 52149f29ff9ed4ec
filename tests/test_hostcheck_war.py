"""Host check build of the WARDiffPathTracer kernels' per-path code (CPU, AddressSanitizer).

The kernels' grid-loop items (warItem for the primal and the path-replay adjoint, geomItem for
the vertex gradients) are compiled for the CPU together with the library's host scene builder
(tests/hostcheck). Every array the kernels read is an exact-size heap allocation, each traversal
stack is one lane's LDS part plus its overflow part, and gradient rows are exact-size too, so any
out-of-bounds access aborts the program with the sanitizer's report. The checks:

* the configuration of the vertex-gradient finite-difference test (displaced sphere in the
  closed box, 64 x 64, 64 auxiliary rays, vMF concentration 10^4): host vertex updates that
  rebuild the scene, primal renders on the moved scenes, then the adjoint and vertex-gradient
  passes on the restored scene; a bounded sample of its frames here, all of them (512 spp
  renders and 256 spp backward passes over 8 frames each) with FHP_HOSTCHECK_FULL=1;
* the same passes at the vertex-gradient workload's settings (16 auxiliary rays, 10^3);
* shallow LDS stacks (2 and 4 entries) so traversal runs through the overflow part of the stack;
* multi-bounce primal and adjoint paths on a sphere field, on textured spheres and on a sphere
  field of transmissive materials (the adjoint instances with the transmission lobes).
Each step must also report no trace error (no ray needed more stack than the tree's bound).
A stack made smaller than the tree needs must report the overflow, still without any
out-of-bounds access.
"""
import os
import re

import numpy as np
import pytest

from falcor_amd import scene as S

hostcheck = pytest.importorskip("tests.hostcheck.run")


def _run(data, steps, dLdI, timeout=900):
    rc, out, err = hostcheck.run(data, steps, dLdI=dLdI, timeout=timeout)
    assert "ERROR: AddressSanitizer" not in err and "runtime error" not in err, err[-4000:]
    assert rc == 0, f"host check failed ({rc}):\n{out}\n{err[-4000:]}"
    for line in out.splitlines():
        m = re.search(r"stack_peak=(\d+) stack_capacity=(\d+) trace_errors=(\d+)", line)
        if m:
            assert int(m.group(3)) == 0 and int(m.group(1)) <= int(m.group(2)), line
    return out


def _ramp(W, H):
    x = (np.arange(W) + 0.5) / W
    return np.repeat(np.broadcast_to(x[None, :], (H, W))[..., None], 3, axis=2).astype(np.float32)


def _fd_case_steps(mesh, W, H, primal_spp, bwd_spp, frame, aux, log_kappa):
    return [f"translate {mesh} 0.004 0 0", f"primal {W} {H} {primal_spp} {frame} 0 24",
            f"translate {mesh} -0.008 0 0", f"primal {W} {H} {primal_spp} {frame} 0 24",
            f"translate {mesh} 0.004 0 0",
            f"adjoint {W} {H} {bwd_spp} {100 + frame} 0 24 64",
            f"geom {W} {H} {bwd_spp} {100 + frame} {aux} {log_kappa} {log_kappa} {mesh} 64 24"]


@pytest.fixture(scope="module")
def box_sphere():
    data = S.displaced_sphere_in_box(0, radius=0.12, segments=(24, 12), sigma=0.05)
    return data, data.mesh_names.index("Sphere")


def test_vertex_gradient_fd_configuration_in_bounds(box_sphere):
    data, mesh = box_sphere
    W = H = 64
    full = os.environ.get("FHP_HOSTCHECK_FULL") == "1"
    frames = range(8) if full else (0, 5)
    steps = []
    for f in frames:
        steps += _fd_case_steps(mesh, W, H, 512 if full else 8, 256 if full else 4, f, 64, 4)
    out = _run(data, steps, _ramp(W, H), timeout=7200 if full else 900)
    g = [float(v) for v in re.findall(r"grad_abs_sum=([0-9.e+-]+)", out)]
    assert len(g) == len(frames) and all(v > 0 for v in g)
    assert "finite=0" not in out


def test_vertex_gradient_workload_settings_in_bounds(box_sphere):
    data, mesh = box_sphere
    W = H = 48
    _run(data, [f"geom {W} {H} 4 7 16 3 3 {mesh} 128 16", f"geom {W} {H} 2 9 16 3 3 {mesh} 1 32"], _ramp(W, H))


@pytest.mark.parametrize("lds", [2, 4])
def test_overflow_stack_in_bounds(box_sphere, lds):
    data, mesh = box_sphere
    W = H = 32
    out = _run(data, [f"primal {W} {H} 2 0 2 {lds}", f"adjoint {W} {H} 2 1 2 {lds} 8",
                      f"geom {W} {H} 1 2 8 3 3 {mesh} 8 {lds}"], _ramp(W, H))
    peaks = [int(v) for v in re.findall(r"stack_peak=(\d+)", out)]
    assert max(peaks) > lds  # the overflow part of the stack was used


def _transmissive_field():
    data = S.sphere_field(2, num_spheres=64, num_emissive=256)
    m = data.materials
    for i in range(64):
        m[i, S.M_SPEC_TRANS] = (0.0, 1.0, 0.6, 0.3)[i % 4]
        m[i, S.M_DIFF_TRANS] = (0.0, 0.0, 0.5, 0.8)[i % 4]
        m[i, S.M_TRANS:S.M_TRANS + 3] = (0.9, 0.7, 0.5)
        m[i, S.M_THIN] = float(i % 8 == 3)
    return data


@pytest.mark.parametrize("which", ["sphere_field", "textured_spheres", "transmissive_field"])
def test_multi_bounce_paths_in_bounds(which):
    data = {"sphere_field": lambda: S.sphere_field(1, num_spheres=64, num_emissive=256),
            "textured_spheres": lambda: S.textured_spheres(3, num_objects=4, tex_res=16),
            "transmissive_field": _transmissive_field}[which]()
    W = H = 48
    _run(data, [f"primal {W} {H} 4 0 4 16", f"adjoint {W} {H} 4 1 4 16 16", f"adjoint {W} {H} 2 2 1 4 3"],
         np.ones((H, W, 3), np.float32))


@pytest.mark.parametrize("which", ["sphere_field", "textured_spheres", "transmissive_field"])
def test_hit_log_replay_equals_retrace(which):
    """The adjoint's replay from the first pass's hit log (one word per closest-hit trace, at most
    maxBounces + 2, in an exact-size array) gives the gradients of a replay that traces again, bit
    for bit (sequential host run: no atomics reorder the sums), and traces half the rays."""
    data = {"sphere_field": lambda: S.sphere_field(1, num_spheres=64, num_emissive=256),
            "textured_spheres": lambda: S.textured_spheres(3, num_objects=4, tex_res=16),
            "transmissive_field": _transmissive_field}[which]()
    W = H = 40
    steps = []
    for B in (0, 1, 4):
        steps += [f"adjoint {W} {H} 3 {B + 1} {B} 16 8", f"adjoint {W} {H} 3 {B + 1} {B} 16 8 log=1"]
    out = _run(data, steps, np.ones((H, W, 3), np.float32))
    fps = re.findall(r"grad_fingerprint=([0-9a-f]+) log=(\d)", out)
    rays = [int(v) for v in re.findall(r"adjoint paths=\d+ rays=(\d+)", out)]
    assert len(fps) == 6 and len(rays) == 6
    for i in range(0, 6, 2):
        assert fps[i][1] == "0" and fps[i + 1][1] == "1"
        assert fps[i][0] == fps[i + 1][0], out
        assert rays[i] == 2 * rays[i + 1], out


@pytest.mark.parametrize("bounces", [1, 3])
def test_vertex_gradients_multibounce_in_bounds(box_sphere, bounces):
    """Vertex gradients of longer paths: the per-thread vertex records (exact size for
    maxBounces + 2 vertices) and every array the reverse sweep reads stay in bounds, for both NEE
    forms, with and without the aux-hit log; the gradients are finite and nonzero."""
    data, mesh = box_sphere
    W = H = 24
    steps = []
    for opt in ("", " neeheld=1", " log=1"):
        steps.append(f"geom {W} {H} 2 5 16 3 3 {mesh} 8 24 b={bounces}{opt}")
    out = _run(data, steps, _ramp(W, H))
    g = [float(v) for v in re.findall(r"grad_abs_sum=([0-9.e+-]+)", out)]
    assert len(g) == 3 and all(v > 0 for v in g)
    assert "finite=0" not in out
    fps = re.findall(r"geom grad_fingerprint=([0-9a-f]+) log=(\d)", out)
    assert fps[0][0] == fps[2][0]  # the aux-hit log changes no bit


def test_aux_log_replay_equals_retrace(box_sphere):
    """The vertex-gradient kernel's second loop over each warp's aux rays reads the first loop's
    hits from the aux-hit log (3 words per aux ray, exact-size array): gradients equal, bit for
    bit, those of a second loop that traces again, for both NEE forms, and fewer rays are traced."""
    data, mesh = box_sphere
    W = H = 24
    steps = []
    for opt in ("", " neeheld=1"):
        for log in ("", " log=1"):
            steps.append(f"geom {W} {H} 2 3 16 3 3 {mesh} 8 24{opt}{log}")
    out = _run(data, steps, _ramp(W, H))
    fps = re.findall(r"geom grad_fingerprint=([0-9a-f]+) log=(\d)", out)
    rays = [int(v) for v in re.findall(r"geom paths=\d+ rays=(\d+)", out)]
    assert len(fps) == 4 and len(rays) == 4
    for i in (0, 2):
        assert fps[i][1] == "0" and fps[i + 1][1] == "1"
        assert fps[i][0] == fps[i + 1][0], out
        assert rays[i + 1] < 0.6 * rays[i], out
    assert "finite=0" not in out


def test_undersized_stack_reports_overflow_in_bounds(box_sphere):
    data, _ = box_sphere
    W = H = 32
    rc, out, err = hostcheck.run(data, [f"primal {W} {H} 2 0 2 2 cap=3"], dLdI=_ramp(W, H))
    assert "ERROR: AddressSanitizer" not in err and "runtime error" not in err, err[-4000:]
    assert rc == 1, (rc, out, err[-2000:])
    m = re.search(r"stack_peak=(\d+) stack_capacity=(\d+) trace_errors=(\d+)", out)
    assert m and int(m.group(2)) == 3 and int(m.group(3)) == 1 and int(m.group(1)) > 3, out


@pytest.mark.parametrize("bounces,held,general", [(0, False, False), (0, False, True), (0, True, True), (1, True, False),
                                                   (2, False, False), (2, True, False)])
def test_vertex_gradient_code_matches_oracle_estimator(box_sphere, oracle_mod, tmp_path, bounces, held, general):
    """The kernel's per-path vertex-gradient code (built for the host) against the oracle's
    restatement of the estimator (central differences of the path radiance with every decision
    held, FP64), same seeds, at 0, 1 and 2 bounces: projections on three translations, a uniform
    scale, a random position field and a random normal field. At 0 bounces both the one-pass
    direct-illumination code and the general reverse sweep over recorded vertices are checked."""
    data, mesh = box_sphere
    W = H = 32
    spp, frame = 8, 5
    out = tmp_path / "grads.bin"
    step = (f"geom {W} {H} {spp} {frame} 16 3 3 {mesh} 4 24 b={bounces} out={out}" + (" neeheld=1" if held else "")
            + (" general=1" if general else ""))
    _run(data, [step], _ramp(W, H))
    off = data.mesh_vertex_offsets
    v0, v1 = int(off[mesh]), int(off[mesh + 1])
    g = np.fromfile(out, np.float64)
    gp, gn = g[:3 * (v1 - v0)].reshape(-1, 3), g[3 * (v1 - v0):].reshape(-1, 3)
    osc = oracle_mod.OracleScene(data)
    rng = np.random.default_rng(11)
    pr = np.zeros((6, osc.num_vertices, 6), np.float32)
    for a in range(3):
        pr[a, v0:v1, a] = 1.0
    pr[3, v0:v1, 0:3] = data.positions[v0:v1] - data.positions[v0:v1].mean(0)
    pr[4, v0:v1, 0:3] = rng.normal(size=(v1 - v0, 3))
    pr[5, v0:v1, 3:6] = rng.normal(size=(v1 - v0, 3))
    ref, _ = osc.war_geom_probe(W, H, spp, frame * spp, _ramp(W, H), pr, aux=16, log10_kappa=3, log10_kappa_screen=3,
                                nee_direction_held=held, max_bounces=bounces)
    got = np.array([(gp * pr[k, v0:v1, 0:3]).sum() + (gn * pr[k, v0:v1, 3:6]).sum() for k in range(6)])
    rel = np.abs(got - ref) / np.abs(ref)
    print(f"bounces {bounces} held {held}: code {np.round(got, 4)} oracle {np.round(ref, 4)} "
          f"rel {np.array2string(rel, precision=2)}", flush=True)
    assert np.all(np.abs(ref) > 0)
    assert np.all(rel <= 2e-3)
