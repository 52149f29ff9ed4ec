"""Pins the oracle (CPU restatement) against known answers, closed forms and invariants.

Every oracle function used as a checker by the GPU parity tests is covered here:
PRNG / bit tricks / hash (known answers), ray-origin offset (closed form), watertight
intersection (FP64 reference + watertightness), camera basis (closed form), BSDF sampling
(unbiasedness against quadrature), microfacet normalisation, light collection and LightBVH
(structural invariants, pdf normalisation, sample/eval pdf consistency) and the integrators
(closed-form furnace scenes, NEE/MIS unbiasedness).
"""
import json
import math
from pathlib import Path

import numpy as np
import pytest

from falcor_amd import scene as S

GOLDEN = json.loads((Path(__file__).parent / "golden" / "known_answers.json").read_text())


# --------------------------------------------------------------------------- PRNG and hashes
def test_splitmix64_published_vector(oracle_mod):
    out = oracle_mod.Oracle.splitmix64(0, 3)
    assert [int(v) for v in out] == GOLDEN["splitmix64_state0"]


def test_xoshiro128ss_known_answers(oracle_mod):
    """Eight outputs from state {1, 2, 3, 4}: the third is the first to depend on rotl(s3, 11), the
    fifth the first through s2 ^= s1 << 9 (tests/golden/make_known_answers.py)."""
    want = GOLDEN["xoshiro128ss_state1234"]
    assert len(want) == 8
    out = oracle_mod.Oracle.xoshiro128ss([1, 2, 3, 4], len(want))
    assert [int(v) for v in out] == want


def test_interleave_known_answers(oracle_mod):
    for case in GOLDEN["interleave"]:
        got = int(oracle_mod.Oracle.interleave([case["x"]], [case["y"]])[0])
        assert got == case["expect"]


def test_interleave_matches_bitwise_definition(oracle_mod):
    rng = np.random.default_rng(0)
    x = rng.integers(0, 1 << 16, 4096, dtype=np.uint32)
    y = rng.integers(0, 1 << 16, 4096, dtype=np.uint32)
    got = oracle_mod.Oracle.interleave(x, y)
    ref = np.zeros_like(x)
    for i in range(16):  # result bit 2i = x_i, 2i+1 = y_i
        ref |= ((x >> i) & 1) << (2 * i)
        ref |= ((y >> i) & 1) << (2 * i + 1)
    assert np.array_equal(got, ref)


def test_sample_generator_seeding_and_quantisation(oracle_mod):
    """sampleNext1D = (next() >> 8) * 2^-24 of xoshiro128** seeded by SplitMix64(interleave | n << 32)."""
    px, py, n = 37, 1021, 5
    st = (n << 32) | int(oracle_mod.Oracle.interleave([px], [py])[0])
    s = oracle_mod.Oracle.splitmix64(st, 2)
    seed = [int(s[0]) & 0xFFFFFFFF, int(s[0]) >> 32, int(s[1]) & 0xFFFFFFFF, int(s[1]) >> 32]
    bits = oracle_mod.Oracle.xoshiro128ss(seed, 64)
    expect = (bits >> 8).astype(np.float64) / float(1 << 24)
    got = oracle_mod.Oracle.sample_gen(px, py, n, 64)
    assert np.array_equal(got.astype(np.float64), expect)
    assert got.min() >= 0.0 and got.max() < 1.0


def test_sample_generator_statistics(oracle_mod):
    """Mean/correlation thresholds in the spirit of SampleGeneratorTests.cpp:156-164."""
    vals = np.stack([oracle_mod.Oracle.sample_gen(x, y, 0, 16) for x in range(32) for y in range(32)])
    assert abs(vals.mean() - 0.5) < 0.01
    c = np.corrcoef(vals[:, 0], vals[:, 1])[0, 1]
    assert abs(c) < 0.1


def test_jenkins_hash_known_answers(oracle_mod):
    k = GOLDEN["jenkins"]
    got = oracle_mod.Oracle.jenkins(np.array(k["inputs"], np.uint32))
    assert [int(v) for v in got] == k["hashes"]


def test_jenkins_hash_is_a_permutation(oracle_mod):
    x = np.arange(1 << 20, dtype=np.uint32)
    h = oracle_mod.Oracle.jenkins(x)
    assert np.unique(h).size == x.size  # perfect hash on this range (HashUtils.slang:36-38)
    assert int(h[0]) != 0


# --------------------------------------------------------------------------- geometry helpers
def test_offset_ray_closed_forms(oracle_mod):
    rng = np.random.default_rng(1)
    n = rng.normal(size=(2000, 3)).astype(np.float32)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    # |p| < 1/16: p + n * 3/65536 exactly.
    p_small = (rng.uniform(-0.06, 0.06, (2000, 3))).astype(np.float32)
    out = oracle_mod.Oracle.offset_ray(p_small, n)
    assert np.array_equal(out, p_small + n * np.float32(3.0 / 65536.0))
    # |p| >= 1/16: integer offset of int(n * 768) ULPs away from zero along the normal sign.
    p_big = (rng.uniform(0.1, 100.0, (2000, 3)) * rng.choice([-1, 1], (2000, 3))).astype(np.float32)
    out = oracle_mod.Oracle.offset_ray(p_big, n)
    ioff = (n * np.float32(768.0)).astype(np.int32)
    bits = p_big.view(np.int32) + np.where(p_big < 0, -ioff, ioff)
    assert np.array_equal(out.view(np.int32), bits)


def _ray_tri_f64(o, d, v):
    e1, e2 = v[1] - v[0], v[2] - v[0]
    m = np.stack([-d, e1, e2], 1)
    t, b1, b2 = np.linalg.solve(m, o - v[0])
    return t, b1, b2


def test_watertight_against_fp64_solution(oracle_mod):
    rng = np.random.default_rng(2)
    n = 4000
    tris = rng.uniform(-1, 1, (n, 9)).astype(np.float32)
    o = rng.uniform(-3, 3, (n, 3)).astype(np.float32)
    # aim at a random point inside the triangle
    bc = rng.dirichlet([1, 1, 1], n)
    tgt = (tris.reshape(n, 3, 3) * bc[:, :, None]).sum(1)
    d = (tgt - o).astype(np.float32)
    rays = np.concatenate([o, d], 1)
    out = oracle_mod.Oracle.intersect(rays, tris)
    hit = out[:, 0] == 1
    assert hit.mean() > 0.999
    for i in np.nonzero(hit)[0][:500]:
        t, b1, b2 = _ray_tri_f64(o[i].astype(np.float64), d[i].astype(np.float64), tris[i].reshape(3, 3).astype(np.float64))
        assert abs(out[i, 1] - t) <= 1e-4 * max(1.0, abs(t))
        assert abs(out[i, 2] - b1) < 1e-4 and abs(out[i, 3] - b2) < 1e-4
    # Rays aimed away from the plane miss or hit at negative t (the caller applies the range).
    out2 = oracle_mod.Oracle.intersect(np.concatenate([o, -d], 1), tris)
    assert np.all((out2[:, 0] == 0) | (out2[:, 1] < 0))


def test_watertight_shared_edge_never_leaks(oracle_mod):
    """Rays through the shared edge of two triangles hit at least one of them (watertightness)."""
    rng = np.random.default_rng(3)
    n = 3000
    a = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    b = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    c1 = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    c2 = (a + b - c1).astype(np.float32)  # opposite side of edge ab
    s = rng.uniform(0.01, 0.99, (n, 1)).astype(np.float32)
    p = (a * (1 - s) + b * s).astype(np.float32)
    o = (p + rng.normal(size=(n, 3)).astype(np.float32) * 2).astype(np.float32)
    d = (p - o).astype(np.float32)
    rays = np.concatenate([o, d], 1)
    h1 = oracle_mod.Oracle.intersect(rays, np.concatenate([a, b, c1], 1))[:, 0]
    h2 = oracle_mod.Oracle.intersect(rays, np.concatenate([b, a, c2], 1))[:, 0]
    assert np.all((h1 + h2) >= 1)


def test_camera_basis_closed_form(oracle_mod):
    data = S.closed_box(0)
    sc = oracle_mod.OracleScene(data)
    U, V, W = sc.camera_basis().reshape(3, 3).astype(np.float64)
    c = data.camera
    w = np.subtract(c.target, c.position)
    w = w / np.linalg.norm(w) * c.focal_distance
    u = np.cross(w, c.up)
    u = u / np.linalg.norm(u)
    v = np.cross(u, w)
    v = v / np.linalg.norm(v)
    th = math.tan(c.fov_y / 2)
    np.testing.assert_allclose(W, w, rtol=1e-6)
    np.testing.assert_allclose(U, u * c.focal_distance * th * c.aspect, rtol=1e-5, atol=1e-3)
    np.testing.assert_allclose(V, v * c.focal_distance * th, rtol=1e-5, atol=1e-3)


# --------------------------------------------------------------------------- sampling and BSDFs
def test_cosine_hemisphere_pdf_and_moments(oracle_mod):
    rng = np.random.default_rng(4)
    u = rng.random((200000, 2)).astype(np.float32)
    w = oracle_mod.Oracle.sample_cosine(u)
    np.testing.assert_allclose(np.linalg.norm(w[:, :3], axis=1), 1.0, atol=1e-5)
    np.testing.assert_allclose(w[:, 3], w[:, 2] / math.pi, rtol=1e-5)
    # E[cos] under cos/pi sampling = 2/3; E[x] = E[y] = 0
    assert abs(w[:, 2].mean() - 2 / 3) < 3e-3
    assert abs(w[:, 0].mean()) < 3e-3 and abs(w[:, 1].mean()) < 3e-3


def test_vmf_mean_cosine(oracle_mod):
    rng = np.random.default_rng(5)
    u = rng.random((200000, 2)).astype(np.float32)
    for kappa in (10.0, 1000.0):
        w = oracle_mod.Oracle.sample_vmf(u, kappa)
        expect = 1 / math.tanh(kappa) - 1 / kappa  # mean cosine of a vMF distribution
        assert abs(w[:, 2].mean() - expect) < 3e-3


def _hemi_grid(n=400):
    # midpoint quadrature over (cos theta, phi)
    ct = (np.arange(n) + 0.5) / n
    ph = (np.arange(2 * n) + 0.5) / (2 * n) * 2 * math.pi
    C, P = np.meshgrid(ct, ph, indexing="ij")
    st = np.sqrt(1 - C ** 2)
    w = np.stack([st * np.cos(P), st * np.sin(P), C], -1).reshape(-1, 3)
    dA = (1.0 / n) * (2 * math.pi / (2 * n))
    return w.astype(np.float32), dA


@pytest.mark.parametrize("mat", [
    [0, 0.8, 0.5, 0.2, 0.5, 0.0, 1.5],   # dielectric, rough
    [0, 0.9, 0.9, 0.9, 0.2, 1.0, 1.5],   # metal
    [0, 0.3, 0.6, 0.9, 0.9, 0.3, 1.5],
])
def test_standard_bsdf_sampling_is_unbiased(oracle_mod, mat):
    """E[weight * g(wo)] over sampled directions equals the quadrature of f * g (g = 1 and g = wo.x)."""
    rng = np.random.default_rng(6)
    wi = np.array([0.3, -0.2, 0.9], np.float32)
    wi /= np.linalg.norm(wi)
    n = 400000
    s = oracle_mod.Oracle.bsdf_sample(np.array(mat, np.float32), np.tile(wi, (n, 1)),
                                      rng.integers(0, 2 ** 31, (n, 2), dtype=np.uint32))
    valid = s[:, 7] > 0
    wgt = np.where(valid[:, None], s[:, 3:6], 0.0)
    grid, dA = _hemi_grid()
    f = oracle_mod.Oracle.bsdf_eval(np.array(mat, np.float32), np.tile(wi, (grid.shape[0], 1)), grid)
    for g_mc, g_q in ((np.ones(n), np.ones(grid.shape[0])), (s[:, 0], grid[:, 0])):
        mc = (wgt * g_mc[:, None]).mean(0)
        q = (f[:, :3] * g_q[:, None]).sum(0) * dA
        np.testing.assert_allclose(mc, q, rtol=0.03, atol=5e-3)
    # The pdf reported with each non-delta sample equals evalPdf of that direction.
    ok = s[:, 7] == 1
    e = oracle_mod.Oracle.bsdf_eval(np.array(mat, np.float32), np.tile(wi, (int(ok.sum()), 1)), s[ok, :3])
    np.testing.assert_allclose(s[ok, 6], e[:, 3], rtol=2e-3, atol=1e-6)
    # The pdf integrated over the upper hemisphere equals the fraction of samples that land
    # there (rough GGX lobes also sample directions below the horizon, which are rejected).
    assert abs(f[:, 3].sum() * dA - ok.mean()) < 0.01


def test_pbrt_diffuse_known_answer(oracle_mod):
    """DiffMaterialTests.cpp:86-114: d(sum f)/d albedo = wo.z / pi = 0.3003115."""
    k = GOLDEN["diff_pbrt_diffuse"]
    wi = np.array(k["wi"], np.float32); wi /= np.linalg.norm(wi)
    wo = np.array(k["wo"], np.float32); wo /= np.linalg.norm(wo)
    mat = np.array([1, *k["base_color"], 0, 0, 1.5], np.float32)
    eps = 1e-2
    f0 = oracle_mod.Oracle.bsdf_eval(mat, wi[None], wo[None])[0, :3]
    for c in range(3):
        m2 = mat.copy(); m2[1 + c] += eps
        f1 = oracle_mod.Oracle.bsdf_eval(m2, wi[None], wo[None])[0, :3]
        g = (f1.sum() - f0.sum()) / eps
        assert abs(g - k["grad_albedo"][c]) < k["tol"]
    # Lambert: sum f = sum(albedo) / pi * wo.z, so its wo.z derivative is the reference's 0.5411268.
    assert abs(f0.sum() / wo[2] - k["grad_wo_z"]) < k["tol"]


# --------------------------------------------------------------------------- lights
def test_light_collection_flux(oracle_mod):
    data = S.sphere_field(1, num_spheres=64, num_emissive=256, box=20.0)
    sc = oracle_mod.OracleScene(data)
    L = sc.lights()
    assert L.shape[0] == 256
    p = L[:, :9].reshape(-1, 3, 3).astype(np.float64)
    area = 0.5 * np.linalg.norm(np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]), axis=1)
    np.testing.assert_allclose(L[:, 12], area, rtol=1e-5)
    lum = L[:, 13] * 0.2126 + L[:, 14] * 0.7152 + L[:, 15] * 0.0722
    np.testing.assert_allclose(L[:, 16], lum * area * math.pi, rtol=1e-5)


def test_light_bvh_invariants(oracle_mod):
    data = S.sphere_field(1, num_spheres=128, num_emissive=1000, box=30.0)
    sc = oracle_mod.OracleScene(data)
    nodes, tri, masks = sc.light_bvh()
    L = sc.lights()
    leaves = np.nonzero(nodes[:, 11] == 1)[0]
    # every active light in exactly one leaf; leaves hold <= 10 lights
    counts = nodes[leaves, 14].astype(int)
    assert counts.max() <= 10 and counts.sum() == L.shape[0]
    assert sorted(tri.tolist()) == list(range(L.shape[0]))

    # Nodes carry the packed precision (half-float extents rounded to nearest, octahedral 16-bit
    # cone axes): a box may miss its content by half a half-float step of its extent.
    def check(i, depth, mask):
        n = nodes[i]
        lo, hi = n[0:3] - n[3:6], n[0:3] + n[3:6]
        tol = 1e-4 + 2.0 ** -11 * np.abs(n[3:6])
        if n[11] == 1:
            off, cnt = int(n[13]), int(n[14])
            ids = tri[off:off + cnt]
            pts = L[ids, :9].reshape(-1, 3)
            assert np.all(pts >= lo - tol) and np.all(pts <= hi + tol)
            np.testing.assert_allclose(n[6], L[ids, 16].sum(), rtol=1e-4)
            for li in ids:
                assert int(masks[li]) == mask  # bitmask encodes the path to the leaf
            if n[10] != -1:  # cone bounds all normals
                assert np.all(L[ids, 9:12] @ n[7:10] >= n[10] - 2e-4)
            return n[6]
        left, right = i + 1, int(n[12])
        fl = check(left, depth + 1, mask)
        fr = check(right, depth + 1, mask | (1 << depth))
        for c in (left, right):
            cl, ch = nodes[c, 0:3] - nodes[c, 3:6], nodes[c, 0:3] + nodes[c, 3:6]
            tc = tol + 2.0 ** -11 * np.abs(nodes[c, 3:6])  # both boxes rounded, possibly towards each other
            assert np.all(cl >= lo - tc) and np.all(ch <= hi + tc)
        np.testing.assert_allclose(n[6], fl + fr, rtol=1e-4)
        return n[6]

    check(0, 0, 0)


@pytest.mark.parametrize("kind", [0, 1])
def test_light_selection_pdf_normalised_and_consistent(oracle_mod, kind):
    data = S.sphere_field(1, num_spheres=128, num_emissive=600, box=30.0)
    sc = oracle_mod.OracleScene(data)
    lib = oracle_mod.oracle_lib()
    rng = np.random.default_rng(7)
    nl = sc.lights().shape[0]
    for _ in range(5):
        pos = rng.uniform(0, 30, 3).astype(np.float32)
        nrm = rng.normal(size=3).astype(np.float32); nrm /= np.linalg.norm(nrm)
        pdfs = np.zeros(nl, np.float32)
        lib.fo_light_selection_pdfs(sc.h, kind, pos.ctypes.data, nrm.ctypes.data, pdfs.ctypes.data)
        # Traversal probabilities split 1 at every node, so the selection pdfs sum to one minus
        # the probability of reaching a node whose children both have zero importance.
        assert pdfs.sum() <= 1.0 + 1e-5
        # sampled selection frequencies follow the pdfs; the pdf returned by sampling equals evalPdf
        m = 20000
        out = np.zeros((m, 4), np.float32)
        seeds = rng.integers(0, 2 ** 31, (m, 2), dtype=np.uint32)
        P, N = np.tile(pos, (m, 1)), np.tile(nrm, (m, 1))  # keep the arrays alive during the call
        lib.fo_light_sample(sc.h, kind, m, P.ctypes.data, N.ctypes.data, seeds.ctypes.data, out.ctypes.data)
        sel = out[:, 0] >= 0
        li = out[sel, 0].astype(int)
        np.testing.assert_allclose(out[sel, 1], pdfs[li], rtol=1e-5)
        assert abs(pdfs.sum() - sel.mean()) < 0.01
        freq = np.bincount(li, minlength=nl) / m
        top = np.argsort(pdfs)[-5:]
        np.testing.assert_allclose(freq[top], pdfs[top], atol=0.01)


# --------------------------------------------------------------------------- integrators
def _furnace(albedo=0.6, le=17.0):
    """Lambertian floor under a huge downward-facing emitter: every cosine-sampled scatter ray
    from the floor hits the emitter, so with maxBounces = 0 each path returns albedo * Le."""
    sb = S.SceneBuilder()
    floor = sb.add_material(S.pbrt_diffuse_material((albedo, albedo, albedo)))
    light = sb.add_material(S.standard_material(base=(0, 0, 0), roughness=1.0, emissive=(le, le, le)))
    quad = S.TriangleMesh.create_quad()
    sb.add_mesh_instance(quad, floor, S.Transform((2.0, 1.0, 2.0)), "floor")
    sb.add_mesh_instance(quad, light, S.Transform.deg((1e4, 1.0, 1e4), (0, 1.0, 0), (180, 0, 0)), "sky")
    sb.camera = S.Camera((0, 0.5, 0.0), (0, 0, 0.0), (0, 0, -1), math.radians(40), 1.0)
    return sb.build()


def test_mpt_furnace_closed_form(oracle_mod):
    sc = oracle_mod.OracleScene(_furnace())
    img, stats = sc.render_mpt(32, 32, spp=2, max_bounces=0, brute=True)
    expect = np.float32(0.6) * np.float32(17.0)
    assert np.all(img == expect)
    # closest-hit rays = primary + (maxBounces + 1) scatter rays per path
    assert stats[0] == 32 * 32 * 2 * 2 and stats[1] == 0


def test_mpt_nee_mis_unbiased(oracle_mod):
    """MinimalPathTracer without NEE and with emissive NEE + MIS estimate the same image."""
    data = S.closed_box(0)
    sc = oracle_mod.OracleScene(data)
    a, _ = sc.render_mpt(24, 24, spp=256, max_bounces=2, brute=True)
    for sampler in (0, 1):
        b, st = sc.render_mpt(24, 24, spp=256, max_bounces=2, nee=True, light_sampler=sampler, brute=True)
        assert st[1] > 0
        assert abs(a.mean() - b.mean()) < 0.03 * a.mean()


def test_war_primal_furnace_mean(oracle_mod):
    """WAR primal (NEE + MIS, uniform emissive sampler) converges to albedo * Le in the furnace."""
    sc = oracle_mod.OracleScene(_furnace())
    img, st = sc.render_war(16, 16, spp=64, max_bounces=0, brute=True)
    assert abs(img.mean() - 0.6 * 17.0) < 0.02 * 0.6 * 17.0
    assert st[1] > 0


def _emissive_ramp_scene():
    """A camera-facing emissive quad (x, y in [-1, 1] at z = 0) whose emissive texture is the ramp
    Le = 1 + 4u = 3 + 2x, seen from (0, 0, 3) with a 20 degree vertical field of view: every camera
    ray hits it inside u in [0.24, 0.76], where bilinear filtering reproduces the ramp exactly.
    Nothing else is in the scene and the base colour is black, so a pixel sample's radiance is the
    emission at its primary hit (with NEE off: a light sample on the quad itself is coplanar with
    the shading point, where computeNEE's areaToSolidAngle / detach(areaToSolidAngle) is inf / inf
    in the reference, PTUtils.slang:236-238)."""
    sb = S.SceneBuilder()
    tw = 64
    ramp = 1.0 + 4.0 * (np.arange(tw) + 0.5) / tw
    tex = sb.add_texture(np.repeat(np.repeat(ramp[None, :, None], 2, axis=0), 3, axis=2).astype(np.float32))
    m = sb.add_material(S.standard_material(base=(0, 0, 0), roughness=1.0, emissive=(1, 1, 1), emissive_texture=tex))
    sb.add_mesh_instance(S.TriangleMesh.create_quad((2.0, 2.0)), m, S.Transform.deg((1, 1, 1), (0, 0, 0), (90, 0, 0)))
    sb.camera = S.Camera((0, 0, 3), (0, 0, 0), (0, 1, 0), math.radians(20.0), 1.0)
    return sb.build()


def test_war_primal_pixel_jitter_known_answer(oracle_mod):
    """tracePaths (WARDiffPathTracer.rt.slang:114-143): SampleGenerator(pixel, frameSeed / 2), three
    2D samples discarded, Gaussian jitter sigma 0.5 from the next two (sampleGaussian :60-63,
    Box-Muller), negated when frameSeed is even (antithetic). On the emissive ramp a sample's value
    is 3 + 6 tan(10 deg) ndc_x of its jittered screen position, predicted here from the independently
    pinned sample generator; a wrong discard count, sigma, flip parity or axis fails it."""
    sc = oracle_mod.OracleScene(_emissive_ramp_scene())
    W = H = 24
    k = 6.0 * math.tan(math.radians(10.0))
    for frame in range(4):  # spp = 1: frameSeed = frame
        img, _ = sc.render_war(W, H, spp=1, frame=frame, max_bounces=0, nee=False)
        want = np.zeros((H, W))
        for py in range(H):
            for px in range(W):
                u = oracle_mod.Oracle.sample_gen(px, py, frame // 2, 8).astype(np.float64)
                jx = math.sqrt(-2.0 * math.log(u[6])) * math.cos(2.0 * math.pi * u[7]) * 0.5
                if frame % 2 == 0:
                    jx = -jx
                want[py, px] = 3.0 + k * (2.0 * (px + 0.5 + jx) / W - 1.0)
        np.testing.assert_allclose(img[..., 0], want, rtol=2e-5, atol=0)
        assert np.array_equal(img[..., 0], img[..., 1]) and np.array_equal(img[..., 0], img[..., 2])
    # The antithetic pair of a seed (frames 2j, 2j + 1) averages to the pixel centre's value.
    a, _ = sc.render_war(W, H, spp=1, frame=0, max_bounces=0, nee=False)
    b, _ = sc.render_war(W, H, spp=1, frame=1, max_bounces=0, nee=False)
    centre = 3.0 + k * (2.0 * (np.arange(W) + 0.5) / W - 1.0)
    np.testing.assert_allclose(0.5 * (a[..., 0] + b[..., 0]), np.broadcast_to(centre, (H, W)), rtol=2e-5)


def test_war_primal_converges_to_fp64_minimal_path_tracer(oracle_mod):
    """The WAR primal and MinimalPathTracer at the same maxBounces (1) estimate the same light
    transport on the closed box: both bound a path at the primary hit plus maxBounces + 1 scatter
    vertices (PTUtils.slang:302-307 terminates at length >= maxBounces + 1 after the length++ of
    each scatter; MinimalPathTracer.rt.slang:342-346, :451 at pathLength >= maxBounces with the
    primary scatter not counted). WAR uses NEE + MIS at both shading vertices, the FP64
    MinimalPathTracer BSDF sampling only. WAR's pixels are Gaussian-filtered (sigma 0.5 px), so
    the FP64 image is rendered at 4 x 4 sample positions per pixel and filtered with the same
    Gaussian; compared are the interior pixels whose filter support does not reach the directly
    visible light (whose edge the 4 x 4 quadrature cannot resolve at 17x the wall radiance). A
    missing NEE or MIS term at the second vertex biases the mean by several percent; a wrong
    jitter width moves the wall edges."""
    data = S.closed_box(0)
    sc = oracle_mod.OracleScene(data)
    W = H = 24
    SS = 4
    war, _ = sc.render_war(W, H, spp=2048, frame=0, max_bounces=1)
    his = [sc.render_mpt_f64(SS * W, SS * H, spp=128, frame=100 + f, max_bounces=1) for f in range(2)]
    c = (np.arange(SS * W) + 0.5) / SS
    bright = np.maximum(his[0], his[1]).max(-1) > 2.0
    refs = np.zeros((2, H, W, 3))
    mask = np.zeros((H, W), bool)
    for py in range(2, H - 2):
        wy = np.exp(-((c - (py + 0.5)) ** 2) / (2 * 0.25))
        near_y = np.abs(c - (py + 0.5)) < 2.0
        for px in range(2, W - 2):
            wx = np.exp(-((c - (px + 0.5)) ** 2) / (2 * 0.25))
            w = wy[:, None] * wx[None, :]
            for k in range(2):
                refs[k, py, px] = (his[k] * w[..., None]).sum((0, 1)) / w.sum()
            mask[py, px] = not (bright & near_y[:, None] & (np.abs(c - (px + 0.5)) < 2.0)[None, :]).any()
    assert mask.sum() > 250
    a = war[mask].astype(np.float64)
    ref = 0.5 * (refs[0] + refs[1])
    b = ref[mask]
    assert abs(a.mean() - b.mean()) < 0.01 * b.mean(), (a.mean(), b.mean())
    # Per pixel, the difference is the two estimates' noise: the FP64 average's own noise is half
    # the RMS of its two independent halves' difference (WAR's is much smaller).
    noise = 0.5 * np.sqrt(((refs[0][mask] - refs[1][mask]) ** 2).mean())
    rms = np.sqrt(((a - b) ** 2).mean())
    assert rms < 1.5 * noise, (rms, noise)


def test_oracle_bvh_matches_brute_force(oracle_mod):
    data = S.sphere_field(1, num_spheres=48, num_emissive=64, box=12.0)
    sc = oracle_mod.OracleScene(data)
    rng = np.random.default_rng(8)
    n = 3000
    o = rng.uniform(-2, 14, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    rays = np.concatenate([o, np.zeros((n, 1), np.float32), d, np.full((n, 1), 1e30, np.float32)], 1)
    a = sc.trace(rays)
    b = sc.trace(rays, brute=True)
    assert np.array_equal(a, b)
    assert (a[:, 0] != 0xFFFFFFFF).mean() > 0.05
