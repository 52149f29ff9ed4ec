"""VBufferRT and MinimalPathTracer's V-buffer input on the GPU against oracle/hitinfo.py and the
traversal oracle.

* PackedHitInfo records are bit-exact: the oracle traces the same primary rays (directions taken
  from the pass's viewW output) and packs the hits with the reference's bit layout.
* viewW matches a float64 camera model, with the jitter of each sample pattern, to 2e-6.
* Depth matches the float64 closed form far (d - near) / ((far - near) d) to 1e-5 (the kernel
  multiplies by the float view-projection matrix).
* MinimalPathTracer renders identical images from its own primary traversal and from a bound
  uncompressed V-buffer (+ viewW).
"""
import numpy as np
import pytest

from falcor_amd import scene as S
from oracle import hitinfo as O

pytestmark = pytest.mark.gpu

W, H = 64, 48


def _scene(comp, data=None):
    from falcor_amd.render_graph import Scene
    data = data if data is not None else S.closed_box(0)
    return data, Scene(data, use_compressed_hit_info=comp)


def _layout(sc):
    i = sc.info()
    return {"compressed": bool(i["hit_info_compressed"]), "type_bits": i["hit_info_type_bits"],
            "inst_bits": i["hit_info_instance_bits"], "prim_bits": i["hit_info_primitive_bits"]}


def _mesh_tri_offsets(data):
    counts = np.bincount(np.asarray(data.tri_mesh), minlength=len(data.mesh_vertex_offsets) - 1)
    return np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)


def _vbuffer(sc, props=None):
    from falcor_amd.render_graph import VBufferRT
    vb = VBufferRT(props or {})
    vb.set_scene(sc)
    return vb


def test_scene_reports_hit_info_layout(gpu):
    data = S.closed_box(0)
    n_mesh = len(data.mesh_vertex_offsets) - 1
    max_prim = int(np.diff(_mesh_tri_offsets(data)).max())
    for comp in (False, True):
        _, sc = _scene(comp, data)
        assert _layout(sc) == O.layout(n_mesh, max_prim, comp)
        assert sc.info()["meshes_contiguous"] == 1


@pytest.mark.parametrize("comp", [False, True])
def test_vbuffer_records_match_oracle(gpu, oracle_mod, comp):
    data, sc = _scene(comp)
    out = _vbuffer(sc).execute_frame(W, H)
    view = out["viewW"].cpu().numpy().reshape(-1, 4)
    d = -view[:, :3]
    cam = data.camera_array().astype(np.float64)
    fwd = (cam[3:6] - cam[0:3]) / np.linalg.norm(cam[3:6] - cam[0:3])
    inv_cos = 1.0 / (d.astype(np.float64) @ fwd)
    rays = np.zeros((W * H, 8), np.float32)
    rays[:, 0:3] = cam[0:3]
    rays[:, 3] = cam[12] * inv_cos
    rays[:, 4:7] = d
    rays[:, 7] = cam[13] * inv_cos
    hits = oracle_mod.OracleScene(data).trace(rays, brute=True)
    tri = hits[:, 0]
    valid = tri != 0xFFFFFFFF
    assert valid.mean() > 0.5  # most primary rays enter the open box
    off = _mesh_tri_offsets(data)
    t = np.where(valid, tri, 0).astype(np.int64)
    inst = np.searchsorted(off, t, side="right") - 1
    prim = t - off[inst]
    L = _layout(sc)
    exp = O.pack_triangle_hits(L, inst, prim, hits[:, 2].view(np.float32), hits[:, 3].view(np.float32), valid)
    got = out["vbuffer"].cpu().numpy().view(np.uint32).reshape(W * H, -1)
    assert got.shape[1] == (2 if comp else 4)
    assert np.array_equal(got, exp)
    # Mask and depth.
    assert np.array_equal(out["mask"].cpu().numpy().ravel(), valid.astype(np.float32))
    idx = np.asarray(data.indices)[t]
    pos = np.asarray(data.positions, np.float64)
    b1 = hits[:, 2].view(np.float32).astype(np.float64)
    b2 = hits[:, 3].view(np.float32).astype(np.float64)
    P = pos[idx[:, 0]] * (1 - b1 - b2)[:, None] + pos[idx[:, 1]] * b1[:, None] + pos[idx[:, 2]] * b2[:, None]
    want = np.where(valid, O.ndc_depth(cam, P), 1.0)
    np.testing.assert_allclose(out["depth"].cpu().numpy().ravel(), want, atol=1e-5)


def test_viewW_and_halton_jitter(gpu):
    data, sc = _scene(False)
    vb = _vbuffer(sc, {"samplePattern": "Halton", "sampleCount": 4})
    cam = data.camera_array()
    for f in range(6):
        out = vb.execute_frame(W, H, outputs=("viewW",))
        hx, hy = O.halton_jitter(f % 4)
        jx, jy = vb.jitter
        assert jx == pytest.approx(hx / W, abs=1e-9) and jy == pytest.approx(hy / H, abs=1e-9), f
        d64 = O.camera_dirs(cam, W, H, jitter=(jx, jy))
        np.testing.assert_allclose(-out["viewW"].cpu().numpy()[..., :3], d64, atol=2e-6)
        assert np.all(out["viewW"].cpu().numpy()[..., 3] == 0)


def test_directx_and_stratified_patterns(gpu):
    _, sc = _scene(False)
    vb = _vbuffer(sc, {"samplePattern": "DirectX"})
    js = []
    for _ in range(16):
        vb.execute_frame(16, 16, outputs=())
        js.append(vb.jitter)
    q = np.round(np.array(js) * 16 * 16).astype(int)  # jitter * frame size * 16: odd sixteenths
    assert np.all(q % 2 == 1) and np.all(np.abs(q) < 8)
    assert len(set(q[:8, 0])) == 8 and len(set(q[:8, 1])) == 8  # one sample per row and column
    assert np.array_equal(q[:8], q[8:])  # period 8
    vb = _vbuffer(sc, {"samplePattern": "Stratified", "sampleCount": 16})
    for rnd in range(2):
        bins = set()
        for _ in range(16):
            vb.execute_frame(8, 8, outputs=())
            x, y = np.array(vb.jitter) * 8 + 0.5
            assert 0 <= x < 1 and 0 <= y < 1
            bins.add((int(x * 4), int(y * 4)))
        assert len(bins) == 16, rnd  # every stratum once per round of 16 frames


def test_mpt_with_vbuffer_matches_internal_primary(gpu):
    import torch
    from falcor_amd.render_graph import MinimalPathTracer
    _, sc = _scene(False)
    out = _vbuffer(sc).execute_frame(W, H)
    for use_graph in (1, 0):
        mpt = MinimalPathTracer({"maxBounces": 3, "useGraph": use_graph})
        mpt.set_scene(sc)
        a = mpt.render(W, H, 4, frame=5).clone()
        mpt.set_input("vbuffer", out["vbuffer"])
        b = mpt.render(W, H, 4, frame=5).clone()
        mpt.set_input("viewW", out["viewW"])
        c = mpt.render(W, H, 4, frame=5).clone()
        mpt.set_input("vbuffer", None)
        mpt.set_input("viewW", None)
        d = mpt.render(W, H, 4, frame=5).clone()
        assert torch.equal(a, b) and torch.equal(a, c) and torch.equal(a, d), use_graph


def test_mpt_with_compressed_vbuffer_is_close(gpu):
    from falcor_amd.render_graph import MinimalPathTracer
    _, sc = _scene(True)
    out = _vbuffer(sc).execute_frame(W, H)
    mpt = MinimalPathTracer({"maxBounces": 3})
    mpt.set_scene(sc)
    a = mpt.render(W, H, 16, frame=1).clone()
    mpt.set_input("vbuffer", out["vbuffer"])
    mpt.set_input("viewW", out["viewW"])
    b = mpt.render(W, H, 16, frame=1).clone()
    rel = float((a - b).abs().sum() / a.abs().sum())
    print(f"compressed V-buffer relative L1 difference {rel:.2e}", flush=True)
    assert rel < 2e-2


def test_invalid_records_read_as_misses(gpu):
    import torch
    from falcor_amd.render_graph import MinimalPathTracer
    data, sc = _scene(False)
    L = _layout(sc)
    out = _vbuffer(sc).execute_frame(W, H)
    vb = out["vbuffer"].clone()
    words = vb.view(-1, 4)
    hit_rows = torch.nonzero(out["mask"].view(-1) > 0).view(-1)[:3].tolist()
    bad = np.zeros((3, 4), np.uint32)
    # None type (all zero); a primitive past the end of mesh 0 (2 triangles); type Volume (2).
    bad[1] = O.pack_triangle_hits(L, [0], [(1 << L["prim_bits"]) - 1], [0.2], [0.2], [True])[0]
    bad[2] = bad[1].copy()
    bad[2, 0] = (2 << (32 - L["type_bits"])) | (bad[2, 0] & ((1 << (32 - L["type_bits"])) - 1))
    assert np.diff(_mesh_tri_offsets(data))[0] < (1 << L["prim_bits"]) - 1
    words[hit_rows] = torch.from_numpy(bad.view(np.int32)).cuda()
    mpt = MinimalPathTracer({"maxBounces": 2})
    mpt.set_scene(sc)
    mpt.set_input("vbuffer", out["vbuffer"])
    ref = mpt.render(W, H, 2, frame=0).clone().view(-1, 4)
    mpt.set_input("vbuffer", vb)
    img = mpt.render(W, H, 2, frame=0).clone().view(-1, 4)
    assert torch.all(img[hit_rows, :3] == 0)  # background (alpha stays 1)
    keep = torch.ones(W * H, dtype=torch.bool, device=img.device)
    keep[hit_rows] = False
    assert torch.equal(img[keep], ref[keep])


def test_reference_graph_vbuffer_mpt_accumulate(gpu):
    """The graph of scripts/MinimalPathTracer.py (ToneMapper aside): VBufferRT (Stratified, 16)
    -> MinimalPathTracer (vbuffer, viewW) -> AccumulatePass."""
    import torch
    from falcor_amd.render_graph import RenderGraph
    _, sc = _scene(False)
    g = RenderGraph("MinimalPathTracer")
    acc = g.create_pass("AccumulatePass", "AccumulatePass", {"enabled": True, "precisionMode": "Single"})
    g.create_pass("MinimalPathTracer", "MinimalPathTracer", {"maxBounces": 3})
    vb = g.create_pass("VBufferRT", "VBufferRT", {"samplePattern": "Stratified", "sampleCount": 16})
    g.add_edge("VBufferRT.vbuffer", "MinimalPathTracer.vbuffer")
    g.add_edge("VBufferRT.viewW", "MinimalPathTracer.viewW")
    g.add_edge("MinimalPathTracer.color", "AccumulatePass.input")
    g.mark_output("AccumulatePass.output")
    jit = []
    for _ in range(4):
        g.execute(sc, W, H)
        jit.append(vb.jitter)
    out = g.get_output("AccumulatePass.output")
    assert acc.frame_count == 4 and len(set(jit)) == 4
    assert torch.isfinite(out).all() and float(out[..., :3].mean()) > 0


def test_vbuffer_errors(gpu):
    import torch
    from falcor_amd import _lib
    from falcor_amd.render_graph import MinimalPathTracer
    data, sc = _scene(False)
    out = _vbuffer(sc).execute_frame(W, H)
    mpt = MinimalPathTracer({})
    mpt.set_scene(sc)
    mpt.set_input("vbuffer", out["vbuffer"])
    with pytest.raises(ValueError, match="vbuffer"):
        mpt.render(W // 2, H, 1)
    mpt.set_input("vbuffer", out["vbuffer"][..., :2].contiguous())
    with pytest.raises(ValueError, match="32-bit words"):
        mpt.render(W, H, 1)
    # Triangles of one mesh split apart: V-buffer encoding is refused.
    idx = np.asarray(data.indices).copy()
    tm = np.asarray(data.tri_material).copy()
    mesh = np.asarray(data.tri_mesh).copy()
    for a in (idx, tm, mesh):
        a[[0, -1]] = a[[-1, 0]]
    from dataclasses import replace
    d2 = replace(data, indices=idx, tri_material=tm, tri_mesh=mesh)
    _, sc2 = _scene(False, d2)
    assert sc2.info()["meshes_contiguous"] == 0
    with pytest.raises(_lib.FhpError, match="contiguous"):
        _vbuffer(sc2).execute_frame(W, H)


def test_testbed_loads_pyscene_and_renders(gpu, tmp_path):
    """Testbed.load_scene with a .pyscene path: the scene renders through MinimalPathTracer."""
    from falcor_amd.render_graph import MinimalPathTracer, Testbed
    (tmp_path / "box.pyscene").write_text(
        "m = StandardMaterial('White')\nm.baseColor = float4(0.7, 0.7, 0.7, 1)\n"
        "lamp = StandardMaterial('Lamp')\nlamp.emissiveColor = float3(10, 10, 10)\n"
        "quad = TriangleMesh.createQuad()\n"
        "sceneBuilder.addMeshInstance(sceneBuilder.addNode('Floor', Transform(scaling=2)), sceneBuilder.addTriangleMesh(quad, m))\n"
        "sceneBuilder.addMeshInstance(sceneBuilder.addNode('Lamp', Transform(translation=float3(0, 1, 0), rotationEulerDeg=float3(180, 0, 0))),"
        " sceneBuilder.addTriangleMesh(quad, lamp))\n"
        "c = Camera()\nc.position = float3(0, 0.5, 2.5)\nc.target = float3(0, 0.3, 0)\nc.up = float3(0, 1, 0)\nc.focalLength = 35\n"
        "sceneBuilder.addCamera(c)\n")
    tb = Testbed(48, 32)
    sc = tb.load_scene(tmp_path / "box.pyscene")
    assert sc.info()["num_triangles"] == 4
    p = MinimalPathTracer({"maxBounces": 2})
    p.set_scene(sc)
    img = p.render(48, 32, 4, frame=0)
    import torch
    assert torch.isfinite(img).all() and float(img[..., :3].mean()) > 0
