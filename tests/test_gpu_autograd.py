"""torch.autograd binding: gradients equal the direct adjoint pass on the same seeds, and a short
material optimization reduces the image loss."""
import numpy as np
import pytest

from falcor_amd import scene as S

pytestmark = pytest.mark.gpu


def test_autograd_material_gradients_equal_direct_pass(gpu):
    import torch
    from falcor_amd import autograd as A
    from falcor_amd.render_graph import Scene, SceneGradients, WARDiffPathTracer
    data = S.closed_box(1)
    sc = Scene(data)
    r = A.DiffRenderer(sc, 48, 48, spp_primal=4, spp_grad=4, props={"maxBounces": 2}, hash_size=16, seed=10)
    params = torch.tensor(A.material_params(sc.materials), device="cuda", requires_grad=True)
    img = r(materials=params)
    target = torch.zeros_like(img)
    loss = ((img - target) ** 2).mean()
    loss.backward()
    # Direct: same parameters, adjoint frame 11 (the renderer's second frame), dL/dI of the loss.
    p = WARDiffPathTracer({"maxBounces": 2, "diffMode": "BackwardDiff"})
    p.set_scene(sc)
    g = SceneGradients(sc, hash_size=16, texture_grads=False)
    p.scene_gradients = g
    p.dL_dI = (2.0 / img.numel()) * img.detach()
    g.clear_all_grads()
    p.backward(48, 48, 4, frame=11)
    g.aggregate_all_grads()
    direct = g.get_grads_buffer(g.MATERIAL).view(-1, A.PARAM_COUNT) / 4
    assert torch.allclose(params.grad, direct, rtol=1e-4, atol=1e-7)
    assert float(params.grad[:, 0:3].abs().sum()) > 0


def test_material_optimization_recovers_albedos(gpu):
    """Adam on the walls' and boxes' albedos from grey toward a target image (the light fixed):
    the albedo error falls (the image loss itself is dominated by Monte Carlo noise)."""
    import torch
    from falcor_amd import autograd as A
    from falcor_amd.render_graph import Scene
    data = S.closed_box(2)
    sc = Scene(data)
    ref = A.DiffRenderer(sc, 32, 32, spp_primal=256, spp_grad=4, props={"maxBounces": 2}, hash_size=16)
    true_p = torch.tensor(A.material_params(sc.materials), device="cuda")
    with torch.no_grad():
        target = ref(materials=true_p)
    init = true_p.clone()
    init[:7, 0:3] = 0.5  # walls and boxes start grey
    params = init.clone().requires_grad_(True)
    r = A.DiffRenderer(sc, 32, 32, spp_primal=32, spp_grad=32, props={"maxBounces": 2}, hash_size=16, seed=100)
    opt = torch.optim.Adam([params], lr=0.03)
    errs = []
    for it in range(30):
        opt.zero_grad()
        loss = ((r(materials=params) - target) ** 2).mean()
        loss.backward()
        params.grad[7:] = 0  # keep the light fixed
        opt.step()
        with torch.no_grad():
            params[:, 0:3].clamp_(0.0, 1.0)
        errs.append(float((params.detach()[:7, 0:3] - true_p[:7, 0:3]).abs().mean()))
    print("albedo error", [round(x, 4) for x in errs[::5]], errs[-1], flush=True)
    assert errs[-1] < 0.6 * float((init[:7, 0:3] - true_p[:7, 0:3]).abs().mean())


def test_autograd_vertex_gradients_equal_direct_pass(gpu):
    import torch
    from falcor_amd import autograd as A
    from falcor_amd.render_graph import Scene, SceneGradients, WARDiffPathTracer
    data = S.displaced_sphere_in_box(0, radius=0.12, segments=(24, 12), sigma=0.05)
    sc = Scene(data)
    mesh = data.mesh_names.index("Sphere")
    props = {"maxBounces": 0, "useWAR": True}
    r = A.DiffRenderer(sc, 32, 32, spp_primal=4, spp_grad=4, props=props, mesh_id=mesh, hash_size=16,
                       texture_grads=False, seed=20)
    pos0, nrm0, _ = sc.get_mesh_vertices_and_indices(mesh)
    pos = torch.tensor(pos0, device="cuda", requires_grad=True)
    nrm = torch.tensor(nrm0, device="cuda", requires_grad=True)
    img = r(positions=pos, normals=nrm)
    img.sum().backward()
    p = WARDiffPathTracer(dict(props, diffMode="BackwardDiff"))
    p.set_scene(sc)
    g = SceneGradients(sc, hash_size=16, texture_grads=False, mesh_id=mesh)
    p.scene_gradients = g
    p.set_mesh_to_optimize(mesh)
    p.dL_dI = torch.ones((32, 32, 3), device="cuda")
    g.clear_all_grads()
    p.backward(32, 32, 4, frame=21)
    g.aggregate_all_grads()
    dp = g.get_grads_buffer(g.MESH_POSITION).view(-1, 3) / 4
    dn = g.get_grads_buffer(g.MESH_NORMAL).view(-1, 3) / 4
    assert torch.allclose(pos.grad, dp, rtol=1e-4, atol=1e-6)
    assert torch.allclose(nrm.grad, dn, rtol=1e-4, atol=1e-6)
    assert float(pos.grad.abs().sum()) > 0
