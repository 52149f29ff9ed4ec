"""Batched inverse rendering of a mesh's shape from WARDiffPathTracer vertex gradients.

The reference's shape-optimization driver (scripts/inv-rendering/shape_optimization/
shape_example.py, run_shape_optimization.py, diff_render_module.py) renders one view per
iteration, takes MeshPosition / MeshNormal gradients from render_grad and steps the mesh with the
largesteps optimizer (lambda_value 60). Here every iteration renders all views in one batched
primal launch (fhp_pass_execute_batch) and back-propagates all of them in one batched adjoint
(fhp_pass_backward_batch); the mesh is updated on the GPU (vertex upload and BVH refit).

The unknown is a per-vertex displacement d along the initial vertex normals n0 (positions
p0 + d n0, normals recomputed from the displaced surface, area-weighted, in torch, so the
MeshNormal gradients flow back through them). As in largesteps (Nicolet et al. 2021, "Large Steps
in Inverse Rendering of Geometry"), the optimizer works on u = (I + lambda L) d, L the mesh's
combinatorial Laplacian, which smooths the noisy per-vertex gradients; d is recovered from u by
conjugate gradients.
"""
from __future__ import annotations

import numpy as np

from . import scene as S
from .render_graph import Scene, SceneGradients, WARDiffPathTracer


def _torch():
    import torch
    return torch


def vertex_normals_torch(pos, idx):
    """Area-weighted vertex normals of positions pos [V, 3] and triangles idx [T, 3] (torch,
    differentiable), as scene.vertex_normals computes them."""
    torch = _torch()
    a, b, c = pos[idx[:, 0]], pos[idx[:, 1]], pos[idx[:, 2]]
    fn = torch.cross(b - a, c - a, dim=1)
    n = torch.zeros_like(pos)
    for k in range(3):
        n = n.index_add(0, idx[:, k], fn)
    return n / torch.clamp(torch.linalg.norm(n, dim=1, keepdim=True), min=1e-30)


def mesh_laplacian(num_vertices: int, idx: np.ndarray, device):
    """Combinatorial Laplacian (degree - adjacency) of the triangle mesh, as a torch sparse matrix."""
    torch = _torch()
    e = np.concatenate([idx[:, [0, 1]], idx[:, [1, 2]], idx[:, [2, 0]]])
    e = np.unique(np.sort(e, axis=1), axis=0)
    i = np.concatenate([e[:, 0], e[:, 1]])
    j = np.concatenate([e[:, 1], e[:, 0]])
    deg = np.bincount(i, minlength=num_vertices).astype(np.float32)
    rows = np.concatenate([i, np.arange(num_vertices)])
    cols = np.concatenate([j, np.arange(num_vertices)])
    vals = np.concatenate([-np.ones(i.shape[0], np.float32), deg])
    L = torch.sparse_coo_tensor(torch.from_numpy(np.stack([rows, cols])).long(), torch.from_numpy(vals),
                                (num_vertices, num_vertices)).coalesce()
    return L.to(device), torch.from_numpy(deg).to(device)


def _cg(apply, b, x0, diag, iters=60, tol=1e-7):
    """Jacobi-preconditioned conjugate gradients for the SPD system apply(x) = b."""
    torch = _torch()
    x = x0.clone()
    r = b - apply(x)
    z = r / diag
    p = z.clone()
    rz = torch.dot(r, z)
    bn = torch.linalg.norm(b).clamp(min=1e-30)
    for _ in range(iters):
        if torch.linalg.norm(r) <= tol * bn:
            break
        Ap = apply(p)
        alpha = rz / torch.dot(p, Ap).clamp(min=1e-30)
        x = x + alpha * p
        r = r - alpha * Ap
        z = r / diag
        rz_new = torch.dot(r, z)
        p = z + (rz_new / rz.clamp(min=1e-30)) * p
        rz = rz_new
    return x


class BatchedShapeRenderer:
    """Renders views of a scene whose mesh `mesh_id` takes given vertex positions and normals, and
    back-propagates an image loss of all views to them (torch.autograd.Function semantics: the
    gradients are render_grad's, the raw adjoint sums divided by spp_grad)."""

    def __init__(self, scene: Scene, mesh_id: int, cameras: list, width: int, height: int, spp_primal: int = 8,
                 spp_grad: int = 8, props: dict | None = None, hash_size: int = 128):
        self.scene, self.mesh_id = scene, mesh_id
        self.cameras = list(cameras)
        self.width, self.height = width, height
        self.spp_primal, self.spp_grad = spp_primal, spp_grad
        base = {"maxBounces": 0, "useWAR": True, "diffMode": "BackwardDiff", "auxSampleCount": 16,
                "Log10vMFConcentration": 3, "Log10vMFConcentrationScreen": 3}
        self.pass_ = WARDiffPathTracer(dict(base, **(props or {})))
        self.pass_.set_scene(scene)
        self.pass_.set_mesh_to_optimize(mesh_id)
        self.grads = SceneGradients(scene, hash_size=hash_size, texture_grads=False, mesh_id=mesh_id)
        self.pass_.scene_gradients = self.grads
        self.frame = 0
        self._fn = self._make_fn()

    def render(self, spp: int | None = None, frame: int | None = None):
        """All views, [num_views, H, W, 3]."""
        torch = _torch()
        f = self.frame if frame is None else frame
        views = [dict(scene=self.scene, width=self.width, height=self.height, spp=spp or self.spp_primal,
                      frame=f * len(self.cameras) + i, camera=c) for i, c in enumerate(self.cameras)]
        return torch.stack([o[..., :3] for o in self.pass_.render_batch(views)])

    def backward(self, dL_dI, frame: int):
        """Raw adjoint sums of sum_views dL_dI[v] . image_v into the vertex gradients."""
        g = self.grads
        g.clear_all_grads()
        views = [dict(dL_dI=dL_dI[i], frame=frame * len(self.cameras) + i, camera=c) for i, c in enumerate(self.cameras)]
        self.pass_.backward_batch(views, self.width, self.height, self.spp_grad)
        g.aggregate_all_grads()
        return (g.get_grads_buffer(g.MESH_POSITION).view(-1, 3).clone(),
                g.get_grads_buffer(g.MESH_NORMAL).view(-1, 3).clone())

    def _make_fn(self):
        torch = _torch()
        r = self

        class _Render(torch.autograd.Function):
            @staticmethod
            def forward(ctx, positions, normals):
                r.scene.set_mesh_vertices(r.mesh_id, positions.detach().contiguous(), normals.detach().contiguous())
                img = r.render(frame=r.frame)
                r.frame += 1
                return img

            @staticmethod
            def backward(ctx, grad):
                gp, gn = r.backward(grad.contiguous(), frame=r.frame)
                r.frame += 1
                inv = 1.0 / r.spp_grad  # render_grad divides the raw sum by spp (common.py:111-112)
                return gp * inv, gn * inv

        return _Render

    def __call__(self, positions, normals):
        return self._fn.apply(positions, normals)


class DisplacementRecovery:
    """Recovers a displacement field of a mesh along its initial normals from target views.

    target: the scene with the true shape (only used to render the target images);
    init: the same scene with the initial shape (same mesh topology). Adam on u = (I + lam L) d
    with the step size decayed by lr_decay per iteration.
    """

    def __init__(self, target: S.SceneData, init: S.SceneData, mesh_name: str = "Sphere", cameras=None,
                 width: int = 256, height: int = 256, spp_target: int = 64, spp_primal: int = 8, spp_grad: int = 8,
                 props: dict | None = None, lr: float = 5e-4, lam: float = 20.0, lr_decay: float = 0.99,
                 hash_size: int = 128):
        torch = _torch()
        dev = torch.device("cuda")
        self.mesh = init.mesh_names.index(mesh_name)
        off = init.mesh_vertex_offsets
        v0, v1 = int(off[self.mesh]), int(off[self.mesh + 1])
        self.cameras = cameras if cameras is not None else S.orbit_cameras(32)
        idx = np.asarray(init.indices).reshape(-1, 3)
        tri = idx[(idx >= v0).all(1) & (idx < v1).all(1)] - v0
        self.idx = torch.from_numpy(tri.astype(np.int64)).to(dev)
        self.p0 = torch.from_numpy(np.asarray(init.positions[v0:v1], np.float32)).to(dev)
        self.n0 = vertex_normals_torch(self.p0, self.idx)
        tp = np.asarray(target.positions[v0:v1], np.float64)
        # true displacement along n0 (target and init share the mesh's parametrization)
        self.d_true = torch.from_numpy(((tp - self.p0.cpu().numpy()) * self.n0.cpu().numpy()).sum(1).astype(np.float32)).to(dev)
        tsc = Scene(target)
        tr = BatchedShapeRenderer(tsc, self.mesh, self.cameras, width, height, spp_primal=spp_target, props=props)
        self.targets = tr.render(spp=spp_target, frame=100000).clone()
        del tr
        self.scene = Scene(init)
        self.renderer = BatchedShapeRenderer(self.scene, self.mesh, self.cameras, width, height, spp_primal, spp_grad,
                                             props, hash_size)
        nv = v1 - v0
        self.L, deg = mesh_laplacian(nv, tri, dev)
        self.lam = lam
        self.diag = 1.0 + lam * deg
        self.u = torch.zeros(nv, device=dev, requires_grad=True)
        self.opt = torch.optim.Adam([self.u], lr=lr)
        # multiplicative step-size decay per iteration: the gradients are Monte Carlo estimates, so
        # late steps are shortened to settle instead of wandering at the noise level
        self.sched = torch.optim.lr_scheduler.ExponentialLR(self.opt, gamma=lr_decay)
        self._d = torch.zeros(nv, device=dev)
        self.history: list[dict] = []

    def _apply(self, x):
        torch = _torch()
        return x + self.lam * torch.sparse.mm(self.L, x[:, None])[:, 0]

    def displacement(self):
        """d = (I + lam L)^-1 u (conjugate gradients, warm-started from the last solution)."""
        self._d = _cg(self._apply, self.u.detach(), self._d, self.diag)
        return self._d

    def positions_normals(self, d):
        p = self.p0 + d[:, None] * self.n0
        return p, vertex_normals_torch(p, self.idx)

    def loss(self, img):
        return ((img - self.targets) ** 2).mean()

    def step(self) -> dict:
        torch = _torch()
        self.opt.zero_grad()
        d0 = self.displacement()
        d = d0.detach().requires_grad_(True)
        p, n = self.positions_normals(d)
        img = self.renderer(p, n)
        loss = self.loss(img)
        loss.backward()
        # chain through d = A^-1 u: dL/du = A^-1 dL/dd (A symmetric)
        gu = _cg(self._apply, d.grad.detach(), torch.zeros_like(d0), self.diag)
        self.u.grad = gu
        self.opt.step()
        self.sched.step()
        rec = dict(loss=float(loss.detach()), rmse=self.rmse(d0))
        self.history.append(rec)
        return rec

    def rmse(self, d=None) -> float:
        """Relative RMS error of the displacement: |d - d_true| / |d_true|."""
        torch = _torch()
        d = self._d if d is None else d
        return float(torch.linalg.norm(d - self.d_true) / torch.linalg.norm(self.d_true))

    def gradient_fd_check(self, direction=None, eps: float = 5e-4, spp: int = 64, frame: int = 50000):
        """Relative error of the adjoint's directional derivative of sum dLdI . I along a
        displacement direction against central differences of the batched primal with common
        random numbers (the adjoint's own seeds), dLdI being the loss gradient at the current shape
        (from a spp-sample render). direction: a per-vertex displacement field, "true" (the
        remaining error d_true - d, the default), or "uniform" (inflation); scaled to unit RMS.
        Returns (adjoint, fd, rel)."""
        torch = _torch()
        d0 = self.displacement().detach()
        p, n = self.positions_normals(d0)
        r = self.renderer
        self.scene.set_mesh_vertices(self.mesh, p.detach().contiguous(), n.detach().contiguous())
        img = r.render(spp=spp, frame=frame)
        dLdI = 2.0 * (img - self.targets) / img.numel()
        if direction is None or (isinstance(direction, str) and direction == "true"):
            direction = self.d_true - d0
        elif isinstance(direction, str) and direction == "uniform":
            direction = torch.ones_like(d0)
        direction = direction / direction.pow(2).mean().sqrt()
        # adjoint: d/dd (dLdI . I) projected on the direction, through positions and normals
        d = d0.clone().requires_grad_(True)
        pp, nn = self.positions_normals(d)
        gp, gn = r.backward(dLdI, frame=frame + 1)
        ((pp * (gp / r.spp_grad)).sum() + (nn * (gn / r.spp_grad)).sum()).backward()
        ad = float((d.grad * direction).sum())
        vals = []
        for sgn in (1.0, -1.0):
            q, m = self.positions_normals(d0 + sgn * eps * direction)
            self.scene.set_mesh_vertices(self.mesh, q.detach().contiguous(), m.detach().contiguous())
            vals.append(float((r.render(spp=r.spp_grad, frame=frame + 1).double() * dLdI.double()).sum()))
        self.scene.set_mesh_vertices(self.mesh, p.detach().contiguous(), n.detach().contiguous())
        fd = (vals[0] - vals[1]) / (2 * eps)
        return ad, fd, abs(ad - fd) / max(abs(fd), 1e-30)
