#!/usr/bin/env python
"""Throughput benchmark of the fhp path-tracing core on one MI355X.

A step is one forward frame of the sphere-field workload: 1024 icospheres (1,310,720 triangles),
4096 emissive triangles sampled through the light BVH, 1920x1080, 4 spp, max 6 bounces,
MinimalPathTracer with emissive next-event estimation (FP32). `value` is closest-hit plus
visibility rays per second over the timed steps (Mrays/s), with all inputs resident in HBM.

The timed steps run the product path: the captured bounce loop, no per-kernel events and no host
synchronisation inside the timed region. The same frames are rendered untimed beforehand for
their exact ray counts and afterwards with HIP events on the launch stream for kernel durations.

Also reported on the same JSON line:
  * roofline of the dominant kernel (kTraceBounce, the BVH4 traversal of every bounce's
    closest-hit and visibility rays): algorithmic bytes per closest-hit ray (32 B ray + 16 B hit
    + 64 B per visited node + 48 B per tested triangle) and per visibility ray (48 B record +
    nodes + triangles), node/triangle counts measured on the same frame, over the kernel's average
    launch duration; beside it the rocprofv3-measured HBM traffic per launch and L2 hit rate
    (committed profile summary) and the measured HBM fraction of peak;
  * batch: the batched forward workload (64 views of 512x512, spp in {1, 2, 4, 8, 16}, over 16
    sphere fields of 10 k - 2 M triangles) in one launch sequence against one sequence per view;
  * diff: ms per forward+adjoint iteration of the differentiable-materials workload (64
    textured spheres, 512x512 RGB fp32 albedo textures, 512x512 image, 16 spp forward +
    16 spp adjoint), batched vs the plain per-frame loop, at maxBounces 1 (and batched at 4),
    the gradient relative error against central finite differences (eps 1e-3) on 256 texels, and
    100 Adam iterations of the texture / roughness optimization (loss and errors before and after);
  * geom: the batched differentiable-geometry workload (20,000-triangle displaced sphere in the
    closed box, 32 views of 256x256 at 8 spp, 16 auxiliary rays, kappa 1e3): ms per iteration of
    the displacement-recovery loop (one batched primal and one batched adjoint), the gradient's
    relative error against central differences, and the displacement error after the iterations;
  * cpu_baseline: the oracle (C++ restatement, all host cores) on the same forward workload.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--skip-cpu] [--skip-diff] [--skip-batch]
                       [--skip-geom] [--geom-iters N] [--only forward|batch|diff|war-iter|geom]
                       [--tune [--tune-out F]]
       python bench.py --copy-calibration   (1 GiB copies: scale for rocprofv3 FETCH_SIZE/WRITE_SIZE)
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Mrays/s fwd; ms per fwd+adjoint iter at 512² 16spp; grad rel-err vs FD"
HBM_PEAK_GBS = 8000.0
NODE_BYTES, TRI_BYTES, RAY_BYTES, HIT_BYTES, VIS_RECORD_BYTES = 64, 48, 32, 16, 48
TUNE_TAG = "sphere_field_1920x1080_4spp_b6_lightbvh_nee"
BATCH_TUNE_TAG = "batch_64views_16scenes_512x512_b6_lightbvh_nee"


TRAFFIC_PROFILE = ROOT / "profiles" / "kTraceBounce_sphere_field_1920x1080_4spp.md"


def profiled_traffic():
    """HBM bytes per launch of the traversal kernel from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (counters need their own profiler runs, so they are not measured here)."""
    try:
        for line in TRAFFIC_PROFILE.read_text().splitlines():
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":")[1])
    except OSError:
        pass
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def profiled_l2_hit_rate():
    """L2 hit rate over the traversal launches from the committed rocprofv3 TCC_HIT / TCC_MISS summary."""
    try:
        for line in TRAFFIC_PROFILE.read_text().splitlines():
            if line.startswith("l2_hit_rate:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def forward_bench(steps: int, warmup: int, tune: bool, tune_out: str | None):
    """Times `steps` frames of the product path (captured bounce loop, no per-kernel events, no
    host synchronisation inside the timed region). The same frames are rendered untimed before:
    once with ray statistics (exact counts: a frame's rays depend only on its seeds) and once with
    per-launch HIP events on the launch stream (kernel durations for the roofline)."""
    import torch
    from falcor_amd import autotune
    from falcor_amd import scene as S
    from falcor_amd.render_graph import MinimalPathTracer, Scene

    W, H, SPP, B = 1920, 1080, 4, 6
    t0 = time.time()
    data = S.sphere_field(1)
    t1 = time.time()
    sc = Scene(data)
    torch.cuda.synchronize()
    t2 = time.time()
    info = sc.info()
    log(f"scene: {info['num_triangles']} tris, {info['num_active_lights']} lights, bvh {info['bvh_nodes']} nodes "
        f"depth {info['bvh_depth']} stack {info['stack_need']}; build {t2 - t1:.2f}s (geometry {t1 - t0:.2f}s)")
    props = {"maxBounces": B, "useEmissiveNEE": True, "emissiveSampler": "LightBVH"}
    p = MinimalPathTracer(dict(props))
    p.set_scene(sc)
    out = torch.empty((H, W, 4), device="cuda")
    launch = autotune.load("MinimalPathTracer", TUNE_TAG)
    if launch is not None and not tune:
        autotune.apply(p, launch)
    launch = {k: int(p.get_property(k)) for k in autotune.LAUNCH_KEYS}
    if tune:
        # Untimed: pick the traversal launch configuration on this frame shape.
        tr = autotune.tune(p, lambda i: p.render(W, H, SPP, frame=1000 + i, out=out), reps=2, warmup=1)
        launch = dict(tr["best"])
        table = sorted(tr["table"], key=lambda r: r[1])
        log("autotune (fastest first): " + "; ".join(
            f"{c['traceBlockSize']}/{c['traceStackDepth']}/{c['traceBlocksPerCU']}: {ms:.3f} ms"
            for c, ms in table))
        if tune_out:
            autotune.save("MinimalPathTracer", TUNE_TAG, launch, path=tune_out)
        autotune.apply(p, launch)
    frames = [100 + i for i in range(steps)]
    for i in range(warmup):
        p.render(W, H, SPP, frame=10 + i, out=out)
    # Ray counts of the timed frames (untimed).
    rays_c = rays_v = tail_c = tail_v = 0
    for f in frames:
        _, st = p.render(W, H, SPP, frame=f, out=out, stats=True)
        rays_c += int(st[0])
        rays_v += int(st[1])
        tail_c += int(st[6])  # of those, traced by the bounce loop's tail launch
        tail_v += int(st[7])
    torch.cuda.synchronize()
    ts = time.perf_counter()
    for f in frames:
        p.render(W, H, SPP, frame=f, out=out)
    torch.cuda.synchronize()
    te = time.perf_counter()
    sec = te - ts
    mrays = (rays_c + rays_v) / sec / 1e6
    # Kernel durations of the same frames (untimed): HIP events around every launch on its stream.
    pt = MinimalPathTracer(dict(props, timeKernels=1, **launch))
    pt.set_scene(sc)
    pt.render(W, H, SPP, frame=10, out=out)
    pt.kernel_times(reset=True)
    for f in frames:
        pt.render(W, H, SPP, frame=f, out=out)
    kt = pt.kernel_times(reset=True)
    # Traversal counts of one frame (separate, untimed run).
    pc = MinimalPathTracer(dict(props, countTraversal=1, **launch))
    pc.set_scene(sc)
    _, cs = pc.render(W, H, SPP, frame=0, out=out, stats=True)
    n_node = cs[2] / max(cs[0], 1)
    n_tri = cs[3] / max(cs[0], 1)
    s_node = cs[4] / max(cs[1], 1)
    s_tri = cs[5] / max(cs[1], 1)
    # Algorithmic bytes of the traversal kernel (kTraceBounce: primary, closest-hit and visibility
    # rays of every bounce up to the tail launch): each closest-hit ray reads its 32 B ray and
    # writes a 16 B hit, each visibility ray reads its 48 B record, and every visited node /
    # tested triangle is read once (per-ray averages of the whole frame).
    bytes_per_ray = RAY_BYTES + HIT_BYTES + NODE_BYTES * n_node + TRI_BYTES * n_tri
    bytes_per_vis = VIS_RECORD_BYTES + NODE_BYTES * s_node + TRI_BYTES * s_tri
    trace_ms, trace_launches = kt["trace"]
    trace_bytes = bytes_per_ray * (rays_c - tail_c) + bytes_per_vis * (rays_v - tail_v)
    achieved = trace_bytes / (trace_ms * 1e-3) / 1e9 if trace_ms > 0 else 0.0
    img = out[..., :3].float()
    res = dict(
        mrays=mrays, ms_per_step=sec / steps * 1e3, rays_closest=rays_c / steps, rays_visibility=rays_v / steps,
        kernel_ms=kt, nodes_per_ray=float(n_node), tris_per_ray=float(n_tri), bytes_per_ray=float(bytes_per_ray),
        bytes_per_vis=float(bytes_per_vis), bytes_per_launch=trace_bytes / max(trace_launches, 1),
        trace_launches_per_step=trace_launches / steps, trace_avg_launch_ms=trace_ms / max(trace_launches, 1),
        achieved_gbs=achieved, image_mean=float(img.mean()), build_s=t2 - t1, info=info, W=W, H=H, spp=SPP, B=B,
        data=data, launch=launch, shadow_nodes_per_ray=float(s_node), shadow_tris_per_ray=float(s_tri),
        tail_rays_closest=tail_c / steps, tail_rays_visibility=tail_v / steps,
        tail_ms_per_step=kt["other"][0] / steps)
    log(f"forward: {mrays:.1f} Mrays/s, {res['ms_per_step']:.2f} ms/frame, {rays_c / steps / 1e6:.2f}M closest + "
        f"{rays_v / steps / 1e6:.2f}M visibility rays/frame; nodes/ray {n_node:.1f} tris/ray {n_tri:.1f} "
        f"(visibility {s_node:.1f} / {s_tri:.1f}); launch {launch}; kernels {kt}")
    errs = sc.trace_errors()
    if errs:
        raise RuntimeError(f"traversal reported errors {errs:#x}")
    return res


def batch_bench(reps: int = 2, tune: bool = False, tune_out: str | None = None):
    """Batched forward: 64 views (512x512, spp in {1,2,4,8,16}, 6 bounces, LightBVH NEE) over 16
    sphere fields of 10k-2M triangles, rendered in one launch sequence (captured bounce loop)
    with the traversal launch the autotuner chose for this workload, against the same views
    rendered one after the other (one launch sequence per view)."""
    import torch
    from falcor_amd import autotune
    from falcor_amd import scene as S
    from falcor_amd.render_graph import MinimalPathTracer, Scene

    t0 = time.time()
    datas, views = S.batch_workload()
    scenes = [Scene(d) for d in datas]
    torch.cuda.synchronize()
    build = time.time() - t0
    vs = [dict(v, scene=scenes[v["scene"]]) for v in views]
    for v in vs:
        v["out"] = torch.empty((v["height"], v["width"], 4), device="cuda")
    p = MinimalPathTracer({"maxBounces": 6, "useEmissiveNEE": True, "emissiveSampler": "LightBVH"})
    launch = autotune.load("MinimalPathTracer", BATCH_TUNE_TAG)
    if tune:
        # Untimed: every traversal launch configuration on the whole batch.
        tr = autotune.tune(p, lambda i: p.render_batch(vs), reps=2, warmup=1)
        launch = dict(tr["best"])
        table = sorted(tr["table"], key=lambda r: r[1])
        log("batch autotune (fastest first): " + "; ".join(
            f"{c['traceBlockSize']}/{c['traceStackDepth']}/{c['traceBlocksPerCU']}: {ms:.2f} ms" for c, ms in table))
        if tune_out:
            autotune.save("MinimalPathTracer", BATCH_TUNE_TAG, launch, path=tune_out)
    if launch is not None:
        autotune.apply(p, launch)
    launch = {k: int(p.get_property(k)) for k in autotune.LAUNCH_KEYS}
    _, st = p.render_batch(vs, stats=True)  # warm-up; also allocates the batch buffers
    rays = int(st[0]) + int(st[1])
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        p.render_batch(vs)
    torch.cuda.synchronize()
    batched = (time.perf_counter() - t) / reps * 1e3
    q = MinimalPathTracer({"maxBounces": 6, "useEmissiveNEE": True, "emissiveSampler": "LightBVH"})
    autotune.apply(q, launch)
    for v in vs:
        q.render_batch([v])
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        for v in vs:
            q.render_batch([v])
    torch.cuda.synchronize()
    seq = (time.perf_counter() - t) / reps * 1e3
    paths = sum(v["width"] * v["height"] * v["spp"] for v in views)
    log(f"batch: {len(views)} views over {len(datas)} scenes ({sum(d.num_triangles for d in datas)} tris, build "
        f"{build:.1f}s), {paths / 1e6:.1f}M paths, {rays / 1e6:.1f}M rays: batched {batched:.1f} ms "
        f"({rays / batched / 1e3:.1f} Mrays/s), per-view {seq:.1f} ms ({rays / seq / 1e3:.1f} Mrays/s)")
    return dict(views=len(views), scenes=len(datas), paths=paths, rays=rays, ms_batched=batched, ms_per_view_loop=seq,
                mrays_batched=rays / batched / 1e3, speedup=seq / batched, traversal_launch=launch)


def cpu_baseline(data, W, H, spp, B, budget_s: float = 12.0, max_frames: int = 12):
    """The oracle (C++ restatement of the same algorithms, scalar BVH traversal) on the host cores:
    whole frames of the forward workload until about budget_s of wall time."""
    import oracle
    osc = oracle.OracleScene(data)
    threads = os.cpu_count() or 1
    rays = 0
    frames = 0
    t = time.perf_counter()
    while frames < max_frames and (frames == 0 or time.perf_counter() - t < budget_s):
        _, st = osc.render_mpt(W, H, spp=spp, frame=frames, max_bounces=B, nee=True, light_sampler=1, threads=threads)
        rays += int(st[0]) + int(st[1])
        frames += 1
    dt = time.perf_counter() - t
    return dict(value=rays / dt / 1e6, unit="Mrays/s", cores=threads, kind="port",
                sample=f"{frames} {W}x{H} frames of the same workload ({spp} spp, {B} bounces, LightBVH NEE), "
                       f"oracle C++ restatement on a std::thread pool of {threads} threads, {dt:.1f} s")


def diff_bench(iters: int = 3, fd_texels: int = 256, bounces=(1, 4), plain_loop: bool = True):
    """ms per forward + adjoint iteration of the differentiable-materials workload (64 textured
    spheres, 512x512 RGB albedo textures, 512x512 image, 16 spp forward + 16 spp adjoint, L2 image
    loss): one 16-spp launch each way ("batched") against the reference's per-frame loop of 1-spp
    launches with a host synchronisation per frame ("plain"), at maxBounces 1 and 4 (the
    reference material driver's setting, sphere_materials_example.py:47). Timed without per-kernel
    events; kernel times come from a separate untimed iteration. Then the gradient's relative
    error against central finite differences on 256 texels."""
    import torch
    from falcor_amd import scene as S
    from falcor_amd.render_graph import Scene, SceneGradients, WARDiffPathTracer

    W = H = 512
    SPP = 16
    tgt = S.textured_spheres(3, target=True)
    init = S.textured_spheres(3, target=False)
    sc_t = Scene(tgt)
    sc = Scene(init)
    n_pix = W * H * 3
    res = {}
    kt = {}
    for B in bounces:
        props = {"maxBounces": B, "useWAR": False}
        pt = WARDiffPathTracer(props)
        pt.set_scene(sc_t)
        target = pt.render(W, H, SPP, frame=1000)[..., :3].clone()
        p = WARDiffPathTracer(dict(props, diffMode="BackwardDiff"))
        p.set_scene(sc)
        g = SceneGradients(sc, hash_size=256, texture_grads=True)
        p.scene_gradients = g

        def batched(frame, p=p, g=g, target=target):
            img = p.render(W, H, SPP, frame=2 * frame)[..., :3]
            p.dL_dI = (2.0 / n_pix) * (img - target)
            g.clear_all_grads()
            p.backward(W, H, SPP, frame=2 * frame + 1)
            g.aggregate_all_grads()
            return img

        def plain(frame, p=p, g=g, target=target):
            # The reference's per-frame loop (scripts/inv-rendering/common.py:80-114): one launch per
            # sample, host synchronisation after each frame, forward then adjoint.
            acc = None
            for k in range(SPP):
                img = p.render(W, H, 1, frame=frame * 2 * SPP + k)[..., :3]
                torch.cuda.synchronize()
                acc = img.clone() if acc is None else acc + img
            img = acc * (1.0 / SPP)
            p.dL_dI = (2.0 / n_pix) * (img - target)
            g.clear_all_grads()
            for k in range(SPP):
                p.backward(W, H, 1, frame=frame * 2 * SPP + SPP + k)
                torch.cuda.synchronize()
            g.aggregate_all_grads()
            return img

        fns = (("batched", batched), ("plain", plain)) if B == 1 and plain_loop else (("batched", batched),)
        for _, fn in fns:
            fn(0)
        torch.cuda.synchronize()
        for name, fn in fns:
            t = time.perf_counter()
            for i in range(iters):
                fn(i + 1)
            torch.cuda.synchronize()
            res[f"{name}_b{B}"] = (time.perf_counter() - t) / iters * 1e3
        # Kernel times of one batched iteration (untimed): HIP events on the launch stream.
        p.set_property("timeKernels", 1)
        p.kernel_times(reset=True)
        batched(iters + 1)
        kt[f"b{B}"] = p.kernel_times(reset=True)
        p.set_property("timeKernels", 0)
        if B == 1:
            keep = (p, g, target)
    p, g, target = keep
    log(f"diff: ms/iter {res}; kernels {kt}")
    if not fd_texels:
        return dict(res, kernel_ms=kt)

    # Gradient check: central differences of L = sum(dLdI * I) on random texels with nonzero
    # gradient, sampling frozen at the current parameters (the function the adjoint differentiates).
    img = p.render(W, H, SPP, frame=7)[..., :3]
    dLdI = ((2.0 / n_pix) * (img - target)).contiguous()
    p.dL_dI = dLdI
    g.clear_all_grads()
    p.backward(W, H, SPP, frame=7)
    g.aggregate_all_grads()
    gtex = g.get_grads_buffer(g.TEXTURE) / SPP
    cand = torch.nonzero(gtex.abs() > 1e-3 * gtex.abs().max()).flatten().cpu().numpy()
    rng = np.random.default_rng(0)
    pick = rng.choice(cand, size=min(fd_texels, cand.size), replace=False)
    texels = torch.from_numpy(init.texels).cuda()
    sc.freeze_sampling(True)
    eps = 1e-3
    fd = []
    t = time.perf_counter()
    d64 = dLdI.double()
    for ti in pick:
        vals = []
        for s in (eps, -eps):
            tx = texels.clone()
            tx[int(ti)] += s
            sc.set_texels(tx)
            im = p.render(W, H, SPP, frame=7)[..., :3]
            vals.append(float((im.double() * d64).sum()))
        fd.append((vals[0] - vals[1]) / (2 * eps))
    sc.set_texels(texels)
    sc.freeze_sampling(False)
    torch.cuda.synchronize()
    ad = gtex[torch.from_numpy(pick).cuda()].double().cpu().numpy()
    fd = np.array(fd)
    rel = float(np.linalg.norm(ad - fd) / max(np.linalg.norm(fd), 1e-30))
    log(f"diff: grad rel-err vs FD on {pick.size} texels = {rel:.2e} ({time.perf_counter() - t:.1f}s)")
    loop = material_loop()
    return dict(loop=loop, ms_per_iter_batched=res["batched_b1"], ms_per_iter_plain=res["plain_b1"],
                speedup_batched_vs_plain=res["plain_b1"] / res["batched_b1"], ms_per_iter_batched_b4=res["batched_b4"],
                grad_rel_err_fd=rel, fd_texels=int(pick.size), kernel_ms=kt)


def material_loop(iters: int = 100) -> dict:
    """The differentiable-materials workload's optimization: `iters` Adam iterations on the 64
    albedo textures and roughness values (falcor_amd.material_optimization), with the loss and the
    errors of the texels seen in the first iteration and of the roughness before and after."""
    import torch
    from falcor_amd import scene as S
    from falcor_amd.material_optimization import TextureRecovery

    rec = TextureRecovery(S.textured_spheres(3, target=True), S.textured_spheres(3, target=False))
    rec.step()
    seen = rec.seen.clone()
    e1 = rec.errors(seen)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(iters - 1):
        rec.step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t) / max(iters - 1, 1) * 1e3
    e = rec.errors(seen)
    out = dict(iters=iters, ms_per_iter=ms, loss_start=rec.history[0]["loss"], loss_end=rec.history[-1]["loss"],
               seen_texels=float(seen.float().mean()), seen_texel_rmse_start=e1["seen_texel_rmse"],
               seen_texel_rmse_end=e["seen_texel_rmse"], roughness_rmse_start=e1["roughness_rmse"],
               roughness_rmse_end=e["roughness_rmse"])
    log(f"diff loop: {json.dumps(out)}")
    return out


def geom_bench(iters: int = 20):
    """Batched differentiable geometry: a 20,000-triangle sphere (seeded normal displacement,
    sigma 0.05 radius) in the closed box, 32 views from inside the box (orbit cameras), 256x256,
    8 spp forward and 8 spp adjoint, vertex gradients by warped-area reparameterization with 16
    auxiliary rays (vMF kappa 1e3): one batched primal launch and one batched adjoint per
    iteration of the displacement-recovery loop (falcor_amd.shape_optimization). Reports ms per
    iteration, the adjoint's directional derivative against central differences of the batched
    primal at the initial shape (uniform inflation and the remaining-error direction, with the
    reference's NEE term and with the held NEE direction), and the relative displacement error
    after `iters` iterations."""
    import torch
    from falcor_amd import scene as S
    from falcor_amd.shape_optimization import DisplacementRecovery

    tgt = S.displaced_sphere_in_box(0, radius=0.12, segments=(100, 100), sigma=0.05, front_wall=True)
    ini = S.displaced_sphere_in_box(0, radius=0.12, segments=(100, 100), sigma=0.0, front_wall=True)
    out = {"views": 32, "width": 256, "height": 256, "spp": 8, "aux": 16, "kappa": 1000.0,
           "sphere_triangles": 20000, "scene_triangles": int(np.asarray(ini.indices).size // 3)}
    fd = {}
    for name, props in (("reference_nee", {}), ("held_nee", {"useSolidAngleNEEWarp": 1})):
        rec = DisplacementRecovery(tgt, ini, props=props, lr=5e-4, lam=20.0, lr_decay=0.99)
        for dname in ("uniform", "true"):
            ad, f, rel = rec.gradient_fd_check(direction=dname)
            fd[f"{name}_{dname}"] = {"adjoint": ad, "fd": f, "rel": rel}
        if name == "held_nee":
            keep = rec
    rec = keep
    rec.step()  # warm-up (allocations)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(iters):
        rec.step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t) / iters * 1e3
    out.update(ms_per_iter=ms, iters=iters + 1, rmse_start=1.0, rmse_end=rec.rmse(rec.displacement()),
               loss_start=rec.history[0]["loss"], loss_end=rec.history[-1]["loss"], fd=fd)
    log(f"geom: {ms:.1f} ms/iter (32 views 256x256 8 spp fwd + adjoint), displacement error 1.0 -> "
        f"{out['rmse_end']:.3f} after {iters + 1} iterations; fd {json.dumps(fd)}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-diff", action="store_true")
    ap.add_argument("--tune", action="store_true",
                    help="autotune the traversal launch on this workload (untimed) instead of using the tuned table")
    ap.add_argument("--tune-out", default=None, help="with --tune: JSON file to record the chosen launch in")
    ap.add_argument("--skip-batch", action="store_true",
                    help="skip the batched forward workload (64 views over 16 scenes)")
    ap.add_argument("--skip-geom", action="store_true",
                    help="skip the batched differentiable-geometry workload (32 views, vertex gradients)")
    ap.add_argument("--geom-iters", type=int, default=20, help="displacement-recovery iterations of the geometry leg")
    ap.add_argument("--only", choices=("forward", "batch", "diff", "war-iter", "geom"), default=None,
                    help="run one leg only and print its result (profiling runs); no JSON line")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device-to-device copy 3 times (scale for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU only (the path shards by views/tiles with no exchange "
              "except image gather and one gradient all-reduce; see DESIGN.md).", file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        sys.exit(1)
    if args.copy_calibration:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("copy calibration: 3 x 1 GiB device-to-device copies")
        return
    from falcor_amd import _lib
    _lib.load()
    if args.only:
        r = {"forward": lambda: forward_bench(args.steps, args.warmup, args.tune, args.tune_out),
             "batch": lambda: batch_bench(tune=args.tune, tune_out=args.tune_out), "diff": diff_bench,
             "war-iter": lambda: diff_bench(iters=1, fd_texels=0, bounces=(1,), plain_loop=False),
             "geom": lambda: geom_bench(args.geom_iters)}[args.only]()
        log(f"{args.only}: " + json.dumps({k: v for k, v in r.items() if isinstance(v, (int, float, str))}))
        return

    fwd = forward_bench(args.steps, args.warmup, args.tune, args.tune_out)
    traffic = profiled_traffic()
    launch_s = fwd["trace_avg_launch_ms"] * 1e-3
    line = {
        "metric": METRIC, "value": round(fwd["mrays"], 3), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(fwd["ms_per_step"], 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded sphere field)",
        "config": {"workload": "forward path tracing, 1024 icospheres (1,310,720 tris), 4096 emissive tris via "
                               "LightBVH NEE, 1920x1080, 4 spp, max 6 bounces",
                   "rays_closest_per_step": int(fwd["rays_closest"]),
                   "rays_visibility_per_step": int(fwd["rays_visibility"]),
                   "traversal_launch": fwd["launch"],
                   "visibility_nodes_per_ray": round(fwd["shadow_nodes_per_ray"], 2),
                   "visibility_tris_per_ray": round(fwd["shadow_tris_per_ray"], 2),
                   "tail_rays_per_step": [int(fwd["tail_rays_closest"]), int(fwd["tail_rays_visibility"])],
                   "tail_ms_per_step": round(fwd["tail_ms_per_step"], 4),
                   "bvh_build_s": round(fwd["build_s"], 2)},
        "roofline": {"bound": "hbm", "achieved": round(fwd["achieved_gbs"], 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(fwd["achieved_gbs"] / HBM_PEAK_GBS, 5), "traffic": traffic,
                     "traffic_source": str(TRAFFIC_PROFILE.relative_to(ROOT)),
                     "limiter": "L2 / Infinity Cache latency of the BVH pointer chase: the working set is cache "
                                "resident, so measured HBM traffic is a fraction of the algorithmic bytes",
                     "measured_hbm_gbs": round(traffic / launch_s / 1e9, 2) if traffic and launch_s > 0 else None,
                     "measured_hbm_frac": round(traffic / launch_s / 1e9 / HBM_PEAK_GBS, 5)
                     if traffic and launch_s > 0 else None,
                     "l2_hit_rate": profiled_l2_hit_rate(),
                     "kernel": "kTraceBounce (BVH4 traversal of the primary, closest-hit and visibility rays)",
                     "bytes_per_launch": round(fwd["bytes_per_launch"]),
                     "bytes_per_ray": round(fwd["bytes_per_ray"], 1), "nodes_per_ray": round(fwd["nodes_per_ray"], 2),
                     "bytes_per_visibility_ray": round(fwd["bytes_per_vis"], 1),
                     "tris_per_ray": round(fwd["tris_per_ray"], 2),
                     "avg_launch_ms": round(fwd["trace_avg_launch_ms"], 4),
                     "launches_per_step": fwd["trace_launches_per_step"]},
    }
    if not args.skip_batch:
        line["batch"] = {k: (round(v, 3) if isinstance(v, float) else v)
                         for k, v in batch_bench(tune=args.tune, tune_out=args.tune_out).items()}
    if not args.skip_diff:
        d = diff_bench()
        line["diff"] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in d.items() if k != "loop"}
        line["diff"]["loop"] = {k: (round(v, 5) if isinstance(v, float) else v) for k, v in d["loop"].items()}
    if not args.skip_geom:
        g = geom_bench(args.geom_iters)
        line["geom"] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in g.items() if k != "fd"}
        line["geom"]["fd"] = {k: {kk: round(vv, 6) for kk, vv in v.items()} for k, v in g["fd"].items()}
    if not args.skip_cpu:
        line["cpu_baseline"] = cpu_baseline(fwd["data"], fwd["W"], fwd["H"], fwd["spp"], fwd["B"])
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
