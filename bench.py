"""Headline benchmark: Path integrator throughput (Mrays/s) at 1920x1080 on
one MI355X, on the procedural terrain workload (1.0M-triangle fBm heightfield
+ 2,000 instanced boxes, Disney materials, 64 point lights + 1 area light,
Russian roulette after depth 3, max depth 12).  With --integrator bdpt it
measures the BDPT caustics workload instead (20k-triangle glass sphere on a
diffuse plane under a 15-degree spot light, 1024x1024, subpaths of up to 8
bounces, 64 frames); with --integrator vcm the VCM workload on the same
scene (1,048,576 light subpaths per frame stored as photons, merge radius
0.5% of the scene diagonal shrinking with alpha = 0.75, vertex connection
and merging, 100 frames).

A step is one 1-spp frame of the whole image (one lt_render launch).  Prints
one JSON line with the metric, the roofline of the path kernel (algorithmic
BVH bytes per launch = nodes_visited * 128 B + tris_tested * 48 B, counted by
the kernels themselves in an untimed counting pass over the same frames,
divided by the kernel's HIP-event duration) and the CPU baseline (the oracle
restatement of the reference Path integrator on the host cores, on a band of
rows of the same frame).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def traffic_from_profile(workload: str, integrator: str):
    """HBM bytes per launch from a committed rocprof counter summary
    (profiles/*_pmc.json) of the same workload and integrator, if any."""
    for p in sorted((ROOT / "profiles").glob("*_pmc.json")):
        try:
            d = json.loads(p.read_text())
        except Exception:
            continue
        if d.get("workload") == workload and d.get("integrator") == integrator:
            return d.get("traffic_bytes_per_launch")
    return None


def cpu_baseline_bdpt(scene, camera, W, H, depth, frame, target_s: float, power: bool = False):
    import oracle

    log("cpu baseline: building host BVH")
    acc = oracle.Accel(scene.desc)
    threads = oracle.hardware_threads()
    y0 = H // 2 - 16
    _, secs, _, _ = acc.render_bdpt(camera, W, H, depth, frame_start=frame, frames=1, window=(0, y0, W, y0 + 32),
                                    threads=threads, power=power)
    per_row = max(secs / 32, 1e-5)
    if per_row * H >= target_s:
        rows, frames = int(max(8, target_s / per_row)), 1
    else:
        rows, frames = H, int(max(1, min(64, target_s / (per_row * H))))
    y0 = max(0, H // 2 - rows // 2)
    log(f"cpu baseline: {rows} rows x {W} px x {frames} frame(s), {threads} threads")
    _, secs, cl, sh = acc.render_bdpt(camera, W, H, depth, frame_start=frame, frames=frames,
                                      window=(0, y0, W, y0 + rows), threads=threads, power=power)
    return {
        "value": round((cl + sh) / secs / 1e6, 4),
        "unit": "Mrays/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"camera rows {y0}..{y0 + rows - 1} (with their light subpaths) of the {W}x{H} frames "
                   f"{frame}..{frame + frames - 1} ({cl + sh} rays, {secs:.2f} s)"),
    }


def cpu_baseline_vcm(scene, camera, W, H, depth, frame, target_s: float, vcm_kw: dict):
    import oracle

    log("cpu baseline: building host BVH")
    acc = oracle.Accel(scene.desc)
    threads = oracle.hardware_threads()
    frames, total_s, cl_all, sh_all = 0, 0.0, 0, 0
    while total_s < target_s and frames < 16:  # whole frames: the eye pass needs the full light pass
        _, secs, cl, sh = acc.render_vcm(camera, W, H, depth, frame_start=frame + frames, frames=1, threads=threads,
                                         **vcm_kw)
        frames += 1
        total_s += secs
        cl_all += cl
        sh_all += sh
    log(f"cpu baseline: {frames} full frame(s), {threads} threads")
    return {
        "value": round((cl_all + sh_all) / total_s / 1e6, 4),
        "unit": "Mrays/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"{frames} full {W}x{H} frame(s) {frame}..{frame + frames - 1} (light and eye passes, "
                   f"{cl_all + sh_all} rays, {total_s:.2f} s)"),
    }


def cpu_baseline_sppm(scene, camera, W, H, depth, target_s: float, base_radius: float):
    import oracle

    log("cpu baseline: building host BVH")
    acc = oracle.Accel(scene.desc)
    threads = oracle.hardware_threads()
    frames, total_s, rays, state, img = 0, 0.0, 0, None, None
    while total_s < target_s and frames < 16:  # whole frames from frame 0: the per-pixel state carries over
        img, secs, r, _, state = acc.render_sppm(camera, W, H, depth, frame_start=frames, frames=1, seed=7,
                                                 base_radius=base_radius, state=state, accum=img, threads=threads)
        frames += 1
        total_s += secs
        rays += r
    log(f"cpu baseline: {frames} full frame(s), {threads} threads")
    return {
        "value": round(rays / total_s / 1e6, 4),
        "unit": "Mrays/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{frames} full {W}x{H} frame(s) 0..{frames - 1} (eye, light and gather passes, {rays} rays, "
                  f"{total_s:.2f} s)",
    }


def cpu_baseline(scene, camera, W, H, depth, sky, frame, target_s: float):
    import numpy as np

    import oracle

    log("cpu baseline: building host BVH")
    acc = oracle.Accel(scene.desc)
    threads = oracle.hardware_threads()
    # calibrate on a band of 32 rows, then size the sample (rows x frames) for
    # about target_s seconds of CPU work
    y0 = H // 2 - 16
    _, secs, _, _ = acc.render_path(camera, W, H, depth, sky_col=sky, frame_start=frame, frames=1,
                                    window=(0, y0, W, y0 + 32), threads=threads)
    per_row = max(secs / 32, 1e-5)
    if per_row * H >= target_s:
        rows, frames = int(max(8, target_s / per_row)), 1
    else:
        rows, frames = H, int(max(1, min(64, target_s / (per_row * H))))
    y0 = max(0, H // 2 - rows // 2)
    log(f"cpu baseline: {rows} rows x {W} px x {frames} frame(s), {threads} threads")
    _, secs, cl, sh = acc.render_path(camera, W, H, depth, sky_col=sky, frame_start=frame, frames=frames,
                                      window=(0, y0, W, y0 + rows), threads=threads)
    return {
        "value": round((cl + sh) / secs / 1e6, 4),
        "unit": "Mrays/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"rows {y0}..{y0 + rows - 1} of the 1920x1080 frames {frame}..{frame + frames - 1}, 1 spp each "
                   f"({cl + sh} rays, {secs:.2f} s)"),
    }


def time_to_rmse(it, ctx, ms_per_frame: float, gt_frames: int, gt_seed: int = 0x6A09E667, tol: float = 1e-3,
                 budget_s: float = 20.0):
    """Time until the running image reaches an RMSE of `tol`, under both
    definitions (rmse/output_rmse.comp:23 and the standard one).

    Ground truth: `gt_frames` frames of the same integrator with another RNG
    seed.  Then the running image restarts from frame 0 and its squared
    error S against the ground truth is read after 1, 2, ..., 16, 24, 32, ...
    up to 1024 frames, until both bounds are met or `budget_s` seconds pass.
    Lumen's RMSE is sqrt(S) / (3 N) (N pixels); the standard RMSE is
    sqrt(S / (3 N)), taken relative to the ground truth's mean.  The first
    frame count meeting each bound times the measured ms per frame is the
    time.  A bound not met within the frames checked reports frames = null
    with the last measured value; nothing is extrapolated.  Beside it, the
    tail of the measurements (the checks at >= 1/4 of the last frame count)
    is fitted to RMSE^2(k) = sigma^2 (1/k + 1/gt_frames), the ground truth's
    own variance included: the fit's prediction at the last check, the noise
    floor sqrt(sigma^2 / gt_frames) that a ground truth of gt_frames frames
    leaves, and the frames a noise-free ground truth would need
    (sigma^2 / tol^2), given only when that floor lies below the bound."""
    import numpy as np

    seed0, frame0 = it.seed, it.frame_num
    it.seed = gt_seed
    it.reset()
    left = gt_frames
    t0 = time.perf_counter()
    while left > 0:
        n = min(left, 64)
        it.render(spp=n)
        it.update()
        left -= n
    gt = it.output_tex
    gt_s = time.perf_counter() - t0
    mean = float(gt[..., :3].mean())
    npx = gt.shape[0] * gt.shape[1]
    it.seed = seed0
    it.reset()
    checks = sorted(set(list(range(1, 17)) + [24, 32, 48, 64, 96, 128, 192, 256, 384, 512, 768, 1024]))
    pts = []
    lumen_k = None
    done = 0
    t_check = time.perf_counter()
    for k in checks:
        it.render(spp=k - done)
        it.update()
        done = k
        S = ctx.sq_error(it.output_tex, gt)
        lum = np.sqrt(S) / (3 * npx)
        std = np.sqrt(S / (3 * npx))
        pts.append((k, lum, std))
        if lumen_k is None and lum <= tol:
            lumen_k = k
        # bounded: stop once both bounds are met or after budget_s seconds
        if (lumen_k is not None and std <= tol * mean) or time.perf_counter() - t_check > budget_s:
            break
    ks = np.array([p[0] for p in pts], np.float64)
    rel2 = (np.array([p[2] for p in pts], np.float64) / mean) ** 2
    std_k = next((p[0] for p in pts if p[2] <= tol * mean), None)
    tail = ks >= ks[-1] / 4
    basis = 1.0 / ks[tail] + 1.0 / gt_frames
    sigma2 = float((basis * rel2[tail]).sum() / (basis * basis).sum())
    floor_rel = float(np.sqrt(sigma2 / gt_frames))
    fit = {"points": int(tail.sum()), "sigma2_rel_per_frame": sigma2,
           "predicted_at_max_checked": float(np.sqrt(sigma2 * (1.0 / ks[-1] + 1.0 / gt_frames))),
           "gt_noise_floor_rel": floor_rel,
           "frames_with_noise_free_gt": round(sigma2 / tol ** 2, 1) if floor_rel < tol else None}
    it.seed, it.frame_num = seed0, frame0
    return {
        "tol": tol,
        "ground_truth": f"{gt_frames} frames, seed {gt_seed:#x} ({gt_s:.1f} s)",
        "ms_per_frame": round(ms_per_frame, 3),
        "lumen_rmse": {"definition": "sqrt(S)/(3N)", "frames": lumen_k,
                       "seconds": None if lumen_k is None else round(lumen_k * ms_per_frame / 1e3, 4),
                       "at_1_frame": float(pts[0][1]),
                       "at_max_checked": {"frames": int(ks[-1]), "rmse": float(pts[-1][1])}},
        "standard_rmse_rel": {"definition": "sqrt(S/(3N)) / mean(ground truth)", "frames": std_k,
                              "seconds": None if std_k is None else round(std_k * ms_per_frame / 1e3, 3),
                              "at_max_checked": {"frames": int(ks[-1]), "rmse_rel": float(pts[-1][2] / mean)},
                              "tail_fit": fit},
    }


def cpu_baseline_adjoint(scene, camera, W, H, depth, albedo, rough):
    """The oracle's differentiable estimator and adjoint on the host cores:
    one variant at 1 spp over the whole image, forward then backward (first
    pass + replay), scaled to paths per second."""
    import numpy as np

    import oracle

    log("cpu baseline: building host BVH")
    acc = oracle.Accel(scene.desc)
    threads = oracle.hardware_threads()
    kw = dict(albedo=albedo[None], rough=rough[None], albedo_id=0, rough_id=1, seed=1)
    t0 = time.perf_counter()
    img = acc.adjoint(camera, W, H, depth, 1, threads=threads, **kw)
    acc.adjoint(camera, W, H, depth, 1, threads=threads, d_images=np.ones_like(img), **kw)
    secs = time.perf_counter() - t0
    return {
        "value": round(W * H / secs / 1e6, 4),
        "unit": "Mpaths/s (forward + adjoint per path)",
        "cores": threads,
        "kind": "port",
        "sample": f"the whole {W}x{H} image, 1 variant, 1 spp, forward and backward ({secs:.2f} s)",
    }


def bench_adjoint(args) -> int:
    """C4: differentiable Cornell box (back wall = 16,384-triangle grid with
    256x256 albedo and roughness textures), 512x512, 32 spp, depth 6, 32
    texture variants per launch, L2 loss against a target render.  A step is
    one optimisation iteration: forward (writing the per-sample cache),
    loss gradient, backward."""
    import numpy as np
    import torch

    import lumen_amd as la
    from lumen_amd.adjoint import Adjoint

    W = args.width or 512
    H = args.height or 512
    V, spp, depth = 32, 32, 6
    scene = la.Scene.procedural("diff_wall", seed=5)
    d = scene.desc
    tex = [np.ctypeslib.as_array(d.textures[t].data, shape=(d.textures[t].height, d.textures[t].width, 4)).copy()
           for t in range(d.num_textures)]
    alb0, rough0 = tex[0][..., :3].astype(np.float32), tex[1][..., 0].astype(np.float32)
    ctx = la.Context(0)
    n_cu, arch = ctx.device_info()
    adj = Adjoint(scene, ctx, W, H, spp=spp, variants=V, albedo_texture=0, roughness_texture=1, max_depth=depth)
    rng = np.random.default_rng(1)
    albs = torch.from_numpy(np.clip(alb0[None] + rng.uniform(-0.2, 0.2, (V,) + alb0.shape), 0.05, 0.95)
                            .astype(np.float32)).cuda()
    roughs = torch.from_numpy(np.clip(rough0[None] + rng.uniform(-0.2, 0.2, (V,) + rough0.shape), 0.1, 0.95)
                              .astype(np.float32)).cuda()
    # target: the base textures, rendered once
    base_a = torch.from_numpy(np.repeat(alb0[None], V, 0)).cuda()
    base_r = torch.from_numpy(np.repeat(rough0[None], V, 0)).cuda()
    target, _ = adj.forward(base_a, base_r, seed=1000, cache=False)
    torch.cuda.synchronize()
    log(f"adjoint: {W}x{H}, {spp} spp, {V} variants, depth {depth}, cache {adj.cache_floats * 4 / 2**30:.2f} GiB")
    steps = args.steps if "--steps" in sys.argv else 5
    warmup = args.warmup if "--warmup" in sys.argv else 1

    def step(k):
        img, cache = adj.forward(albs, roughs, seed=k)
        dl = 2.0 * (img - target)
        ga, gr = adj.backward(albs, roughs, dl, seed=k, cache=cache)
        return img, ga, gr

    for k in range(warmup):
        step(k)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True),
           torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    t0 = time.perf_counter()
    for k in range(steps):
        e0, e1, e2 = ev[k]
        e0.record()
        img, cache = adj.forward(albs, roughs, seed=warmup + k)
        e1.record()
        ga, gr = adj.backward(albs, roughs, 2.0 * (img - target), seed=warmup + k, cache=cache)
        e2.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    fwd_ms = sum(a.elapsed_time(b) for a, b, _ in ev) / steps
    bwd_ms = sum(b.elapsed_time(c) for _, b, c in ev) / steps
    loss = float(((img - target) ** 2).sum())
    paths = V * W * H * spp
    ms = wall / steps * 1e3
    log(f"timed {steps} iterations in {wall:.3f}s: forward {fwd_ms:.1f} ms, backward {bwd_ms:.1f} ms, loss {loss:.4g}")
    # algorithmic bytes of one iteration from the counting variants of the
    # same kernels (untimed): BVH8 nodes x 128 B + triangles x 48 B, plus the
    # sample cache, written forward and read backward (12 B of radiance and
    # 4 B of flags per sample, a 16 B hit record per closest-hit ray of the
    # forward), and the images / their adjoints (12 B per pixel and variant);
    # texel reads and the gradient atomics are not counted
    adj.set_counting(True)
    img_c, cache_c = adj.forward(albs, roughs, seed=warmup + steps)
    cf = adj.counts()
    adj.backward(albs, roughs, 2.0 * (img_c - target), seed=warmup + steps, cache=cache_c)
    cb = adj.counts()
    adj.set_counting(False)
    cache_b = 12 * spp * V * W * H
    if depth <= 10:  # the forward records its paths
        cache_b += 4 * spp * V * W * H + 16 * cf["closest"]
    img_b = 12 * V * W * H
    fwd_bytes = cf["nodes"] * 128 + cf["tris"] * 48 + cache_b + img_b
    bwd_bytes = cb["nodes"] * 128 + cb["tris"] * 48 + cache_b + img_b

    def line(kernel, nbytes, kms, counts):
        ach = nbytes / (kms * 1e-3) / 1e9
        return {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None, "kernel": kernel, "kernel_ms_avg": round(kms, 3),
                "algorithmic_bytes_per_launch": int(nbytes), "rays_per_launch": counts["closest"] + counts["shadow"]}

    roof = line("k_adjoint<true> (path replay with texel-gradient atomics)", bwd_bytes, bwd_ms, cb)
    roof["other_kernels"] = {"k_adjoint<false>": line("k_adjoint<false> (forward, writes the sample cache)",
                                                      fwd_bytes, fwd_ms, cf)}
    roof["replay"] = ("the replay reads each path's hit records and traces only the occlusion rays the forward "
                      "skipped; its time goes to the texel-gradient atomics and the BSDF derivatives")
    roof["traffic"] = traffic_from_profile(f"diff-cornell-wall16k-tex256-{W}x{H}-spp{spp}-v{V}-d{depth}", "adjoint")
    roof["note"] = ("the scene's BVH8 (16k-triangle wall and box) stays in L2 and the Infinity Cache, so the "
                    "algorithmic bytes can exceed the HBM peak; traffic is what reaches HBM")
    log(f"adjoint roofline: {roof}")
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_adjoint(scene, adj.camera, W, H, depth, alb0, rough0)
    out = {
        "metric": "adjoint ms/iter (C4, 1 GPU)",
        "value": round(ms, 2),
        "unit": "ms/iter",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": round(ms, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scene and texture variants)",
        "config": {
            "workload": f"diff-cornell-wall16k-tex256-{W}x{H}-spp{spp}-v{V}-d{depth}",
            "variants": V, "spp": spp, "width": W, "height": H, "max_depth": depth,
            "paths_per_iter": paths,
            "forward_ms": round(fwd_ms, 2), "backward_ms": round(bwd_ms, 2),
            "mpaths_per_s_forward_plus_replay": round(2 * paths / (ms * 1e-3) / 1e6, 1),
        },
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    adj.close()
    ctx.close()
    return 0


def stream_copy(mib: int) -> int:
    """Copies mib MiB device-to-device 5 times with the plain copy kernel;
    reports read+write GB/s (the attainable HBM bandwidth)."""
    import lumen_amd as la

    ctx = la.Context(0)
    n = mib << 20
    src, dst = ctx.alloc(n), ctx.alloc(n)
    times = [ctx.stream_copy(dst, src, n) for _ in range(5)]
    best = min(times)
    print(json.dumps({"stream_copy_bytes": n, "ms": [round(t, 4) for t in times],
                      "read_plus_write_GBps": round(2 * n / (best * 1e-3) / 1e9, 1)}), flush=True)
    ctx.free(src)
    ctx.free(dst)
    ctx.close()
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=32)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--integrator", choices=["path", "bdpt", "vcm", "sppm", "adjoint"], default="path")
    ap.add_argument("--mis", choices=["balance", "power"], default="power",
                    help="BDPT MIS heuristic (the caustics workload specifies the power heuristic)")
    ap.add_argument("--width", type=int, default=0, help="default 1920 (path) / 1024 (bdpt, vcm)")
    ap.add_argument("--height", type=int, default=0, help="default 1080 (path) / 1024 (bdpt, vcm)")
    ap.add_argument("--estimator", choices=["corrected", "reference"], default="corrected",
                    help="VCM estimator (the reference's shader or the corrected sums)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--strategy", choices=["persistent", "wavefront", "megakernel"], default="persistent",
                    help="Path execution form: persistent wavefront stages, one persistent kernel with path "
                         "regeneration, or the per-pixel megakernel")
    ap.add_argument("--gt-frames", type=int, default=-1,
                    help="ground-truth frames of the time-to-RMSE measurement (0 skips it; default 1024 for "
                         "path and bdpt, 128 for vcm)")
    ap.add_argument("--stream-copy", type=int, default=0, metavar="MIB",
                    help="only run the HBM copy kernel on MIB MiB (bandwidth / counter calibration)")
    args = ap.parse_args()
    if args.stream_copy:
        return stream_copy(args.stream_copy)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (each GPU box here has one GPU); "
              "run with --gpus 1", file=sys.stderr)
        return 2

    if args.integrator == "adjoint":
        return bench_adjoint(args)

    import numpy as np

    import lumen_amd as la

    bdpt = args.integrator == "bdpt"
    vcm = args.integrator == "vcm"
    sppm = args.integrator == "sppm"
    caustics = bdpt or vcm or sppm
    W = args.width or (1024 if caustics else 1920)
    H = args.height or (1024 if caustics else 1080)
    kind = "caustics" if caustics else "terrain"
    log(f"generating {kind} scene")
    t0 = time.time()
    scene = la.Scene.procedural(kind, seed=7 if caustics else 1)
    log(f"scene: {scene.num_triangles} triangles in {time.time() - t0:.1f}s")
    ctx = la.Context(0)
    n_cu, arch = ctx.device_info()
    vcm_kw = dict(radius_factor=1.0, alpha=0.75, use_vm=True, corrected=args.estimator == "corrected")
    diag = float(np.linalg.norm(np.array(scene.desc.scene_max[:]) - np.array(scene.desc.scene_min[:])))
    sppm_r0 = 0.005 * diag  # 0.5% of the scene diagonal, the VCM workload's r_0
    if bdpt:
        it = la.BDPT(scene, ctx, W, H, path_length=8, mis=args.mis)
        it.seed = 7
    elif vcm:
        # r_0 = 0.5% of the diagonal = 1% of the bounding radius (half diagonal)
        it = la.VCM(scene, ctx, W, H, path_length=8, radius_factor=1.0, alpha=0.75, enable_vm=True,
                    estimator=args.estimator)
        it.seed = 7
    elif sppm:
        it = la.SPPM(scene, ctx, W, H, path_length=8, base_radius=sppm_r0)
        it.seed = 7
    else:
        it = la.Path(scene, ctx, W, H, strategy=args.strategy)
    it.init()
    bst = it.create_accel()
    log(f"bvh8: {bst.num_nodes} nodes, depth {bst.max_depth}, build {bst.build_ms:.1f} ms on {arch} ({n_cu} CUs)")
    # launch shapes from the autotuner's table for this device and workload, if it has them
    from lumen_amd import autotune
    tune_key = ("bdpt" if bdpt else f"vcm-{args.estimator}" if vcm else "sppm" if sppm else f"path-{args.strategy}")
    tuned = autotune.apply(ctx, tune_key, kind)
    log(f"tuning: {ctx.get_tuning()} ({'from the autotuner table' if tuned else 'defaults'})")
    depth = it.path_length
    sky = tuple(scene.integrator_cfg().sky_col)
    if bdpt and args.steps == 32 and "--steps" not in sys.argv:
        args.steps = 64  # the caustics workload is 64 spp
    if (vcm or sppm) and args.steps == 32 and "--steps" not in sys.argv:
        args.steps = 100  # the VCM workload is 100 iterations

    for _ in range(args.warmup):
        it.render(spp=1)
        it.update()
    ctx.synchronize()
    first = it.frame_num

    # timed region: exactly K steps, synchronised on both sides; HIP events on
    # the context stream time the whole frame and each of its phases
    it.set_timing(True)
    kernel_ms, phase_ms = [], []
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        it.render(spp=1)
        st = it.stats()
        kernel_ms.append(st.kernel_ms)
        phase_ms.append(list(st.phase_ms[:st.num_phases]))
        it.update()
    ctx.synchronize()
    wall = time.perf_counter() - t0
    it.set_timing(False)
    log(f"timed {args.steps} frames in {wall:.3f}s")

    # untimed counting pass over the same frame indices (ray, BVH and photon
    # work counts are a function of the frame index only; SPPM's per-pixel
    # state is replayed from frame 0 to reach the same frames)
    it.set_counting(True)
    c = dict(rays=0, closest=0, nodes=0, tris=0, snodes=0, stris=0, tested=0, gathered=0, gbytes=0)
    if sppm:
        it.frame_num = 0
        for _ in range(first):
            it.render(spp=1)
            it.update()
    for k in range(args.steps):
        it.frame_num = first + k
        it.render(spp=1)
        s = it.stats()
        c["rays"] += s.closest_rays + s.shadow_rays
        c["closest"] += s.closest_rays
        c["nodes"] += s.nodes_visited
        c["tris"] += s.tris_tested
        c["snodes"] += s.shadow_nodes_visited
        c["stris"] += s.shadow_tris_tested
        c["tested"] += s.photons_tested
        c["gathered"] += s.photons_gathered
        c["gbytes"] += s.gather_bytes
    it.set_counting(False)
    it.frame_num = first + args.steps

    K = args.steps
    rays = c["rays"]
    mrays = rays / wall / 1e6
    avg_ms = sum(kernel_ms) / len(kernel_ms)
    nph = min(len(p) for p in phase_ms)
    ph_avg = [sum(p[i] for p in phase_ms) / K for i in range(nph)]
    bvh_bytes = (c["nodes"] * 128 + c["tris"] * 48) / K
    occl_bytes = (c["snodes"] * 128 + c["stris"] * 48) / K

    def line(kernel, bytes_per_launch, ms, extra=None):
        ach = bytes_per_launch / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
        d = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
             "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None, "kernel": kernel, "kernel_ms_avg": round(ms, 4),
             "algorithmic_bytes_per_launch": int(bytes_per_launch)}
        if extra:
            d.update(extra)
        return d

    if bdpt:
        workload = f"caustics-sphere20k-{W}x{H}-bdpt-{args.mis}-d8"
        out_integrator = f"bdpt-{args.mis}"
        # the walk traces the closest-hit rays, the connect pass the occlusion rays
        roof = line("k_bdpt_walk", bvh_bytes - occl_bytes, ph_avg[0],
                    {"other_kernels": {"k_bdpt_connect": line("k_bdpt_connect", occl_bytes, ph_avg[1]),
                                       "k_bdpt_resolve_ms": round(ph_avg[-1], 4)}})
    elif vcm and args.estimator == "corrected":
        workload = f"caustics-sphere20k-{W}x{H}-vcm-corrected-d8"
        out_integrator = "vcm-corrected"
        roof = line("k_vcm_merge", c["gbytes"] / K, ph_avg[2],
                    {"photons_tested_per_launch": c["tested"] // K, "photons_merged_per_launch": c["gathered"] // K,
                     "other_kernels": {
                         "light_pass_and_grid_plus_eye_pass": line("k_vcm_light + grid + k_vcm_eye", bvh_bytes,
                                                                   ph_avg[0] + ph_avg[1]),
                         "phase_ms": [round(v, 4) for v in ph_avg]}})
    elif vcm:
        workload = f"caustics-sphere20k-{W}x{H}-vcm-reference-d8"
        out_integrator = "vcm-reference"
        roof = line("one frame's launch sequence (light pass, photon grid, eye pass)", bvh_bytes, avg_ms,
                    {"phase_ms": [round(v, 4) for v in ph_avg]})
    elif sppm:
        workload = f"caustics-sphere20k-{W}x{H}-sppm-d8"
        out_integrator = "sppm"
        roof = line("k_sppm_eye + k_sppm_light (the traversal kernels)", bvh_bytes, ph_avg[0] + ph_avg[2],
                    {"other_kernels": {"k_sppm_gather": line("k_sppm_gather", c["gbytes"] / K, ph_avg[3]),
                                       "phase_ms": [round(v, 4) for v in ph_avg]},
                     "photons_tested_per_launch": c["tested"] // K,
                     "photons_gathered_per_launch": c["gathered"] // K})
    else:
        workload = "terrain-1M+2000-boxes-1920x1080-path-d12"
        out_integrator = f"path-{args.strategy}"
        roof = line("k_path_mega" if args.strategy == "megakernel" else
                    "k_path_regen (one launch per frame)" if args.strategy == "persistent" else
                    "one frame's launch sequence (trace + shade per depth)", bvh_bytes, avg_ms)
    roof["traffic"] = traffic_from_profile(workload, out_integrator)

    gt_frames = args.gt_frames if args.gt_frames >= 0 else (128 if (vcm or sppm) else 1024)
    rmse = None
    if gt_frames > 0:
        log(f"time to RMSE: ground truth of {gt_frames} frames")
        rmse = time_to_rmse(it, ctx, wall / args.steps * 1e3, gt_frames)
        log(f"time to RMSE: {rmse}")

    cpu = None
    if not args.no_cpu:
        if bdpt:
            cpu = cpu_baseline_bdpt(scene, it.camera, W, H, depth, first, args.cpu_seconds, args.mis == "power")
        elif vcm:
            cpu = cpu_baseline_vcm(scene, it.camera, W, H, depth, first, args.cpu_seconds, vcm_kw)
        elif sppm:
            cpu = cpu_baseline_sppm(scene, it.camera, W, H, depth, args.cpu_seconds, sppm_r0)
        else:
            cpu = cpu_baseline(scene, it.camera, W, H, depth, sky, first, args.cpu_seconds)

    out = {
        "metric": ("Mrays/s (BDPT, 1 GPU)" if bdpt else "Mrays/s (VCM, 1 GPU)" if vcm else
                   "Mrays/s (SPPM, 1 GPU)" if sppm else "Mrays/s (Path, 1080p, 1 GPU)"),
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scene)",
        "config": {
            "workload": workload,
            "integrator": out_integrator,
            "width": W,
            "height": H,
            "spp_per_step": 1,
            "max_depth": depth,
            "triangles": scene.num_triangles,
            "rays_per_step": rays // args.steps,
            "closest_rays_per_step": c["closest"] // args.steps,
            "shadow_rays_per_step": (rays - c["closest"]) // args.steps,
            "tuning": tuned,
        },
        "roofline": roof,
        "cpu_baseline": cpu,
        "time_to_rmse": rmse,
        "bvh_build_ms": round(bst.build_ms, 2),
    }
    print(json.dumps(out), flush=True)
    it.destroy()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
