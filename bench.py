"""Benchmark: Kerr a=0.9 1920x1080x32 spp frames on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload kerr_disk_1080p|kerr_disk_4k|batch64|box_led|box_mesh]

A step is one full frame through the C-ABI of synthetic seeded rays, timed on
two paths:

  plain      mgk_render_frame: ray generation + persistent RKF45 integration
             (FP32 bulk, FP64 near zone and controller) + spectral shading
  surrogate  mgk_render_frame_surrogate: ray generation + fused bf16 MLP
             inference + decode with the fallback rules + RKF45 for the
             fallback queue and the judge sample + the same shading; the
             network is trained first on the GPU (untimed setup, reported)

Each path's per-ray radiance error is measured against the FP64 integrator
on a seeded 2^16-ray subset of the frame with identical per-ray samples.
`value` is the rays/s of the fastest path whose error is within 1e-3; the
JSON line carries both paths, the roofline of the integration kernel
(algorithmic FP32 FLOPs per frame / its HIP-event time) and of the inference
kernel (MLP FLOPs / its HIP-event time), and the CPU baseline (the oracle's
FP32 restatement of the reference integrator on this host's cores).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Kerr rays/s (1080p, a=0.9): MLP surrogate vs plain RKF45 at ≤1e-3 radiance err"
PEAK_FP32_TFLOPS = 157.3          # MI355X FP32 vector peak (vendor figure, counts packed FP32)
PEAK_FP32_UNPACKED_TFLOPS = 78.6  # the same without packed FP32 (this build compiles packed FP32 off)
PEAK_BF16_TFLOPS = 2500.0         # MI355X dense bf16 MFMA peak
PEAK_HBM_TBS = 8.0                # MI355X HBM3E peak
ERROR_BAR = 1e-3
# Algorithmic operations of one reference RKF45 trial (6 Kerr right-hand sides,
# stage sums, error norm, controller), counted by oracle/opcount.cpp:
# 453 add/sub + 921 mul + 40 div + 13 sqrt + 1 pow.
FLOP_PER_TRIAL = {2: 1428, 1: 1272, 0: 0}
# MLP forward, 5 -> 4 x 128 ReLU -> 6 outputs: 2 (5*128 + 3*128*128 + 128*6)
FLOP_PER_INFERENCE = 101120
# HBM bytes moved by the integration of one Kerr 1080p x 32 frame (all zone
# passes): rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, the factors being what
# the same counters read for a copy kernel of known size (bench.py --copy-probe).
TRACE_TRAFFIC_BYTES = {"kerr_disk_1080p": (2975543 + 1247368) * 2.0 * 1024 + (4245581 + 542684) * 1024.0}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def subset(n, seed, n_sub=1 << 16):
    rng = np.random.default_rng(seed ^ 0xE44)
    return np.sort(rng.choice(n, size=min(n_sub, n), replace=False)).astype(np.int64)


def radiance_error(ctx, M, integ, sc, spec, cam, spp, seed, test_hits=None, n_sub=1 << 16, chord_scene=None):
    """Mean relative per-ray Y error against the FP64 path on a seeded subset of
    the frame's rays, with identical per-ray samples.  test_hits: the frame's
    hit records of the path under test ((n, 4) device tensor); when None the
    subset is traced with `integ`, and the endpoint-direction error and the
    drift of the constants of motion (H, L, Carter's Q) are measured too."""
    import torch
    from oracle import invariants as INV
    it64 = M.integrator()
    C.memmove(C.byref(it64), C.byref(integ), C.sizeof(integ))
    it64.precision = M.PREC_FP64
    rays = ctx.generate_rays(it64, cam, seed, 0, spp)            # float64 initial conditions
    # the FP32 reference path starts from its own float initial conditions
    rays_test = rays if integ.precision != M.PREC_FP32_REFERENCE else ctx.generate_rays(integ, cam, seed, 0, spp)
    n = rays.shape[1]
    q = subset(n, seed, n_sub)
    qt = torch.from_numpy(q.astype(np.int32)).to(rays.device)
    out, fin = {}, {}
    runs = [("fp64", it64, rays)] + ([("test", integ, rays_test)] if test_hits is None else [])
    for name, it, rr in runs:
        scn = chord_scene if (name == "test" and chord_scene is not None) else sc
        hits, final = ctx.trace_final(it, scn, list(cam.position), rr, n=len(q), queue=qt)
        out[name] = hits[qt.long()].contiguous()
        fin[name] = final[:, qt.long()].cpu().numpy()
        del hits, final
    del rays_test
    if test_hits is not None:
        out["test"] = test_hits[torch.from_numpy(q).to(test_hits.device)].contiguous()
    ic = rays[:, qt.long()].cpu().numpy()
    side = int(np.sqrt(len(q)))
    assert side * side == len(q)
    vcam = M.abi.Camera()
    C.memmove(C.byref(vcam), C.byref(cam), C.sizeof(cam))
    vcam.x_resolution = vcam.y_resolution = side
    y = {k: ctx.shade(sc, spec, vcam, seed, 0, 1, v)[..., 1].cpu().numpy().ravel().astype(np.float64)
         for k, v in out.items()}
    ht = M.hits_to_numpy(out["test"])
    hr = M.hits_to_numpy(out["fp64"])
    del rays
    torch.cuda.empty_cache()
    eps = 1e-3 * max(y["fp64"].mean(), 1e-30)
    err = np.abs(y["test"] - y["fp64"]) / np.maximum(y["fp64"], eps)
    res = {"mean_rel_Y": float(err.mean()), "p99_rel_Y": float(np.quantile(err, 0.99)),
           "frac_within_1e-3": float(np.mean(err <= 1e-3)),
           "class_agreement": float(np.mean(ht["cls"] == hr["cls"])), "rays": len(q)}
    # endpoint direction of the rays that escape on both paths
    esc = (ht["cls"] == M.HIT_ESCAPE) & (hr["cls"] == M.HIT_ESCAPE)
    if esc.any():
        def unit(h):
            c = h["a0"].astype(np.float64)
            s_ = np.sqrt(np.maximum(0.0, 1.0 - c * c))
            return np.stack([s_ * np.cos(h["a1"]), c, s_ * np.sin(h["a1"])])
        d = np.linalg.norm(unit(ht[esc]) - unit(hr[esc]), axis=0)
        ang = 2.0 * np.arcsin(np.minimum(1.0, 0.5 * d))
        res["endpoint_direction_err_rad"] = {"mean": float(ang.mean()), "p99": float(np.quantile(ang, 0.99)),
                                             "max": float(ang.max()), "rays": int(esc.sum())}
    # drift of the constants of motion between the initial conditions and termination
    if integ.metric_type == M.KERR:
        a_eq = float(np.float32(-integ.spin if integ.time_reversed else integ.spin))
        drift = {}
        for k, f in fin.items():
            h, dl, dq = INV.drift(ic[0:3], ic[4:8], f[0:3], f[4:8], a_eq)
            ok = np.isfinite(h) & np.isfinite(dq)
            drift[k] = {"H_over_p2": {"mean_abs": float(np.abs(h[ok]).mean()), "max_abs": float(np.abs(h[ok]).max())},
                        "L": {"mean_abs": float(np.abs(dl[ok]).mean()), "max_abs": float(np.abs(dl[ok]).max())},
                        "Q": {"mean_abs": float(np.abs(dq[ok]).mean()), "max_abs": float(np.abs(dq[ok]).max())}}
        res["constraint_drift"] = {"test" if k == "test" else "fp64": v for k, v in drift.items()}
    return res


def frame_error(ctx, M, integ, sc, spec, cam, spp, seed, frame):
    """Box scenes: paths that reflect around the hole are chaotic at the
    reference's tolerance, so the frame is compared with the FP64 frame of the
    same seed in aggregate: the relative difference of the mean Y, and of the
    mean Y over 32 x 32-pixel blocks (mean absolute, relative to the frame's
    mean)."""
    import torch
    it64 = M.integrator()
    C.memmove(C.byref(it64), C.byref(integ), C.sizeof(integ))
    it64.precision = M.PREC_FP64
    ref = ctx.render_frame(it64, sc, spec, cam, seed, spp)
    torch.cuda.synchronize()
    y, yr = frame[..., 1].double(), ref[..., 1].double()
    H, W = y.shape
    b = 32
    hb, wb = H // b, W // b
    blk = y[:hb * b, :wb * b].reshape(hb, b, wb, b).mean(dim=(1, 3))
    blk_r = yr[:hb * b, :wb * b].reshape(hb, b, wb, b).mean(dim=(1, 3))
    m = float(yr.mean())
    return {"mean_Y_rel_diff": float(abs(y.mean() - yr.mean()) / max(m, 1e-30)),
            "block32_mean_abs_rel_diff": float((blk - blk_r).abs().mean() / max(m, 1e-30)),
            "note": "aggregate comparison with the FP64 frame of the same seed (whole paths are chaotic)"}


def cpu_baseline(M, integ, sc, cam, spp, seed, budget_s=12.0, single_budget_s=5.0):
    """Oracle FP32 restatement of the reference integrator on the host cores,
    and the same on one core (a smaller sample) as `single_core`."""
    from oracle import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    it = M.integrator()
    C.memmove(C.byref(it), C.byref(integ), C.sizeof(integ))
    it.precision = M.PREC_FP32_REFERENCE
    npx = cam.x_resolution * cam.y_resolution
    rng = np.random.default_rng(seed)

    boxes = bool(sc.flags & M.SCENE_BOXES)
    if sc.flags & M.SCENE_MESHES:
        O.set_meshes(*M.workload_meshes("box_mesh"))

    def run(n, threads=threads):
        idx = np.sort(rng.choice(npx * spp, size=n, replace=False))
        rays = O.generate_rays(O.F32, it, cam, seed, (idx % npx).astype(np.uint32), (idx // npx).astype(np.uint32))
        t = time.perf_counter()
        if boxes:
            # whole paths with their reflections; the sample's paths keep their own Philox streams
            # only approximately (ray i of the sample stands for a path of the frame)
            _, steps, _ = O.trace_paths(O.F32, it, sc, cam, seed, 0, rays, threads=threads)
            trials = steps
        else:
            _, steps, trials = O.trace_rays(O.F32, it, sc, list(cam.position), rays, threads=threads)
        return time.perf_counter() - t, int(steps.sum()), int(trials.sum())

    dt, _, _ = run(1 << 14)
    n = int(min(1 << 25, max(1 << 16, (1 << 14) * budget_s / max(dt, 1e-3))))
    dt, steps, trials = run(n)
    what = "whole paths (reflections included)" if boxes else "rays"
    res = {"value": n / dt, "unit": "rays/s", "cores": threads, "kind": "port",
           "sample": f"{n} seeded {what} of the same {cam.x_resolution}x{cam.y_resolution}x{spp} frame, FP32 "
                     f"reference integrator (oracle/oracle.hpp) with OpenMP over rays, {dt:.1f} s",
           "steps_per_s": steps / dt}
    if not boxes:
        res["trials_per_s"] = trials / dt
    dt1, _, _ = run(1 << 10, threads=1)
    n1 = int(min(1 << 22, max(1 << 11, (1 << 10) * single_budget_s / max(dt1, 1e-3))))
    dt1, steps1, _ = run(n1, threads=1)
    res["single_core"] = {"value": n1 / dt1, "unit": "rays/s", "cores": 1,
                          "sample": f"{n1} seeded {what} of the same frame on one thread, {dt1:.1f} s",
                          "steps_per_s": steps1 / dt1}
    return res


def time_frames(render, steps, warmup, sync_stats=None):
    import torch
    for i in range(warmup):
        render(i)
        if sync_stats:
            sync_stats()
        log(f"  warmup {i} done")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    per = []
    for i in range(steps):
        per.append(render(1000 + i))
    torch.cuda.synchronize()
    return time.perf_counter() - t0, per


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="kerr_disk_1080p")
    ap.add_argument("--chord", default=None, choices=["affine", "hermite"],
                    help="mesh chord of a workload with meshes (default: the workload's)")
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp32_reference", "fp64"])
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-error", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--train-rays", type=int, default=1 << 24)
    ap.add_argument("--epochs", type=int, default=6)
    ap.add_argument("--mlp-only", action="store_true",
                    help="only time the inference kernel over one 1080p x 32 frame of feature rows (counter runs)")
    ap.add_argument("--manager", action="store_true",
                    help="only time the render manager (command queue + worker) against direct one-sample renders")
    ap.add_argument("--surrogate-only", action="store_true",
                    help="only train the surrogate and time its frames (profiler runs); prints the surrogate results")
    ap.add_argument("--rules", default="analytic,default,trace_disk",
                    help="comma-separated fallback rule sets of the surrogate frames")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run a copy kernel over BYTES three times (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        n = args.copy_probe // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 1.0, out=dst)   # an elementwise kernel: reads and writes BYTES each
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": n * 4, "copy_probe_bytes_written": n * 4, "launches": 3}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    import magik_amd as M

    if args.manager:
        bench_manager(M, args)
        return

    if args.mlp_only:
        n = 1920 * 1080 * 32
        sur = M.Surrogate(0, seed=args.seed)
        xf = torch.zeros((n, 16), dtype=torch.float32, device="cuda:0")
        xf[:, :5].uniform_(-1.0, 1.0)
        sur.forward(xf)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(3):
            sur.forward(xf)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / 3
        print(json.dumps({"kernel": "k_mlp_infer", "rows": n, "ms": ms,
                          "tflops": FLOP_PER_INFERENCE * n / (ms * 1e-3) / 1e12}), flush=True)
        return

    prec = {"mixed": M.PREC_MIXED, "fp32_reference": M.PREC_FP32_REFERENCE, "fp64": M.PREC_FP64}[args.precision]
    if args.workload == "batch64":
        bench_batch(M, args, prec)
        return
    integ, sc, spec, cam, spp = M.workload(args.workload, precision=prec)
    W, H = cam.x_resolution, cam.y_resolution
    n_paths = W * H * spp
    ctx = M.Context(0)
    ctx.set_scene(sc)
    meshes = M.workload_meshes(args.workload) if sc.flags & M.SCENE_MESHES else None
    if meshes:
        ctx.set_meshes(*meshes)
        if args.chord:
            sc.chord = M.CHORD_HERMITE if args.chord == "hermite" else M.CHORD_AFFINE
    out = torch.zeros((H, W, 3), dtype=torch.float32, device="cuda:0")
    log(f"bench: {args.workload} {W}x{H}x{spp} = {n_paths} paths, precision {args.precision}")
    if args.surrogate_only:
        sur_res, mlp_roof = bench_surrogate(M, ctx, args, integ, sc, spec, cam, spp, n_paths, out)
        ctx.close()
        print(json.dumps({"workload": args.workload, "paths_per_frame": n_paths, "surrogate": sur_res,
                          "roofline_mlp": mlp_roof}), flush=True)
        return

    # ---- plain RKF45 path ----
    acc = {"ms": 0.0, "steps": 0, "trials": 0, "fp64": 0}

    def plain_frame(seed_off):
        ctx.render_frame(integ, sc, spec, cam, args.seed + seed_off, spp, out)
        st = ctx.stats()                      # waits for this frame's integration
        return st

    log("plain RKF45 frames")
    wall, per = time_frames(plain_frame, args.steps, args.warmup)
    for st in per:
        acc["ms"] += st.ms
        acc["steps"] += st.n_steps
        acc["trials"] += st.n_trials
        acc["fp64"] += st.n_trials_fp64
    assert torch.isfinite(out).all()
    plain_rate = n_paths * args.steps / wall
    log(f"  {args.steps} frames in {wall:.3f} s: {plain_rate:.4e} rays/s")
    flop = FLOP_PER_TRIAL[integ.metric_type] * acc["trials"] / args.steps
    trace_ms = acc["ms"] / args.steps
    achieved = flop / (trace_ms * 1e-3) / 1e12
    plain = {"rays_per_s": plain_rate, "ms_per_frame": wall / args.steps * 1e3,
             "geodesic_steps_per_s": acc["steps"] / wall, "rkf45_trials_per_s": acc["trials"] / wall,
             "fp64_trial_fraction": acc["fp64"] / max(acc["trials"], 1), "trace_ms_per_frame": trace_ms}
    boxes = bool(sc.flags & M.SCENE_BOXES)
    if boxes:
        plain["reflections_per_path"] = sum(st.n_reflections for st in per) / (n_paths * args.steps)
    if not args.no_error and boxes:
        plain["frame_err_vs_fp64"] = frame_error(ctx, M, integ, sc, spec, cam, spp, args.seed, out)
        log(f"  frame error {plain['frame_err_vs_fp64']}")
    elif not args.no_error:
        plain["radiance_err_vs_fp64"] = radiance_error(ctx, M, integ, sc, spec, cam, spp, args.seed)
        log(f"  radiance error {plain['radiance_err_vs_fp64']}")
    if prec == M.PREC_MIXED and not boxes and (sc.flags & M.SCENE_DISK_ON_GEODESIC):
        # the same frames with the disk crossing on the step's chord (one trial
        # per disk path fewer; the chord's O(h^2) error instead)
        sc_chord = M.scene(**{f: getattr(sc, f) for f, _ in sc._fields_})
        sc_chord.flags &= ~M.SCENE_DISK_ON_GEODESIC
        log("plain RKF45 frames, disk crossing on the chord")

        def chord_frame(seed_off):
            ctx.render_frame(integ, sc_chord, spec, cam, args.seed + seed_off, spp, out)
            return ctx.stats()

        wallc, perc = time_frames(chord_frame, args.steps, 1)
        plain["chord_rays_per_s"] = n_paths * args.steps / wallc
        plain["chord_trace_ms_per_frame"] = sum(st.ms for st in perc) / args.steps
        log(f"  {plain['chord_rays_per_s']:.4e} rays/s")
        if not args.no_error:
            plain["chord_radiance_err_vs_fp64_geodesic"] = radiance_error(ctx, M, integ, sc, spec, cam, spp, args.seed,
                                                                          chord_scene=sc_chord)
            log(f"  radiance error against the FP64 frame with the crossing on the geodesic "
                f"{plain['chord_radiance_err_vs_fp64_geodesic']['mean_rel_Y']:.3g}")
    if prec == M.PREC_MIXED and not boxes:
        # the plain path in the reference's FP32 arithmetic (the integrator the
        # reference ships), for the speed ratios
        it32 = M.integrator()
        C.memmove(C.byref(it32), C.byref(integ), C.sizeof(integ))
        it32.precision = M.PREC_FP32_REFERENCE
        log("plain FP32 reference-arithmetic frames")

        def ref_frame(seed_off):
            ctx.render_frame(it32, sc, spec, cam, args.seed + seed_off, spp, out)
            return ctx.stats()

        wall32, _ = time_frames(ref_frame, max(1, min(args.steps, 2)), 1)
        plain["fp32_reference_rays_per_s"] = n_paths * max(1, min(args.steps, 2)) / wall32
        log(f"  {plain['fp32_reference_rays_per_s']:.4e} rays/s")
        if not args.no_error:
            plain["fp32_reference_radiance_err_vs_fp64"] = radiance_error(ctx, M, it32, sc, spec, cam, spp, args.seed)
            log(f"  radiance error {plain['fp32_reference_radiance_err_vs_fp64']}")

    # ---- display side of the splat over the last frame ----
    display_res = bench_display(M, ctx, out)

    # ---- surrogate path ----
    sur_res, mlp_roof = None, None
    if not args.no_surrogate and args.workload in ("kerr_disk_1080p", "kerr_disk_4k") and prec != M.PREC_FP32_REFERENCE:
        sur_res, mlp_roof = bench_surrogate(M, ctx, args, integ, sc, spec, cam, spp, n_paths, out)

    if sur_res:
        for v in sur_res["configs"].values():
            v["speedup_vs_plain"] = v["rays_per_s"] / plain["rays_per_s"]
            if "fp32_reference_rays_per_s" in plain:
                v["speedup_vs_plain_fp32_reference"] = v["rays_per_s"] / plain["fp32_reference_rays_per_s"]

    # ---- headline: the fastest path within the error bar ----
    cands = [("plain_rkf45", plain["rays_per_s"], plain.get("radiance_err_vs_fp64", {}).get("mean_rel_Y"))]
    if sur_res:
        for k, v in sur_res["configs"].items():
            cands.append((f"surrogate_{k}", v["rays_per_s"], v.get("radiance_err_vs_fp64", {}).get("mean_rel_Y")))
    ok = [c for c in cands if c[2] is not None and c[2] <= ERROR_BAR] or cands[:1]
    best = max(ok, key=lambda c: c[1])
    value = best[1]
    res = {
        "metric": METRIC,
        "value": value,
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": n_paths / value * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if prec != M.PREC_FP64 else "f64",
        "data": (f"synthetic (seeded Philox rays, Kerr a={integ.spin:.3f} hole inside a diffuse box with an LED "
                 "panel, the reference's default stage; seeded Lambertian reflections"
                 + ("; procedural triangle meshes)" if meshes else ")")) if boxes else
                (f"synthetic (seeded Philox rays, Kerr a={integ.spin:.2f} disk + 1e5-star field scene; surrogate "
                 "trained on seeded FP64-traced rays)"),
        "config": {"workload": (f"{args.workload}: Kerr a={integ.spin:.3f}, M=6.73e26 kg, 40x20x40 m diffuse box, "
                                f"20x20 m LED panel, camera (2,0,9) m, {W}x{H}x{spp} spp, 4 hero wavelengths, "
                                f"tol 1e-3/1e-10, step limit 512, {sc.max_reflections} reflections"
                                + (f", meshes: {sum(len(f) for _, f in meshes[0])} triangles in {len(meshes[0])} "
                                   f"meshes, {len(meshes[1])} instances, "
                                   f"{'Hermite' if sc.chord == M.CHORD_HERMITE else 'affine'} chords" if meshes else ""))
                               if boxes else
                               (f"{args.workload}: Kerr a={integ.spin:.2f}, r_obs=50M, theta=80 deg, FOV 25 deg, "
                                f"{W}x{H}x{spp} spp, 4 hero wavelengths, tol 1e-7, cap 1e4"),
                   "paths_per_frame": n_paths, "value_path": best[0], "error_bar": ERROR_BAR},
        "plain": plain,
        "surrogate": sur_res,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "frac_of_unpacked_peak": achieved / PEAK_FP32_UNPACKED_TFLOPS,
                     "traffic": TRACE_TRAFFIC_BYTES.get(args.workload) if prec == M.PREC_MIXED else None,
                     "traffic_unit": "bytes per frame (all zone passes)",
                     "kernel": "k_trace_zone (far + near passes of one frame)",
                     "note": "integration kernel of the plain path: FP32 vector ALU (no MFMA on this path); "
                             f"algorithmic FLOPs = {FLOP_PER_TRIAL[integ.metric_type]} per RKF45 trial x trials "
                             "per frame / HIP-event time of the frame's integration on the library stream"},
        "roofline_mlp": mlp_roof,
        "display": display_res,
    }
    if not args.no_cpu_baseline:
        res["cpu_baseline"] = cpu_baseline(M, integ, sc, cam, spp, args.seed)
        log(f"cpu baseline {res['cpu_baseline']['value']:.3e} rays/s on {res['cpu_baseline']['cores']} threads, "
            f"{res['cpu_baseline']['single_core']['value']:.3e} rays/s on one")
    ctx.close()
    print(json.dumps(res), flush=True)


def bench_manager(M, args, window_s=6.0):
    """The render manager at the stage's defaults (the box scene, 2195 x 1860,
    one sample per worker iteration, GT7 display into the "test" AOV): samples
    per second from its statistics over a time window, against the same
    one-sample frames rendered and display-transformed directly."""
    import torch
    m = M.Manager(0)
    try:
        m.push(M.Manager.aov_command("add_aov", "test", 3))
        while not m.dispatch():
            time.sleep(0.001)
        t0 = time.time()
        while m.stats()[0].x_sample < 3 and time.time() - t0 < 120:
            time.sleep(0.01)
        s0, t0 = m.stats()[0].x_sample, time.perf_counter()
        time.sleep(window_s)
        st = m.stats()[0]
        s1, t1 = st.x_sample, time.perf_counter()
        err = m.error()
    finally:
        m.close()
    mgr_rate = (s1 - s0) / (t1 - t0)
    # direct: the same frames through the library, one sample per call
    integ, sc, _, cam, _ = M.workload("box_led", precision=M.PREC_MIXED)
    spec = M.spectral(n_hero=4, hero_min=360.0, hero_max=830.0, n_bin=24, bin_min=360.0, bin_max=830.0)
    ctx = M.Context(0)
    ctx.set_scene(sc)
    disp = M.display()
    xyz = torch.zeros((cam.y_resolution, cam.x_resolution, 3), dtype=torch.float32, device="cuda:0")
    ctx.render_frame(integ, sc, spec, cam, 0, 1, xyz)
    torch.cuda.synchronize()
    k = 5
    t = time.perf_counter()
    for i in range(k):
        ctx.render_frame(integ, sc, spec, cam, i, 1, xyz)
        M.display_transform(ctx, disp, xyz, out=xyz)
    torch.cuda.synchronize()
    direct_rate = k / (time.perf_counter() - t)
    ctx.close()
    W, H = cam.x_resolution, cam.y_resolution
    res = {"metric": "render manager samples/s (box scene, 2195x1860, 1 sample per iteration)", "value": mgr_rate,
           "unit": "samples/s", "direct_samples_per_s": direct_rate, "manager_overhead_frac": 1 - mgr_rate / direct_rate,
           "manager_paths_per_s": mgr_rate * W * H, "latest_iteration_ms": st.latest_render_loop_iteration_time_seconds * 1e3,
           "latest_sample_ms": st.latest_wavefront_sample_time_seconds * 1e3, "samples": int(s1 - s0),
           "window_s": t1 - t0, "error": int(err)}
    log(f"manager {mgr_rate:.2f} samples/s vs direct {direct_rate:.2f} samples/s")
    print(json.dumps(res), flush=True)


def bench_display(M, ctx, frame, reps=20):
    """mgk_display_transform over one rendered frame (XYZ, H x W x 3 float32),
    per view transform: the mean launch time from HIP events on the stream the
    kernel runs on, and the algorithmic HBM rate (12 bytes read + 12 written
    per pixel) against the 8 TB/s peak.  The timed launches are queued behind a
    short spin kernel, so that they run back to back and the events see kernel
    time rather than the host's call overhead."""
    import torch
    n = frame.numel() // 3
    dst = torch.empty_like(frame)
    names = {M.VIEW_NONE: "none", M.VIEW_REINHARD: "reinhard_simple",
             M.VIEW_REINHARD_EXTENDED: "reinhard_extended", M.VIEW_GT7: "gt7_jzazbz",
             M.VIEW_AGX: "agx", M.VIEW_KHRONOS_PBR_NEUTRAL: "khronos_pbr_neutral"}
    res = {"pixels": n, "bytes_per_pixel": 24, "display": "Rec.709 (1/2.4 power)", "views": {}}
    for view, name in names.items():
        disp = M.display(view_transform=view)
        M.display_transform(ctx, disp, frame, out=dst)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda._sleep(5_000_000)   # ~2 ms at 2.4 GHz: time for the host to queue the launches
        e0.record()
        for _ in range(reps):
            M.display_transform(ctx, disp, frame, out=dst)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        gbs = 24.0 * n / (ms * 1e-3) / 1e9
        res["views"][name] = {"ms": ms, "achieved_GB_s": gbs, "frac_of_hbm_peak": gbs / (PEAK_HBM_TBS * 1e3)}
    log("display: " + ", ".join(f"{k} {v['ms'] * 1e3:.1f} us" for k, v in res["views"].items()))
    return res


def bench_batch(M, args, prec):
    """The batched uneven workload: 64 seeded cameras (spin, pose, resolution
    per camera) integrated in one persistent launch per frame."""
    import torch
    integ, sc, spec, cams, spp = M.batch_workload(seed=args.seed, precision=prec)
    ctx = M.Context(0)
    ctx.set_scene(sc)
    rays, off = ctx.generate_rays_batch(integ, cams, args.seed, spp)
    n = rays.shape[1]
    log(f"batch64: {len(cams)} cameras, {n} paths (resolutions {sorted(set(c.camera.x_resolution for c in cams))})")

    del rays
    torch.cuda.empty_cache()

    def frame(i):
        r_, _ = ctx.generate_rays_batch(integ, cams, args.seed + i, spp)
        hits, _ = ctx.trace_batch(integ, sc, r_, want_steps=False)
        st = ctx.stats()
        del r_, hits
        return st

    wall, per = time_frames(frame, args.steps, args.warmup)
    st = per[-1]
    res = {"metric": "batched uneven workload rays/s (64 cameras, one persistent launch per frame)",
           "value": n * args.steps / wall, "unit": "rays/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32" if prec != M.PREC_FP64 else "f64", "data": "synthetic (seeded cameras and rays)",
           "config": {"workload": f"batch64: {len(cams)} cameras, a~U[0,0.99], incl~U[10,89] deg, r~U[20,100] M, "
                                  f"res in {{128,256,512,1024}}^2, {spp} spp, {n} paths"},
           "trace_ms_per_frame": st.ms, "geodesic_steps_per_frame": st.n_steps, "rkf45_trials_per_frame": st.n_trials}
    ctx.close()
    print(json.dumps(res), flush=True)


def bench_surrogate(M, ctx, args, integ, sc, spec, cam, spp, n_paths, out):
    import torch
    log("surrogate: training data + training (untimed setup)")
    ranges = M.surrogate_ranges(spin=(0.0, 0.99), r=(20.0, 100.0), theta=(10.0, 89.0), screen=12.0,
                                disk_r_max=float(sc.disk_r_out))
    sur = M.Surrogate(0, seed=args.seed, ranges=ranges)
    fp64 = M.integrator(metric=M.KERR, precision=M.PREC_FP64, spin=integ.spin)
    tsc = M.scene(celestial_radius_m=sc.celestial_radius_m, flags=M.SCENE_DISK, disk_r_in=sc.disk_r_in,
                  disk_r_out=sc.disk_r_out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, x, t, _ = sur.generate_data(ctx, fp64, tsc, seed=args.seed + 17, n=args.train_rays)
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    t0 = time.perf_counter()
    hist = sur.train(x, t, batch=8192, epochs=args.epochs, lr=2e-3, lr_final=2e-5)
    t_train = time.perf_counter() - t0
    del x, t
    torch.cuda.empty_cache()
    log(f"  data {t_data:.2f} s, training {t_train:.2f} s, last epoch {hist[-1].tolist()}")
    res = {"training": {"rays": args.train_rays, "epochs": args.epochs, "data_s": t_data, "train_s": t_train,
                        "samples_per_s": args.train_rays * args.epochs / t_train,
                        "ranges": "a in [0, 0.99], r_obs in [20, 100] M, theta in [10, 89] deg, |alpha|,|beta| <= 12 M",
                        "last_epoch": {"ce": float(hist[-1, 0]), "disk_r": float(hist[-1, 1]),
                                       "dir": float(hist[-1, 2]), "accuracy": float(hist[-1, 3])}},
           "configs": {}}
    hits = torch.zeros((n_paths, 4), dtype=torch.int32, device="cuda:0")
    # analytic: the network's class checked, and a disk path's radius refined,
    # by the Mino-time integrals (the configuration meant for the error bar);
    # default: the network's answers as they are; trace_disk: every disk path traced
    rules = {"analytic": M.fallback(analytic=True),
             "default": M.fallback(),
             "trace_disk": M.fallback(trace_disk=True)}
    wanted = [r for r in args.rules.split(",") if r]
    if any(r not in rules for r in wanted):
        raise SystemExit(f"--rules: unknown rule set in {args.rules!r}; known: {', '.join(rules)}")
    rules = {k: rules[k] for k in wanted}
    for name, fb in rules.items():
        log(f"surrogate frames ({name})")
        stats = []

        def frame(seed_off):
            _, st = M.render_frame_surrogate(ctx, sur, integ, sc, spec, cam, args.seed + seed_off, spp, fb=fb,
                                             out=out, stats=True)
            stats.append(st)
            return st

        wall, per = time_frames(frame, args.steps, args.warmup)
        rate = n_paths * args.steps / wall
        st = per[-1]
        cfg = {"rays_per_s": rate, "ms_per_frame": wall / args.steps * 1e3, "speedup_vs_plain": None,
               "fallback_fraction": st["n_fallback"] / st["n_paths"],
               "stage_ms": {k: st[k] for k in ("ms_raygen", "ms_mlp", "ms_trace", "ms_shade", "ms_total")},
               "stats": st}
        log(f"  {rate:.4e} rays/s, fallback {cfg['fallback_fraction']:.4f}, stages {cfg['stage_ms']}")
        if not args.no_error:
            M.render_frame_surrogate(ctx, sur, integ, sc, spec, cam, args.seed, spp, fb=fb, out=out, hits=hits,
                                     stats=False)
            cfg["radiance_err_vs_fp64"] = radiance_error(ctx, M, integ, sc, spec, cam, spp, args.seed, test_hits=hits)
            log(f"  radiance error {cfg['radiance_err_vs_fp64']}")
        res["configs"][name] = cfg
    del hits
    torch.cuda.empty_cache()
    # inference kernel alone: the fused MLP over one frame's worth of feature rows
    xf = torch.zeros((n_paths, 16), dtype=torch.float32, device="cuda:0")
    xf[:, :5].uniform_(-1.0, 1.0)
    yf = sur.forward(xf)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 3
    e0.record()
    for _ in range(reps):
        sur.forward(xf)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    achieved = FLOP_PER_INFERENCE * n_paths / (ms * 1e-3) / 1e12
    mlp_roof = {"bound": "mfma", "achieved": achieved, "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / PEAK_BF16_TFLOPS, "traffic": None, "kernel": "k_mlp_infer",
                "ms_per_frame": ms,
                "note": f"{FLOP_PER_INFERENCE} FLOP per path (5 -> 4x128 ReLU -> 6) x {n_paths} paths / "
                        "HIP-event time of the kernel on the launch stream"}
    log(f"  k_mlp_infer {ms:.2f} ms per frame, {achieved:.1f} TFLOP/s")
    del xf, yf
    sur.close()
    torch.cuda.empty_cache()
    return res, mlp_roof


if __name__ == "__main__":
    main()
