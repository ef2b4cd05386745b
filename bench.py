"""Benchmark: Kerr a=0.9 1920x1080x32 spp frames on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload kerr_disk_1080p]

A step is one full frame through the C-ABI (mgk_render_frame: ray generation +
persistent RKF45 integration + spectral shading) of synthetic seeded rays.
Prints one JSON line: rays/s (completed paths per second of wall time over K
frames), geodesic steps/s, RKF45 trials/s, the roofline of the integration
kernel (algorithmic FP32 FLOPs per launch / its HIP-event time on the
library's stream), the radiance error of the frame's rays against the FP64
path on a seeded 2^16-ray subset, and the CPU baseline (the oracle's FP32
restatement of the reference integrator on this host's cores).
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
PEAK_FP32_TFLOPS = 157.3          # MI355X FP32 vector/matrix peak (vendor figure)
# Algorithmic operations of one reference RKF45 trial (6 Kerr right-hand sides,
# stage sums, error norm, controller), counted by oracle/opcount.cpp:
# 453 add/sub + 921 mul + 40 div + 13 sqrt + 1 pow.
FLOP_PER_TRIAL = {2: 1428, 1: 1272, 0: 0}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def radiance_error(ctx, M, integ, sc, spec, cam, spp, seed, n_sub=1 << 16):
    """Mean relative per-ray Y error of `integ` against the FP64 path on a
    seeded subset of the frame's rays, with identical per-ray samples."""
    import torch
    rays = ctx.generate_rays(integ, cam, seed, 0, spp)           # float64 storage (mixed / FP64)
    n = rays.shape[1]
    rng = np.random.default_rng(seed ^ 0xE44)
    q = np.sort(rng.choice(n, size=min(n_sub, n), replace=False)).astype(np.int32)
    qt = torch.from_numpy(q).to(rays.device)
    out = {}
    for name, prec in (("test", integ.precision), ("fp64", M.PREC_FP64)):
        it = M.integrator()
        C.memmove(C.byref(it), C.byref(integ), C.sizeof(integ))
        it.precision = prec
        hits, _ = ctx.trace(it, sc, list(cam.position), rays, n=len(q), queue=qt, want_steps=False)
        out[name] = hits[qt.long()].contiguous()
        del hits
    side = int(np.sqrt(len(q)))
    assert side * side == len(q)
    vcam = M.abi.Camera()
    C.memmove(C.byref(vcam), C.byref(cam), C.sizeof(cam))
    vcam.x_resolution = vcam.y_resolution = side
    y = {k: ctx.shade(sc, spec, vcam, seed, 0, 1, v)[..., 1].cpu().numpy().ravel().astype(np.float64)
         for k, v in out.items()}
    cls_t = out["test"][:, 0].cpu().numpy()
    cls_r = out["fp64"][:, 0].cpu().numpy()
    del rays
    torch.cuda.empty_cache()
    eps = 1e-3 * max(y["fp64"].mean(), 1e-30)
    err = np.abs(y["test"] - y["fp64"]) / np.maximum(y["fp64"], eps)
    return float(err.mean()), float(np.mean(cls_t == cls_r)), len(q)


def cpu_baseline(M, integ, sc, cam, spp, seed, budget_s=12.0):
    """Oracle FP32 restatement of the reference integrator on the host cores."""
    from oracle import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    it = M.integrator()
    C.memmove(C.byref(it), C.byref(integ), C.sizeof(integ))
    it.precision = M.PREC_FP32_REFERENCE
    npx = cam.x_resolution * cam.y_resolution
    rng = np.random.default_rng(seed)

    def run(n):
        idx = rng.choice(npx * spp, size=n, replace=False)
        rays = O.generate_rays(O.F32, it, cam, seed, (idx % npx).astype(np.uint32), (idx // npx).astype(np.uint32))
        t = time.perf_counter()
        _, steps, trials = O.trace_rays(O.F32, it, sc, list(cam.position), rays, threads=threads)
        return time.perf_counter() - t, int(steps.sum()), int(trials.sum())

    dt, _, _ = run(1 << 14)
    n = int(min(1 << 22, max(1 << 16, (1 << 14) * budget_s / max(dt, 1e-3))))
    dt, steps, trials = run(n)
    return {"value": n / dt, "unit": "rays/s", "cores": threads, "kind": "port",
            "sample": f"{n} seeded rays of the same 1920x1080x32 Kerr frame, FP32 reference integrator "
                      f"(oracle/oracle.hpp) with OpenMP over rays, {dt:.1f} s",
            "steps_per_s": steps / dt, "trials_per_s": trials / dt}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="kerr_disk_1080p")
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp32_reference", "fp64"])
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-error", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (every GPU box here has one GPU); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    import magik_amd as M

    prec = {"mixed": M.PREC_MIXED, "fp32_reference": M.PREC_FP32_REFERENCE, "fp64": M.PREC_FP64}[args.precision]
    integ, sc, spec, cam, spp = M.workload(args.workload, precision=prec)
    W, H = cam.x_resolution, cam.y_resolution
    n_paths = W * H * spp
    ctx = M.Context(0)
    ctx.set_scene(sc)
    out = torch.zeros((H, W, 3), dtype=torch.float32, device="cuda:0")
    log(f"bench: {args.workload} {W}x{H}x{spp} = {n_paths} paths, precision {args.precision}")
    for i in range(args.warmup):
        ctx.render_frame(integ, sc, spec, cam, args.seed, spp, out)
        ctx.stats()
        log(f"warmup {i} done")
    torch.cuda.synchronize()
    kern_ms, steps, trials, fp64 = 0.0, 0, 0, 0
    t0 = time.perf_counter()
    for i in range(args.steps):
        ctx.render_frame(integ, sc, spec, cam, args.seed + i, spp, out)
        st = ctx.stats()                      # waits for this frame's integration kernel
        kern_ms += st.ms
        steps += st.n_steps
        trials += st.n_trials
        fp64 += st.n_trials_fp64
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    log(f"timed {args.steps} frames in {wall:.3f} s")
    assert torch.isfinite(out).all()

    flop = FLOP_PER_TRIAL[integ.metric_type] * trials / args.steps
    achieved = flop / (kern_ms / args.steps * 1e-3) / 1e12
    res = {
        "metric": METRIC,
        "value": n_paths * args.steps / wall,
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if prec != M.PREC_FP64 else "f64",
        "data": "synthetic (seeded Philox rays, Kerr a=0.9 disk + 1e5-star field scene)",
        "config": {"workload": f"{args.workload}: Kerr a=0.9, r_obs=50M, theta=80 deg, FOV 25 deg, "
                               f"{W}x{H}x{spp} spp, 4 hero wavelengths, tol 1e-7, cap 1e4",
                   "path": f"plain RKF45 ({args.precision})", "paths_per_frame": n_paths},
        "geodesic_steps_per_s": steps / wall,
        "rkf45_trials_per_s": trials / wall,
        "fp64_trial_fraction": fp64 / max(trials, 1),
        "trace_kernel_ms": kern_ms / args.steps,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None,
                     "note": "integration kernel: FP32 vector ALU (no MFMA on this path); algorithmic FLOPs = "
                             f"{FLOP_PER_TRIAL[integ.metric_type]} per RKF45 trial x trials per launch"},
    }
    if not args.no_error:
        e, agree, n = radiance_error(ctx, M, integ, sc, spec, cam, spp, args.seed)
        res["radiance_err_vs_fp64"] = {"mean_rel_Y": e, "class_agreement": agree, "rays": n}
        log(f"radiance error {e:.3e}, class agreement {agree:.5f}")
    if not args.no_cpu_baseline:
        res["cpu_baseline"] = cpu_baseline(M, integ, sc, cam, spp, args.seed)
        log(f"cpu baseline {res['cpu_baseline']['value']:.3e} rays/s")
    ctx.close()
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
