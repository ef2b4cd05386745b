"""mgk_trace_rays_noalloc against mgk_trace_rays, through the C-ABI.

The entry point takes its counters and (mixed precision) its zone hand-off
buffers from the library's fixed per-device array instead of allocating.  Two
calls in a row, the second with fewer rays, must give what mgk_trace_rays gives
on the same input: bit for bit for the FP32-reference and FP64 precisions
(every decision there is per lane), and for the mixed precision the same class
on >= 99.9 % of rays with g and the hit coordinate within FP32 rounding on
>= 99.8 % (its renormalisation runs on a wave-uniform schedule, so a ray's last
bits may depend on its wave-mates).  A mixed-precision call beyond the array's
16384 rays is refused before anything is launched.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import magik_amd as M
from magik_amd import abi

pytestmark = pytest.mark.gpu

SEED = 77


def noalloc_trace(ctx, integ, sc, cam_pos, rays, n):
    stride = rays.shape[1]
    hits = torch.zeros((stride, 4), dtype=torch.int32, device=rays.device)
    steps = torch.zeros(stride, dtype=torch.int32, device=rays.device)
    cp = (C.c_float * 3)(*cam_pos)
    r = ctx.lib.mgk_trace_rays_noalloc(ctx.h, C.byref(integ), C.byref(sc), cp, C.c_void_p(rays.data_ptr()), stride, n, None,
                                       C.c_void_p(hits.data_ptr()), C.c_void_p(steps.data_ptr()), ctx._stream())
    return r, hits, steps


@pytest.mark.parametrize("prec", [M.PREC_MIXED, M.PREC_FP64, M.PREC_FP32_REFERENCE])
def test_noalloc_trace_matches_trace_two_calls(gpu_ctx, prec):
    # 128 x 64 x 2 = 16384 rays (the mixed-precision maximum), then 96 x 54 x 2 = 10368
    for W, H in ((128, 64), (96, 54)):
        integ, sc, spec, cam, _ = M.workload("kerr_disk_1080p", W=W, H=H, spp=2, precision=prec)
        gpu_ctx.set_scene(sc)
        rays = gpu_ctx.generate_rays(integ, cam, SEED, 0, 2)
        n = rays.shape[1]
        print(f"noalloc: precision {prec}, {W}x{H}x2 = {n} rays", flush=True)
        ref_hits, ref_steps = gpu_ctx.trace(integ, sc, list(cam.position), rays)
        r, hits, steps = noalloc_trace(gpu_ctx, integ, sc, list(cam.position), rays, n)
        torch.cuda.synchronize()
        assert r == abi.MAGIK_SUCCESS, r
        a, b = M.hits_to_numpy(hits), M.hits_to_numpy(ref_hits)
        assert np.all(a["cls"] != M.HIT_NONE)
        if prec == M.PREC_MIXED:
            same = a["cls"] == b["cls"]
            assert same.mean() >= 0.999, same.mean()
            m = same & (a["cls"] != M.HIT_HORIZON)
            close = np.ones(n, dtype=bool)
            for f in ("g", "a0"):
                d = np.abs(a[f].astype(np.float64) - b[f]) / np.maximum(np.abs(b[f].astype(np.float64)), 1e-6)
                close[m] &= d[m] <= 1e-5
            assert close.mean() >= 0.998, close.mean()
        else:
            assert torch.equal(hits, ref_hits) and torch.equal(steps, ref_steps)


def test_noalloc_trace_refuses_mixed_beyond_its_array(gpu_ctx):
    integ, sc, spec, cam, _ = M.workload("kerr_disk_1080p", W=128, H=64, spp=3, precision=M.PREC_MIXED)
    rays = gpu_ctx.generate_rays(integ, cam, SEED, 0, 3)
    torch.cuda.synchronize()
    r, hits, _ = noalloc_trace(gpu_ctx, integ, sc, list(cam.position), rays, rays.shape[1])   # 24576 rays
    torch.cuda.synchronize()
    assert r == abi.MGK_ERROR_CAPACITY_EXCEEDED
    assert torch.all(hits == 0)
