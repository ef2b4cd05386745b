"""GPU parity of reflective matter and the bounce path (mgk_trace_paths,
mgk_render_frame on a box scene) with the CPU oracle, through the C-ABI.

Paths that reflect around a hole are chaotic at the reference's tolerance
(1e-3): the oracle's own FP32 and FP64 paths agree on the outcome for 99.9 %
of unreflected paths but on only 74-86 % after 9-10 reflections, and FP64 at
two tolerances agrees on 87 % of whole paths.  So:
- per path, the first segment and the first reflection are compared;
- whole paths are compared by their outcome fractions, within binomial noise;
- two invariants that do not depend on the path hold per path at full length:
  g = 1 in flat space, and g = sqrt(g_tt(emitter) / g_tt(camera)) in
  Schwarzschild."""
import math

import numpy as np
import pytest

import magik_amd as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SEED = 17


def setup(metric, precision, W=48, H=40, spp=2, max_reflections=10, spin=0.999, tol=None):
    integ, sc, spec, cam, _ = M.workload("box_led", W=W, H=H, spp=spp, precision=precision)
    integ.metric_type = metric
    integ.spin = spin if metric == M.KERR else 0.0
    if tol is not None:
        integ.abs_tol = integ.rel_tol = tol
    sc.max_reflections = max_reflections
    return integ, sc, spec, cam, spp


def oracle_prec(precision):
    return O.F32 if precision == M.PREC_FP32_REFERENCE else O.F64


def oracle_paths(integ, sc, cam, spp, precision, want_at=False):
    n = cam.x_resolution * cam.y_resolution
    pix = np.tile(np.arange(n, dtype=np.uint32), spp)
    smp = np.repeat(np.arange(spp, dtype=np.uint32), n)
    rays = O.generate_rays(oracle_prec(precision), integ, cam, SEED, pix, smp)
    return rays, O.trace_paths(oracle_prec(precision), integ, sc, cam, SEED, 0, rays, want_at=want_at)


def gpu_paths(ctx, integ, sc, cam, spp, want_final=False):
    ctx.set_scene(sc)
    rays = ctx.generate_rays(integ, cam, SEED, 0, spp)
    rays0 = rays.cpu().numpy().T.astype(np.float64)
    out = ctx.trace_paths(integ, sc, cam, SEED, rays, spp, want_final=want_final)
    hits = M.hits_to_numpy(out[0])
    steps = out[1].cpu().numpy().astype(np.int64)
    final = out[2].cpu().numpy().T if want_final else None
    return rays0, hits, steps, final, ctx.stats()


@pytest.mark.parametrize("precision", [M.PREC_FP64, M.PREC_MIXED, M.PREC_FP32_REFERENCE])
@pytest.mark.parametrize("max_reflections", [0, 1])
def test_first_segments_match_oracle(gpu_ctx, precision, max_reflections):
    integ, sc, spec, cam, spp = setup(M.KERR, precision, max_reflections=max_reflections)
    _, hits, steps, _, _ = gpu_paths(gpu_ctx, integ, sc, cam, spp)
    _, (ref, ref_steps, refl) = oracle_paths(integ, sc, cam, spp, precision)
    # most first segments end on a wall: 4 % of the paths reach the LED directly, 10 % after one reflection
    assert (hits["cls"] == M.HIT_BOX_EMITTER).mean() > (0.02 if max_reflections == 0 else 0.05)
    agree = hits["cls"] == ref["cls"]
    bar = 0.999 if max_reflections == 0 else (0.995 if precision == M.PREC_FP64 else 0.99)
    em = agree & (ref["cls"] == M.HIT_BOX_EMITTER)
    # FP32 stages (mixed, FP32 reference): 2e-3 on the first segment; after a reflection the hit
    # point's FP32 rounding moves the second segment, and the tail grows to the level of the
    # oracle's own FP32-vs-FP64 spread (up to 1e-2), while the median stays at the 1e-4 level
    g_tol = 1e-6 if precision == M.PREC_FP64 else (2e-3 if max_reflections == 0 else 1e-2)
    err = np.abs(hits["g"][em] / ref["g"][em] - 1)
    print(f"class agreement {agree.mean():.5f}, emitter g error q99.9 {np.quantile(err, 0.999):.3g}, "
          f"median {np.median(err):.3g}")
    assert agree.mean() >= bar, agree.mean()
    assert np.quantile(err, 0.999) <= g_tol, np.quantile(err, 0.999)
    assert np.median(err) <= (1e-7 if precision == M.PREC_FP64 else 5e-4)
    if max_reflections == 0:
        assert (hits["cls"] == M.HIT_BOX_DIFFUSE).any()


def test_flat_space_paths_see_no_redshift(gpu_ctx):
    for precision in (M.PREC_FP64, M.PREC_FP32_REFERENCE):
        integ, sc, spec, cam, spp = setup(M.MINKOWSKI, precision)
        _, hits, _, _, st = gpu_paths(gpu_ctx, integ, sc, cam, spp)
        em = hits["cls"] == M.HIT_BOX_EMITTER
        assert em.mean() > 0.3 and st.n_reflections > hits.size
        np.testing.assert_allclose(hits["g"][em], 1.0, rtol=1e-6)


@pytest.mark.parametrize("precision", [M.PREC_FP64, M.PREC_MIXED])
def test_schwarzschild_paths_keep_the_static_redshift(gpu_ctx, precision):
    """Every emitter hit, after any number of reflections off static matter:
    g = sqrt(g_tt(emitter) / g_tt(camera)), with the emitter radius from the
    path's terminal state."""
    integ, sc, spec, cam, spp = setup(M.SCHWARZSCHILD, precision, tol=1e-8)
    rays0, hits, _, final, st = gpu_paths(gpu_ctx, integ, sc, cam, spp, want_final=True)
    em = hits["cls"] == M.HIT_BOX_EMITTER
    assert em.sum() > 500 and st.n_reflections > hits.size
    r_o = np.linalg.norm(rays0[em, :3], axis=1)
    r_e = np.linalg.norm(final[em, :3], axis=1)
    expect = np.sqrt((1 - 2 / r_e) / (1 - 2 / r_o))
    # g is stored in FP32; the mixed path finds the chord point with an FP32 box test
    tol = 2e-6 if precision == M.PREC_FP64 else 2e-5
    err = np.abs(hits["g"][em] / expect - 1)
    assert np.quantile(err, 0.99) <= tol, np.quantile(err, 0.99)


@pytest.mark.parametrize("precision", [M.PREC_FP64, M.PREC_MIXED])
def test_whole_path_outcomes_match_oracle_statistically(gpu_ctx, precision):
    integ, sc, spec, cam, spp = setup(M.KERR, precision, W=64, H=54, spp=4)
    _, hits, steps, _, st = gpu_paths(gpu_ctx, integ, sc, cam, spp)
    _, (ref, ref_steps, refl) = oracle_paths(integ, sc, cam, spp, M.PREC_FP64)
    n = hits.size
    for cls in (M.HIT_BOX_EMITTER, M.HIT_BOX_DIFFUSE, M.HIT_HORIZON, M.HIT_STEP_LIMIT):
        p_gpu, p_ref = (hits["cls"] == cls).mean(), (ref["cls"] == cls).mean()
        sigma = math.sqrt(max(p_ref * (1 - p_ref), 1.0 / n) * 2.0 / n)
        assert abs(p_gpu - p_ref) <= 4.5 * sigma, (cls, p_gpu, p_ref)
    # reflections per path
    mean_ref = refl.mean()
    assert st.n_reflections / n == pytest.approx(mean_ref, rel=0.05)
    # emitter redshift distribution
    g_gpu, g_ref = hits["g"][hits["cls"] == M.HIT_BOX_EMITTER], ref["g"][ref["cls"] == M.HIT_BOX_EMITTER]
    assert g_gpu.mean() == pytest.approx(g_ref.mean(), rel=2e-3)


def test_paths_are_deterministic(gpu_ctx):
    """FP64 whole paths are reproducible bit for bit (every decision is per
    lane).  The mixed path renormalises on a wave-uniform schedule, so its
    results depend on which rays share a wave, decided at run time: its first
    segments agree to FP32 rounding."""
    integ, sc, spec, cam, spp = setup(M.KERR, M.PREC_FP64)
    _, h1, s1, _, _ = gpu_paths(gpu_ctx, integ, sc, cam, spp)
    _, h2, s2, _, _ = gpu_paths(gpu_ctx, integ, sc, cam, spp)
    assert np.array_equal(h1, h2) and np.array_equal(s1, s2)
    integ, sc, spec, cam, spp = setup(M.KERR, M.PREC_MIXED, max_reflections=0)
    _, h1, _, _, _ = gpu_paths(gpu_ctx, integ, sc, cam, spp)
    _, h2, _, _, _ = gpu_paths(gpu_ctx, integ, sc, cam, spp)
    same = h1["cls"] == h2["cls"]
    assert same.mean() >= 0.999
    em = same & (h1["cls"] == M.HIT_BOX_EMITTER)
    assert np.quantile(np.abs(h1["g"][em] / h2["g"][em] - 1), 0.99) <= 1e-5


def test_box_frame_renders_and_shades_like_the_oracle(gpu_ctx):
    """mgk_render_frame follows whole paths; shading its hits matches the
    oracle's shading of the same hits (the LED emitter's spectrum and g^5)."""
    import torch
    integ, sc, spec, cam, spp = setup(M.KERR, M.PREC_MIXED, W=64, H=54, spp=2)
    gpu_ctx.set_scene(sc)
    xyz = gpu_ctx.render_frame(integ, sc, spec, cam, SEED, spp)
    torch.cuda.synchronize()
    f = xyz.cpu().numpy()
    assert np.isfinite(f).all() and f[..., 1].mean() > 0
    _, hits, _, _, _ = gpu_paths(gpu_ctx, integ, sc, cam, spp)
    h = torch.from_numpy(np.ascontiguousarray(hits).view(np.int32).reshape(-1, 4)).to("cuda:0")
    mine = gpu_ctx.shade(sc, spec, cam, SEED, 0, spp, h).cpu().numpy().reshape(-1, 3).astype(np.float64)
    W, _ = M.color_weights(spec)
    npx = cam.x_resolution * cam.y_resolution
    ref = O.shade(sc, spec, cam, SEED, 0, spp, np.arange(npx), hits, W)
    scale = np.maximum(np.abs(ref), 1e-6 * np.abs(ref).max())
    assert np.quantile(np.abs(mine - ref) / scale, 0.999) <= 1e-4
    # the frame is the shading of the same paths
    np.testing.assert_allclose(f.reshape(-1, 3), mine, rtol=1e-6, atol=1e-6 * np.abs(mine).max())


def test_invalid_box_scenes_are_rejected(gpu_ctx):
    integ, sc, spec, cam, spp = setup(M.KERR, M.PREC_MIXED, W=8, H=8, spp=1)
    rays = gpu_ctx.generate_rays(integ, cam, SEED, 0, spp)
    bad = M.workload("box_led", W=8, H=8, spp=1)[1]
    bad.box[1].matter = 7
    with pytest.raises(RuntimeError):
        gpu_ctx.trace_paths(integ, bad, cam, SEED, rays, spp)
    bad = M.workload("box_led", W=8, H=8, spp=1)[1]
    bad.n_box = 5
    with pytest.raises(RuntimeError):
        gpu_ctx.trace_paths(integ, bad, cam, SEED, rays, spp)
