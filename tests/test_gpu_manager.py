"""The render manager on the GPU (include/magik_manager.h): commands pushed
and dispatched from the DCC side are applied by the worker thread, which
renders the stage's box scene one sample per iteration; a fetched frame after
n samples equals the library's n-sample frame (mgk_render_frame, same
Philox samples) through the same display transform.  The progressive mean
sums per-sample XYZ where the frame sums inside the shading loop, so the two
agree to float rounding, not bit for bit."""
import time

import numpy as np
import pytest

import magik_amd as M
from magik_amd import abi

pytestmark = pytest.mark.gpu


def stage_structs(W, H, metric=M.KERR, reflections=10, white_K=5600.0):
    """The manager's stage defaults as the library's structs (its wavelengths
    drawn through the CMF inverse CDF, as the reference's stage draws them)."""
    integ, sc, _, cam, _ = M.workload("box_led", W=W, H=H, spp=1, precision=M.PREC_MIXED)
    integ.metric_type = metric
    sc.max_reflections = reflections
    spec = M.spectral(n_hero=4, hero_min=360.0, hero_max=830.0, n_bin=24, bin_min=360.0, bin_max=830.0, sigma=10.0,
                      white_K=white_K, lambda_sampling=M.LAMBDA_CMF)
    return integ, sc, spec, cam


def wait(pred, timeout=60.0, what=""):
    t0 = time.time()
    while time.time() - t0 < timeout:
        if pred():
            return
        time.sleep(0.01)
    raise AssertionError(f"timed out waiting for {what}")


def run_until_paused(m, min_samples):
    """Samples until at least min_samples, pause, and fetch the final frame."""
    def samples():
        st, _ = m.stats()
        return st.x_sample
    wait(lambda: samples() >= min_samples or m.error() != 0, what="samples")
    assert m.error() == 0, m.error()
    m.push(M.Manager.command("pause"))
    wait(m.dispatch, what="dispatch of the pause")
    time.sleep(0.3)
    n0 = samples()
    time.sleep(0.2)
    n = samples()
    assert n == n0, "still sampling after the pause"
    frame = None

    def fetched():
        nonlocal frame
        r, new = m.fetch()
        assert r == abi.MAGIK_SUCCESS
        if new:
            frame = m.extract("test")
        return frame is not None
    wait(fetched, what="a fetched frame")
    return n, frame


def direct(gpu_ctx, W, H, n, disp, **kw):
    import torch
    integ, sc, spec, cam = stage_structs(W, H, **kw)
    gpu_ctx.set_scene(sc)
    xyz = gpu_ctx.render_frame(integ, sc, spec, cam, 0, n)
    out = M.display_transform(gpu_ctx, disp, xyz)
    torch.cuda.synchronize()
    return out.cpu().numpy(), xyz.cpu().numpy()


def compare(a, b, xyz=None):
    assert a.shape == b.shape
    fin = np.isfinite(a) & np.isfinite(b)
    assert fin.mean() > 0.999
    # relative to the value: without a view transform display values reach 1e3 (out-of-gamut colours)
    d = np.where(fin, np.abs(a - b) / np.maximum(np.abs(b), 1e-2), 0.0)
    if xyz is not None:
        w = int(np.argmax(d.max(axis=-1)))
        y, x = divmod(w, a.shape[1])
        print(f"worst pixel ({y},{x}): manager {a[y, x]}, frame {b[y, x]}, frame XYZ {xyz[y, x]}")
    d = d[fin]
    assert np.quantile(d, 0.999) <= 1e-4 and np.median(d) <= 1e-6, (np.quantile(d, 0.999), np.median(d))


def test_progressive_frame_equals_the_n_sample_frame(gpu_ctx):
    W, H = 48, 40
    m = M.Manager(0)
    try:
        for cmd in (M.Manager.command("camera_resolution", W, H), M.Manager.aov_command("add_aov", "test", 3)):
            assert m.push(cmd) == abi.MAGIK_SUCCESS
        wait(m.dispatch, what="dispatch")
        n, frame = run_until_paused(m, 3)
        assert frame.shape == (H, W, 3) and n >= 3
        compare(frame, *direct(gpu_ctx, W, H, n, M.display()))

        # new settings restart the accumulation; render resumes
        cmds = [M.Manager.command("schwarzschild"), M.Manager.command("camera_resolution", 40, 32),
                M.Manager.command("view_transform", M.VIEW_NONE), M.Manager.command("display_colorspace", M.DISPLAY_SRGB),
                M.Manager.command("white_point", 6500.0), M.Manager.command("reflected_limit", 3),
                M.Manager.command("render")]
        for cmd in cmds:
            assert m.push(cmd) == abi.MAGIK_SUCCESS
        wait(m.dispatch, what="dispatch")
        n2, frame2 = run_until_paused(m, 2)
        assert frame2.shape == (32, 40, 3)
        disp = M.display(view_transform=M.VIEW_NONE, display_colorspace=M.DISPLAY_SRGB, white_point_K=6500.0)
        compare(frame2, *direct(gpu_ctx, 40, 32, n2, disp, metric=M.SCHWARZSCHILD, reflections=3, white_K=6500.0))
        st, _ = m.stats()
        assert st.total_wavefront_sample_time_seconds > 0 and st.latest_wavefront_history_length > 0
        print(f"manager: {n} + {n2} samples, latest sample {st.latest_wavefront_sample_time_seconds * 1e3:.2f} ms, "
              f"loop iteration {st.latest_render_loop_iteration_time_seconds * 1e3:.2f} ms")
    finally:
        m.close()


def test_queue_histories_follow_the_latest_samples_paths(gpu_ctx):
    """The statistics' per-iteration queue counts equal the counts that the
    same sample's paths give, traced directly with their step counts
    (mgk_trace_paths): advance_A[k] = paths with more than k steps,
    splat[k] = paths ending at step k + 1 other than by the step limit,
    advance_B[k] = advance_A[k + 1], and bxdf - splat = the reflections."""
    import torch
    W, H = 40, 32
    m = M.Manager(0)
    try:
        for cmd in (M.Manager.command("camera_resolution", W, H), M.Manager.aov_command("add_aov", "test", 3)):
            assert m.push(cmd) == abi.MAGIK_SUCCESS
        wait(m.dispatch, what="dispatch")
        n, _ = run_until_paused(m, 2)
        wait(lambda: m.stats()[1], what="statistics after the pause")
        st, _ = m.stats()
        q = m.queue_histories()
    finally:
        m.close()
    assert st.x_sample == n and st.h_latest_advance_A_queue_count_history is None
    length = st.latest_wavefront_history_length
    assert all(v.shape == (length,) for v in q.values())

    integ, sc, _, cam = stage_structs(W, H)
    gpu_ctx.set_scene(sc)
    rays = gpu_ctx.generate_rays(integ, cam, 0, n - 1, 1)
    hits, steps = gpu_ctx.trace_paths(integ, sc, cam, 0, rays, 1, sample_begin=n - 1)
    torch.cuda.synchronize()
    n_refl = gpu_ctx.stats().n_reflections
    s = np.maximum(steps.cpu().numpy().astype(np.int64), 1)
    lim = hits[:, 0].cpu().numpy() == M.HIT_STEP_LIMIT
    assert length == s.max() and n_refl > 0
    k = np.arange(length)
    alive = (s[None, :] > k[:, None]).sum(axis=1)
    ends = np.bincount(s[~lim], minlength=length + 1)[1:]
    np.testing.assert_array_equal(q["advance_A"], alive)
    np.testing.assert_array_equal(q["intersect"], alive)
    np.testing.assert_array_equal(q["advance_B"], np.append(alive[1:], 0))
    np.testing.assert_array_equal(q["splat"], ends)
    np.testing.assert_array_equal(q["redistribute"], ends)
    extra = q["bxdf"].astype(np.int64) - ends
    assert extra.min() >= 0 and extra.sum() == n_refl
    assert q["advance_A"][0] == W * H
    print(f"queue histories: {length} iterations, {lim.sum()} paths at the step limit, {n_refl} reflections")


def test_headless_manager_has_no_swapchain(gpu_ctx):
    m = M.Manager(0, display=abi.MAGIK_DISPLAY_HEADLESS)
    try:
        m.push(M.Manager.command("camera_resolution", 16, 16))
        m.push(M.Manager.aov_command("add_aov", "test", 3))
        wait(m.dispatch, what="dispatch")
        wait(lambda: m.stats()[0].x_sample >= 1, what="a sample")
        r, new = m.fetch()
        assert r == 902 and not new          # MAGIK_ERROR_AOV_SWAPCHAIN_NOT_INITALIZED
        assert m.error() == 0
    finally:
        m.close()


def test_no_test_aov_renders_nothing_and_removal_stops(gpu_ctx):
    m = M.Manager(0)
    try:
        m.push(M.Manager.command("camera_resolution", 16, 16))
        m.push(M.Manager.aov_command("add_aov", "other", 3))
        wait(m.dispatch, what="dispatch")
        time.sleep(0.3)
        assert m.stats()[0].x_sample == 0    # renderer.cpp:112-121: only the "test" AOV is rendered
        m.push(M.Manager.aov_command("add_aov", "test", 3))
        wait(m.dispatch, what="dispatch")
        wait(lambda: m.stats()[0].x_sample >= 2, what="samples")
        m.push(M.Manager.aov_command("remove_aov", "test", 3))
        wait(m.dispatch, what="dispatch")
        time.sleep(0.3)
        n = m.stats()[0].x_sample
        time.sleep(0.3)
        assert m.stats()[0].x_sample == n
        assert m.error() == 0
    finally:
        m.close()
