"""GPU parity of triangle meshes (mgk_set_meshes, SCENE_MESHES): the two-level
hierarchy and the affine and Hermite chords in the persistent kernels,
against the oracle's brute-force restatement, through the C-ABI.

- Single segments (no reflection): outcome, triangle and instance agree hit
  by hit, and the hit points agree to the precision of the path.
- A hierarchy of 20,480 triangles in flat space: every hit the oracle finds
  over all triangles, the traversal finds, on the same triangle.
- Whole paths: the flat-space and Schwarzschild redshift invariants hold per
  path after any number of reflections off mesh matter, and the outcome
  fractions agree with the oracle's within binomial noise."""
import math

import numpy as np
import pytest

import magik_amd as M
from magik_amd import mesh
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SEED = 23


def set_both(ctx, meshes, instances):
    ctx.set_meshes(meshes, instances)
    O.set_meshes(meshes, instances)


def small_meshes():
    return [mesh.icosphere(3), mesh.grid_plane(1.0, 16)], [
        (0, M.MATTER_DIFFUSE, mesh.affine(1.5, (-4.0, -2.0, -3.0))),
        (0, M.MATTER_DIFFUSE, mesh.affine(2.0, (4.5, 1.5, -7.0))),
        (1, M.MATTER_DIFFUSE, mesh.affine((8.0, 1.0, 8.0), (0.0, -6.0, -5.0), rotate_y_deg=30.0)),
        (0, M.MATTER_ILLUMINANT, mesh.affine(0.8, (0.0, 4.0, -4.0)))]


def setup(metric, precision, chord, W=48, H=40, spp=1, max_reflections=0, tol=None):
    integ, sc, spec, cam, _ = M.workload("box_mesh", W=W, H=H, spp=spp, precision=precision)
    integ.metric_type = metric
    integ.spin = 0.999 if metric == M.KERR else 0.0
    if tol is not None:
        integ.abs_tol = integ.rel_tol = tol
    sc.chord = chord
    sc.max_reflections = max_reflections
    return integ, sc, spec, cam, spp


def oracle_prec(precision):
    return O.F32 if precision == M.PREC_FP32_REFERENCE else O.F64


def run_both(ctx, integ, sc, cam, spp, precision, oracle_precision=None):
    ctx.set_scene(sc)
    rays = ctx.generate_rays(integ, cam, SEED, 0, spp)
    rays0 = rays.cpu().numpy().T.astype(np.float64)
    hits_t, steps_t, final_t = ctx.trace_paths(integ, sc, cam, SEED, rays, spp, want_final=True)
    hits = M.hits_to_numpy(hits_t)
    final = final_t.cpu().numpy().T
    n = cam.x_resolution * cam.y_resolution
    pix = np.tile(np.arange(n, dtype=np.uint32), spp)
    smp = np.repeat(np.arange(spp, dtype=np.uint32), n)
    op = oracle_prec(precision) if oracle_precision is None else oracle_precision
    orays = O.generate_rays(op, integ, cam, SEED, pix, smp)
    ref, _, refl, at = O.trace_paths(op, integ, sc, cam, SEED, 0, orays, want_at=True)
    return rays0, hits, final, ref, refl, at, ctx.stats()


def on_triangle(meshes, instances, hits, final, L):
    """(plane distance in metres, barycentric u, v) of each mesh hit's
    terminal point on the triangle its record names (a0 triangle, a1 instance)."""
    out = []
    for h, x in zip(hits, final):
        v, f = meshes[instances[int(h["a1"])][0]]
        Mx = np.asarray(instances[int(h["a1"])][2], np.float64)
        xo = np.linalg.solve(Mx[:, :3], x[:3] * L - Mx[:, 3])
        t = v[f[int(h["a0"])]].astype(np.float64)
        e1, e2 = t[1] - t[0], t[2] - t[0]
        n = np.cross(e1, e2)
        n /= np.linalg.norm(n)
        u, w, dist = np.linalg.solve(np.stack([e1, e2, n], 1), xo - t[0])
        out.append((abs(dist) * np.linalg.norm(Mx[:, :3] @ n), u, w))
    return np.array(out).reshape(-1, 3)


@pytest.mark.parametrize("chord", [M.CHORD_AFFINE, M.CHORD_HERMITE])
@pytest.mark.parametrize("precision", [M.PREC_FP64, M.PREC_MIXED, M.PREC_FP32_REFERENCE])
def test_first_segments_match_oracle(gpu_ctx, precision, chord):
    """FP64: hit by hit against the oracle.  FP32 stages (mixed, FP32
    reference): their trajectories and step sequences part from the oracle's,
    and a chord's hit point moves with its end points, so the hits are compared
    by outcome and by distance, and each hit point is checked to lie on the
    triangle its record names.  Tolerance 1e-8 keeps the steps short enough
    for the distances to measure the chords rather than the integration."""
    meshes, inst = small_meshes()
    set_both(gpu_ctx, meshes, inst)
    integ, sc, spec, cam, spp = setup(M.KERR, precision, chord, tol=1e-8)
    print(f"case: precision {precision} chord {chord}", flush=True)
    _, hits, final, ref, _, at, _ = run_both(gpu_ctx, integ, sc, cam, spp, precision)
    mesh_hit = np.isin(ref["cls"], [M.HIT_MESH_DIFFUSE, M.HIT_MESH_EMITTER])
    assert mesh_hit.mean() > 0.05
    agree = hits["cls"] == ref["cls"]
    fp64 = precision == M.PREC_FP64
    bar = 0.999 if fp64 else 0.995
    both = agree & mesh_hit
    same_tri = (hits["a0"][both] == ref["a0"][both]) & (hits["a1"][both] == ref["a1"][both])
    d = np.linalg.norm(final[both, :3] - at[both, :3], axis=1)
    geo = on_triangle(meshes, inst, hits[both], final[both], M.length_scale(integ.mass_kg))
    print(f"class agreement {agree.mean():.5f}, mesh hits {mesh_hit.mean():.3f}, same triangle {same_tri.mean():.4f}, "
          f"hit point distance median {np.median(d):.3g} q99 {np.quantile(d, 0.99):.3g} (M), plane distance max "
          f"{geo[:, 0].max():.3g} m")
    assert agree.mean() >= bar
    # the terminal point lies on the named triangle (FP32 chord parameter in all but the FP64 path)
    assert geo[:, 0].max() <= (1e-5 if fp64 else 1e-3)   # the library holds the inverse maps in float
    eps = 1e-6 if fp64 else 2e-3
    assert (geo[:, 1] >= -eps).all() and (geo[:, 2] >= -eps).all() and (geo[:, 1] + geo[:, 2] <= 1 + eps).all()
    if fp64:
        assert same_tri.mean() > 0.99
        assert np.quantile(d, 0.99) <= 1e-6
    else:
        assert np.median(d) <= 1e-3 and np.quantile(d, 0.99) <= 3e-2
    em = both & (ref["cls"] == M.HIT_MESH_EMITTER)
    if em.any():
        err = np.abs(hits["g"][em] / ref["g"][em] - 1)
        assert np.quantile(err, 0.99) <= (1e-6 if fp64 else 2e-3)


def test_hierarchy_finds_every_brute_force_hit(gpu_ctx):
    """Flat space, 20,480 triangles in two instances: the GPU's two-level
    traversal against the oracle's loop over every triangle."""
    meshes = [mesh.icosphere(5)]
    inst = [(0, M.MATTER_DIFFUSE, mesh.affine(4.0, (2.0, 0.0, 1.0))),
            (0, M.MATTER_DIFFUSE, mesh.affine((3.0, 1.5, 3.0), (-6.0, 1.0, -4.0), rotate_y_deg=20.0))]
    set_both(gpu_ctx, meshes, inst)
    integ, sc, spec, cam, spp = setup(M.MINKOWSKI, M.PREC_FP64, M.CHORD_AFFINE, W=64, H=48)
    _, hits, final, ref, _, at, _ = run_both(gpu_ctx, integ, sc, cam, spp, M.PREC_FP64)
    m = ref["cls"] == M.HIT_MESH_DIFFUSE
    assert m.sum() > 1000
    assert np.array_equal(hits["cls"], ref["cls"])
    same = (hits["a0"][m] == ref["a0"][m]) & (hits["a1"][m] == ref["a1"][m])
    assert same.mean() > 0.999     # ties on shared edges may pick either triangle
    np.testing.assert_allclose(final[m, :3], at[m, :3], atol=1e-7)


@pytest.mark.parametrize("precision", [M.PREC_FP64, M.PREC_MIXED])
def test_paths_off_mesh_matter_keep_the_redshift_invariants(gpu_ctx, precision):
    meshes, inst = small_meshes()
    set_both(gpu_ctx, meshes, inst)
    for metric in (M.MINKOWSKI, M.SCHWARZSCHILD):
        if metric == M.MINKOWSKI and precision == M.PREC_MIXED:
            continue   # the mixed path of flat space is the FP64 one
        integ, sc, spec, cam, spp = setup(metric, precision, M.CHORD_HERMITE, W=64, H=54, spp=2, max_reflections=10, tol=1e-8)
        gpu_ctx.set_scene(sc)
        rays = gpu_ctx.generate_rays(integ, cam, SEED, 0, spp)
        rays0 = rays.cpu().numpy().T.astype(np.float64)
        hits_t, _, final_t = gpu_ctx.trace_paths(integ, sc, cam, SEED, rays, spp, want_final=True)
        hits = M.hits_to_numpy(hits_t)
        final = final_t.cpu().numpy().T
        st = gpu_ctx.stats()
        em = hits["cls"] == M.HIT_MESH_EMITTER
        assert em.sum() > 30 and st.n_reflections > hits.size
        if metric == M.MINKOWSKI:
            np.testing.assert_allclose(hits["g"][em], 1.0, rtol=1e-6)
        else:
            r_o = np.linalg.norm(rays0[em, :3], axis=1)
            r_e = np.linalg.norm(final[em, :3], axis=1)
            expect = np.sqrt((1 - 2 / r_e) / (1 - 2 / r_o))
            err = np.abs(hits["g"][em] / expect - 1)
            assert np.quantile(err, 0.99) <= (2e-6 if precision == M.PREC_FP64 else 2e-5), np.quantile(err, 0.99)


@pytest.mark.parametrize("precision", [M.PREC_FP64, M.PREC_MIXED])
def test_whole_path_outcomes_match_oracle_statistically(gpu_ctx, precision):
    meshes, inst = small_meshes()
    set_both(gpu_ctx, meshes, inst)
    integ, sc, spec, cam, spp = setup(M.KERR, precision, M.CHORD_HERMITE, W=64, H=54, spp=4, max_reflections=10)
    _, hits, _, ref, refl, _, st = run_both(gpu_ctx, integ, sc, cam, spp, precision, oracle_precision=O.F64)
    n = hits.size
    for cls in (M.HIT_MESH_EMITTER, M.HIT_BOX_EMITTER, M.HIT_MESH_DIFFUSE, M.HIT_BOX_DIFFUSE, M.HIT_HORIZON):
        p_gpu, p_ref = (hits["cls"] == cls).mean(), (ref["cls"] == cls).mean()
        sigma = math.sqrt(max(p_ref * (1 - p_ref), 1.0 / n) * 2.0 / n)
        assert abs(p_gpu - p_ref) <= 4.5 * sigma, (cls, p_gpu, p_ref)
    assert st.n_reflections / n == pytest.approx(refl.mean(), rel=0.05)


def test_mesh_frame_renders_and_is_deterministic(gpu_ctx):
    import torch
    meshes, inst = small_meshes()
    gpu_ctx.set_meshes(meshes, inst)
    # FP64: every decision per lane, so whole paths repeat bit for bit
    integ, sc, spec, cam, spp = setup(M.KERR, M.PREC_FP64, M.CHORD_HERMITE, W=64, H=54, spp=2, max_reflections=10)
    gpu_ctx.set_scene(sc)
    a = gpu_ctx.render_frame(integ, sc, spec, cam, SEED, spp)
    b = gpu_ctx.render_frame(integ, sc, spec, cam, SEED, spp)
    torch.cuda.synchronize()
    fa, fb = a.cpu().numpy(), b.cpu().numpy()
    assert np.isfinite(fa).all() and fa[..., 1].mean() > 0
    assert np.array_equal(fa, fb)


def test_invalid_meshes_are_rejected(gpu_ctx):
    v, f = mesh.box_mesh((0, 0, 0), (1, 1, 1))
    good = (0, M.MATTER_DIFFUSE, mesh.affine())
    with pytest.raises(RuntimeError):
        gpu_ctx.set_meshes([(v, np.where(f == 7, 9, f))], [good])                      # vertex index out of range
    with pytest.raises(RuntimeError):
        gpu_ctx.set_meshes([(v, f)], [(0, 5, mesh.affine())])                           # matter
    with pytest.raises(RuntimeError):
        gpu_ctx.set_meshes([(v, f)], [(1, M.MATTER_DIFFUSE, mesh.affine())])            # mesh index
    with pytest.raises(RuntimeError):
        gpu_ctx.set_meshes([(v, f)], [(0, M.MATTER_DIFFUSE, mesh.affine((1.0, 0.0, 1.0)))])   # singular map
    # a scene with meshes on a context without them, and an unknown chord mode
    gpu_ctx.set_meshes([], [])
    integ, sc, spec, cam, spp = setup(M.KERR, M.PREC_MIXED, M.CHORD_AFFINE, W=8, H=8)
    rays = gpu_ctx.generate_rays(integ, cam, SEED, 0, spp)
    with pytest.raises(RuntimeError):
        gpu_ctx.trace_paths(integ, sc, cam, SEED, rays, spp)
    gpu_ctx.set_meshes([(v, f)], [good])
    sc.chord = 2
    with pytest.raises(RuntimeError):
        gpu_ctx.trace_paths(integ, sc, cam, SEED, rays, spp)
    gpu_ctx.set_meshes([], [])
