"""Pins of the oracle's triangle-mesh pieces: the segment-triangle test, the
affine and Hermite chords, instancing and mesh reflection.  The reference
reserves these stages (wavefront.cpp:94-100) without implementing them, so
each piece is checked against mathematics it does not restate:
- the Moller-Trumbore test, against a linear solve of o + t d = v0 + u e1 + v e2;
- a tessellated box, against the analytic box's slab test, hit by hit;
- an instanced mesh, against its vertices transformed on the host;
- the Hermite chord in flat space (straight geodesics), against the affine chord;
- the Hermite chord in Schwarzschild, against the crossing point of a finely
  stepped geodesic: it must be far closer than the affine chord's;
- the Hermite curve itself, against a cubic it must reproduce exactly;
- whole paths off mesh matter: g = 1 in flat space, g = sqrt(g_tt(e)/g_tt(o))
  in Schwarzschild."""
import math

import numpy as np
import pytest

import magik_amd as M
from magik_amd import mesh
from oracle import oracle as O

C_LIGHT, G = 299792458.0, 6.67430e-11


def solve_tri(o, d, v0, e1, e2):
    A = np.stack([-d, e1, e2], 1)
    if abs(np.linalg.det(A)) < 1e-12:
        return None
    t, u, v = np.linalg.solve(A, o - v0)
    return t, u, v


@pytest.mark.parametrize("prec", [O.F32, O.F64])
def test_seg_tri_matches_a_linear_solve(prec):
    rng = np.random.default_rng(3)
    n_hit = 0
    for _ in range(3000):
        v0, e1, e2 = rng.normal(size=3), rng.normal(size=3), rng.normal(size=3)
        o = rng.normal(size=3) * 2
        d = rng.normal(size=3) * 3
        if rng.random() < 0.5:
            # aimed at a point of the triangle's plane near it, reached at t in (0, 1.2)
            uu, vv = rng.uniform(-0.2, 1.0, 2)
            d = (v0 + uu * e1 + vv * e2 - o) / rng.uniform(0.05, 1.2)
        if prec == O.F32:
            v0, e1, e2, o, d = (x.astype(np.float32).astype(np.float64) for x in (v0, e1, e2, o, d))
        ref = solve_tri(o, d, v0, e1, e2)
        t = O.seg_tri(prec, o, d, v0, e1, e2, 1.0)
        if ref is None:
            continue
        rt, u, v = ref
        margin = 1e-3 if prec == O.F32 else 1e-9
        inside = u > margin and v > margin and u + v < 1 - margin and margin < rt < 1 - margin
        outside = u < -margin or v < -margin or u + v > 1 + margin or rt < -margin or rt > 1 + margin
        if inside:
            n_hit += 1
            assert t is not None
            assert t == pytest.approx(rt, rel=1e-4 if prec == O.F32 else 1e-10, abs=1e-6 if prec == O.F32 else 1e-12)
        elif outside:
            assert t is None
    assert n_hit > 300


def rays_of(integ, cam, n_pix):
    pix = np.arange(n_pix, dtype=np.uint32)
    return O.generate_rays(O.F64, integ, cam, 5, pix, np.zeros(n_pix, dtype=np.uint32))


def box_scene_pair(metric=M.MINKOWSKI):
    integ, sc, spec, cam, _ = M.workload("box_led", W=40, H=32, spp=1, precision=M.PREC_FP64)
    integ.metric_type = metric
    integ.spin = 0.999 if metric == M.KERR else 0.0
    box = M.scene(celestial_radius_m=1000.0, flags=M.SCENE_BOXES, max_reflections=0,
                  boxes=[((-20.0, -10.0, -20.0), (20.0, 10.0, 20.0), M.MATTER_DIFFUSE)])
    tri = M.scene(celestial_radius_m=1000.0, flags=M.SCENE_MESHES, max_reflections=0)
    return integ, cam, box, tri


@pytest.mark.parametrize("metric", [M.MINKOWSKI, M.KERR])
def test_tessellated_box_hits_like_the_analytic_box(metric):
    integ, cam, box, tri = box_scene_pair(metric)
    O.set_meshes([mesh.box_mesh((-20, -10, -20), (20, 10, 20))], [(0, M.MATTER_DIFFUSE, mesh.affine())])
    rays = rays_of(integ, cam, 40 * 32)
    hb, _, _, ab = O.trace_paths(O.F64, integ, box, cam, 5, 0, rays, want_at=True)
    hm, _, _, am = O.trace_paths(O.F64, integ, tri, cam, 5, 0, rays, want_at=True)
    wall = hb["cls"] == M.HIT_BOX_DIFFUSE
    assert wall.mean() > 0.9
    # every wall hit is a mesh hit, every other outcome (the horizon in Kerr) the same
    assert np.array_equal(hm["cls"] == M.HIT_MESH_DIFFUSE, wall)
    assert np.array_equal(hm["cls"][~wall], hb["cls"][~wall])
    # the same chord crossing the same planes: the same points (the mesh stores its vertices in float)
    np.testing.assert_allclose(am[wall, :3], ab[wall, :3], rtol=0, atol=1e-6)
    np.testing.assert_allclose(am[wall, 4:], ab[wall, 4:], rtol=0, atol=0)
    assert (hm["a1"][wall] == 0).all() and ((hm["a0"][wall] >= 0) & (hm["a0"][wall] < 12)).all()
    O.set_meshes([], [])


def test_instance_map_equals_pre_transformed_vertices():
    v, f = mesh.icosphere(2)
    Mx = mesh.affine((1.5, 0.7, 2.0), (3.0, -1.0, -6.0), rotate_y_deg=35.0)
    vw = (v.astype(np.float64) @ Mx[:, :3].T + Mx[:, 3]).astype(np.float32)
    rng = np.random.default_rng(5)
    hits_a, hits_b = [], []
    for meshes, inst, out in (([(v, f)], [(0, M.MATTER_DIFFUSE, Mx)], hits_a), ([(vw, f)], [(0, M.MATTER_DIFFUSE, mesh.affine())], hits_b)):
        O.set_meshes(meshes, inst)
        r = np.random.default_rng(5)
        for _ in range(400):
            p0 = np.array([3.0, -1.0, -6.0]) + r.normal(size=3) * 0.5 + np.array([0, 0, 8.0])
            p1 = np.array([3.0, -1.0, -9.0]) + r.normal(size=3) * 1.0
            out.append(O.mesh_chord(O.F64, p0, p1))
    n = 0
    for a, b in zip(hits_a, hits_b):
        assert (a is None) == (b is None) or abs((a or b)[0] - 1) < 1e-6
        if a is not None and b is not None:
            n += 1
            assert a[3] == b[3]
            assert a[0] == pytest.approx(b[0], abs=1e-5)
            np.testing.assert_allclose(a[1], b[1], atol=1e-4)
    assert n > 200
    O.set_meshes([], [])


def test_hermite_curve_reproduces_a_cubic_and_its_plane_crossing():
    # x(s) = a + b s + c s^2 + e s^3 is its own cubic Hermite interpolant; the
    # crossing of the plane z = z0 is then found to rounding by Newton's method
    a, b, c, e = np.array([0.1, 0.2, 3.0]), np.array([0.5, -0.3, -2.0]), np.array([0.2, 0.1, -1.5]), np.array([0.0, 0.2, 0.4])
    x = lambda s: a + b * s + c * s * s + e * s ** 3
    dx = lambda s: b + 2 * c * s + 3 * e * s * s
    z0 = 1.2
    quad = (np.array([[-5, -5, z0], [5, -5, z0], [5, 5, z0], [-5, 5, z0]], np.float32), np.array([[0, 1, 2], [0, 2, 3]], np.uint32))
    O.set_meshes([quad], [(0, M.MATTER_DIFFUSE, mesh.affine())])
    r = O.mesh_chord(O.F64, x(0), x(1), dx(0), dx(1), hermite=True)
    zf = float(np.float32(z0))   # the plane as stored
    s_true = [s for s in np.roots([e[2], c[2], b[2], a[2] - zf]) if abs(s.imag) < 1e-12 and 0 <= s.real <= 1][0].real
    assert r is not None
    assert r[0] == pytest.approx(s_true, abs=1e-12)
    np.testing.assert_allclose(r[1], x(s_true), atol=1e-12)
    # the affine chord crosses at the straight line's parameter
    ra = O.mesh_chord(O.F64, x(0), x(1))
    assert ra[0] == pytest.approx((zf - x(0)[2]) / (x(1)[2] - x(0)[2]), abs=1e-12)
    O.set_meshes([], [])


def test_hermite_chord_is_the_affine_chord_in_flat_space():
    integ, cam, _, tri = box_scene_pair(M.MINKOWSKI)
    O.set_meshes([mesh.icosphere(3, 4.0)], [(0, M.MATTER_DIFFUSE, mesh.affine(1.0, (1.0, 0.5, -2.0)))])
    rays = rays_of(integ, cam, 40 * 32)
    ha, _, _, aa = O.trace_paths(O.F64, integ, tri, cam, 5, 0, rays, want_at=True)
    tri.chord = M.CHORD_HERMITE
    hh, _, _, ah = O.trace_paths(O.F64, integ, tri, cam, 5, 0, rays, want_at=True)
    assert (ha["cls"] == M.HIT_MESH_DIFFUSE).sum() > 150
    assert np.array_equal(ha["cls"], hh["cls"]) and np.array_equal(ha["a0"], hh["a0"])
    m = ha["cls"] == M.HIT_MESH_DIFFUSE
    np.testing.assert_allclose(ah[m, :3], aa[m, :3], atol=1e-9)
    O.set_meshes([], [])


def test_hermite_chord_follows_the_curved_geodesic():
    """Schwarzschild, M = 1 m: rays from r = 20 M pass the hole and cross the
    plane z = -8 M.  With steps up to 40 M the affine chord misses the crossing
    by the geodesic's sagitta; the Hermite chord must be far closer to the
    crossing of the same ray stepped finely."""
    mass = C_LIGHT ** 2 / G   # GM/c^2 = 1 m
    integ = M.integrator(metric=M.SCHWARZSCHILD, precision=M.PREC_FP64, spin=0.0, mass_kg=mass, abs_tol=1e-6, rel_tol=1e-6,
                         step_limit=20000)
    integ.h_max = 40.0
    cam = M.camera_pinhole([0.0, 0.0, 20.0], [1, 0, 0, 0, 1, 0, 0, 0, 1], 40.0, 24, 24)
    quad = (np.array([[-400, -400, -8], [400, -400, -8], [400, 400, -8], [-400, 400, -8]], np.float32),
            np.array([[0, 1, 2], [0, 2, 3]], np.uint32))
    O.set_meshes([quad], [(0, M.MATTER_DIFFUSE, mesh.affine())])
    sc = M.scene(celestial_radius_m=5000.0, flags=M.SCENE_MESHES, max_reflections=0)
    pix = np.arange(24 * 24, dtype=np.uint32)
    rays = O.generate_rays(O.F64, integ, cam, 5, pix, np.zeros_like(pix))
    out = {}
    for name, chord, h_max, tol in (("affine", M.CHORD_AFFINE, 40.0, 1e-6), ("hermite", M.CHORD_HERMITE, 40.0, 1e-6),
                                    ("fine", M.CHORD_AFFINE, 0.02, 1e-12)):
        integ.h_max, integ.abs_tol, integ.rel_tol = h_max, tol, tol
        sc.chord = chord
        h, _, _, at = O.trace_paths(O.F64, integ, sc, cam, 5, 0, rays, want_at=True)
        out[name] = (h, at)
    m = (out["fine"][0]["cls"] == M.HIT_MESH_DIFFUSE) & (out["affine"][0]["cls"] == M.HIT_MESH_DIFFUSE) & \
        (out["hermite"][0]["cls"] == M.HIT_MESH_DIFFUSE)
    assert m.sum() > 250
    ref = out["fine"][1][m, :3]
    e_aff = np.linalg.norm(out["affine"][1][m, :3] - ref, axis=1)
    e_her = np.linalg.norm(out["hermite"][1][m, :3] - ref, axis=1)
    print(f"crossing error: affine median {np.median(e_aff):.3g} max {e_aff.max():.3g}, "
          f"hermite median {np.median(e_her):.3g} max {e_her.max():.3g}")
    assert np.median(e_aff) > 1e-3                      # the test bends the rays enough to tell the chords apart
    assert np.median(e_her) < np.median(e_aff) / 20
    # (the largest errors, the same for both, come from rays that pass close to the photon sphere,
    # where the coarse and fine integrations part before the plane)
    assert np.quantile(e_her, 0.9) < np.quantile(e_aff, 0.9) / 5
    O.set_meshes([], [])


def test_paths_off_mesh_matter_keep_the_static_redshift():
    """Whole paths in the box_mesh room: any number of reflections off static
    mesh and box matter leaves g = 1 in flat space and g = sqrt(g_tt(e)/g_tt(o))
    at the illuminant in Schwarzschild."""
    meshes, inst = M.workload_meshes("box_mesh")
    meshes = [mesh.icosphere(2), mesh.grid_plane(1.0, 8)]
    O.set_meshes(meshes, inst)
    for metric in (M.MINKOWSKI, M.SCHWARZSCHILD):
        integ, sc, spec, cam, _ = M.workload("box_mesh", W=32, H=24, spp=1, precision=M.PREC_FP64)
        integ.metric_type = metric
        integ.spin = 0.0
        integ.abs_tol = integ.rel_tol = 1e-8
        rays = rays_of(integ, cam, 32 * 24)
        h, _, refl, at = O.trace_paths(O.F64, integ, sc, cam, 5, 0, rays, want_at=True)
        em = h["cls"] == M.HIT_MESH_EMITTER
        assert (h["cls"] == M.HIT_MESH_DIFFUSE).any() or refl.max() > 0
        assert em.sum() > 3 and refl[em].max() >= 1
        if metric == M.MINKOWSKI:
            np.testing.assert_allclose(h["g"][em], 1.0, rtol=1e-6)
        else:
            r_o = np.linalg.norm(rays[em, :3], axis=1)      # the lens point, in units of M
            r_e = np.linalg.norm(at[em, :3], axis=1)
            expect = np.sqrt((1 - 2 / r_e) / (1 - 2 / r_o))
            np.testing.assert_allclose(h["g"][em], expect, rtol=2e-6)
    O.set_meshes([], [])


# ---- the library's hierarchy build, on the host (no device needed) ----
def test_hierarchy_build_is_sound_on_the_host():
    """mgk_mesh_hierarchy_info builds the hierarchy mgk_set_meshes would
    upload and walks it as the device does: skip links forward, children
    inside parents, every triangle in exactly one leaf, one instance per
    top-level leaf.  Degenerate inputs (all triangles at one point, a single
    triangle, 100k random triangles) and invalid ones are covered."""
    from magik_amd import abi
    rng = np.random.default_rng(9)
    n = 100_000
    c = rng.uniform(-10, 10, (n, 1, 3))
    v = (c + 0.1 * rng.uniform(-1, 1, (n, 3, 3))).reshape(-1, 3).astype(np.float32)
    f = np.arange(3 * n, dtype=np.uint32).reshape(-1, 3)
    point = (np.zeros((3, 3), np.float32), np.zeros((17, 3), np.uint32))
    cases = [([(v, f)], n), ([mesh.icosphere(4)], 5120), ([point], 17), ([mesh.box_mesh((0, 0, 0), (1, 1, 1))], 12)]
    for meshes, tris in cases:
        inst = [(0, M.MATTER_DIFFUSE, mesh.affine(1.0, (3.0 * k, 0.0, 0.0), rotate_y_deg=10.0 * k)) for k in range(5)]
        r, info = mesh.hierarchy_info(meshes, inst)
        assert r == abi.MAGIK_SUCCESS, (r, tris)
        assert info.n_triangles == tris and 1 <= info.max_leaf <= 8 and info.max_depth < 64
        assert info.world_max[0] > info.world_min[0]
    # a mesh that no instance uses is built and checked like the others
    r, info = mesh.hierarchy_info([mesh.icosphere(2), mesh.box_mesh((0, 0, 0), (1, 1, 1)), mesh.icosphere(1)],
                                  [(1, M.MATTER_DIFFUSE, mesh.affine())])
    assert r == abi.MAGIK_SUCCESS and info.n_triangles == 320 + 12 + 80
    v1, f1 = mesh.box_mesh((0, 0, 0), (1, 1, 1))
    good = (0, M.MATTER_DIFFUSE, mesh.affine())
    for meshes, inst in (([(v1, np.where(f1 == 7, 8, f1))], [good]),                         # vertex index out of range
                         ([(v1, f1)], [(0, 5, mesh.affine())]),                               # matter
                         ([(v1, f1)], [(1, M.MATTER_DIFFUSE, mesh.affine())]),                # mesh index
                         ([(v1, f1)], [(0, M.MATTER_DIFFUSE, mesh.affine((1.0, 0.0, 1.0)))])):  # singular map
        assert mesh.hierarchy_info(meshes, inst)[0] == abi.MGK_ERROR_INVALID_ARGUMENT
