"""Pins of the oracle's box pieces, i.e. reflective matter and the bounce path
(traceables.cuh:17-74, bxdf_utilities.cuh:56-151, matter_examples.cuh:36-76,
brdf.cuh:12-29, kernel_enforce_ray_limits.cu:48-50).

Each piece is checked against something it does not restate:
- the slab test, against a face-by-face segment test;
- the white LED spectrum, against the reference (golden vectors);
- the self-intersection offset, against its ulp-count definition;
- one reflection in flat space, against closed forms;
- one reflection in Kerr, against the null condition; in Schwarzschild, against
  the unit local energy;
- the Lambertian lobe, against the moments of the cosine-weighted distribution.

Two invariants pin whole paths whatever their number of reflections off
static matter: g = 1 in flat space, and g = sqrt(g_tt(emitter) / g_tt(camera))
in Schwarzschild."""
import json
import math
import os

import numpy as np
import pytest

import magik_amd as M
from oracle import oracle as O

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.json")))
BMIN, BMAX = np.array([-20.0, -10.0, -20.0]), np.array([20.0, 10.0, 20.0])


def faces_hit(o, d, length, bmin, bmax):
    """First crossing of the box surface along the segment o + s d, 0 <= s <= length,
    by testing each face rectangle; (s, normal code) or None."""
    best = None
    for axis in range(3):
        for val in (bmin[axis], bmax[axis]):
            if d[axis] == 0:
                continue
            s = (val - o[axis]) / d[axis]
            if s < 0 or s > length:
                continue
            p = o + s * d
            if all(bmin[k] <= p[k] <= bmax[k] for k in range(3) if k != axis):
                if best is None or s < best[0]:
                    n = -1.0 if d[axis] > 0 else 1.0    # the face normal against the segment
                    best = (s, axis + (3 if n < 0 else 0))
    return best


@pytest.mark.parametrize("prec", [O.F32, O.F64])
def test_slab_test_matches_face_by_face_crossings(prec):
    rng = np.random.default_rng(11)
    n_hit = n_inside = 0
    for _ in range(4000):
        o = rng.uniform(-35, 35, 3)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        length = rng.uniform(0.5, 90.0)
        if prec == O.F32:
            o, d = o.astype(np.float32).astype(np.float64), d.astype(np.float32).astype(np.float64)
        ok, t, code = O.aabb_hit(prec, o, d, 1.0 / length, BMIN, BMAX)
        ref = faces_hit(o, d, length, BMIN, BMAX)
        # skip grazing cases within rounding of an edge or of the segment's end
        if ref is not None and (abs(ref[0] - length) < 1e-4 or
                                np.sort(np.abs(np.concatenate([o + ref[0] * d - BMIN, o + ref[0] * d - BMAX])))[1] < 1e-4):
            continue
        assert ok == (ref is not None), (o, d, length)
        if ok:
            n_hit += 1
            n_inside += bool(np.all((o > BMIN) & (o < BMAX)))
            assert t == pytest.approx(ref[0] / length, rel=1e-5 if prec == O.F32 else 1e-12, abs=1e-7)
            assert code == ref[1]
    assert n_hit > 600 and n_inside > 150


def test_white_led_equals_the_reference():
    for e in GOLD["white_led"]:
        ref = float(np.uint32(e["f32_bits"]).view(np.float32))
        assert O.white_led(O.F32, e["lam"]) == pytest.approx(ref, rel=2e-6, abs=1e-30)
        assert O.white_led(O.F64, e["lam"]) == pytest.approx(ref, rel=2e-6, abs=1e-30)


def test_offset_moves_along_the_normal_by_1024_ulps():
    for x in (5.0, -5.0, 37.25, -0.75, 0.01, -0.02, 0.0):
        for code in range(6):
            q = np.array([x, x, x, 3.0])
            out = O.offset_geodesic(q, code)
            axis, sgn = code % 3, (-1.0 if code >= 3 else 1.0)
            for k in range(3):
                if k != axis:
                    assert out[k] == q[k]
            xf, of = np.float32(x), np.float32(out[axis])
            if abs(x) < 1.0 / 32:
                assert out[axis] - x == pytest.approx(sgn * 4.0 / 65536.0, rel=1e-6)
            else:
                assert np.sign(out[axis] - x) == sgn                        # along the normal
                bits = int(np.array(of).view(np.int32)) - int(np.array(xf).view(np.int32))
                assert abs(bits) == 1024                                    # int(1024 |n|) units in the last place
            assert out[3] == q[3]


def test_flat_space_specular_reflection_closed_form():
    integ = M.integrator(metric=M.MINKOWSKI, precision=M.PREC_FP64, spin=0.0)
    at = np.array([1.0, 2.0, 3.0, 0.0, 0.6, 0.0, 0.8, -1.0])      # unit-energy null momentum, covariant
    out, c = O.reflect_diffuse(O.F64, integ, at, -1.0, 3, 0.3, 0.7)   # -x face: specular
    np.testing.assert_allclose(out[4:], [-0.6, 0.0, 0.8, -1.0], atol=1e-15)
    assert c == pytest.approx(-1.0, abs=1e-15)
    assert out[0] < 1.0 and out[1:3].tolist() == [2.0, 3.0]          # offset back into the room along -x


def test_flat_space_lambertian_reflection_closed_form():
    integ = M.integrator(metric=M.MINKOWSKI, precision=M.PREC_FP64, spin=0.0)
    at = np.array([1.0, -9.0, 3.0, 0.0, 0.0, -2.0, 0.0, -2.0])     # falling onto the floor, energy 2
    for u0, u1 in ((0.1, 0.2), (0.5, 0.9), (0.95, 0.05)):
        out, c = O.reflect_diffuse(O.F64, integ, at, -1.0, 1, u0, u1)   # +y face (the floor)
        th, ph = math.acos(1 - 2 * u0), 2 * float(np.float32(math.pi)) * u1   # the reference's pi is a float
        d = np.array([math.cos(ph) * math.sin(th), math.sin(ph) * math.sin(th), math.cos(th)]) + [0.0, 1.0, 0.0]
        d /= np.linalg.norm(d)
        np.testing.assert_allclose(out[4:7], d, atol=1e-14)
        assert out[7] == pytest.approx(-1.0)
        # z = -1 / (u.p_in) = -1 / -2; the new contraction (u.p_out) z = -1/2 carries the factor
        assert c == pytest.approx(-0.5, abs=1e-15)


def test_lambertian_lobe_is_cosine_weighted():
    integ = M.integrator(metric=M.MINKOWSKI, precision=M.PREC_FP64, spin=0.0)
    at = np.array([0.0, -9.0, 0.0, 0.0, 0.0, -1.0, 0.0, -1.0])
    k = 160
    cos = []
    for i in range(k):
        for j in range(k):
            out, _ = O.reflect_diffuse(O.F64, integ, at, -1.0, 1, (i + 0.5) / k, (j + 0.5) / k)
            cos.append(out[5])
    cos = np.array(cos)
    assert cos.min() >= 0.0
    assert cos.mean() == pytest.approx(2.0 / 3.0, abs=2e-3)      # E[cos] of p(w) = cos / pi
    assert (cos ** 2).mean() == pytest.approx(0.5, abs=2e-3)     # E[cos^2]


def test_kerr_reflection_is_null_and_leaves_the_wall():
    integ = M.integrator(metric=M.KERR, precision=M.PREC_FP64, spin=0.999)
    rng = np.random.default_rng(5)
    for _ in range(40):
        q = np.array([rng.uniform(-39, 39), rng.uniform(-19, 19), -40.0, 7.0])   # on the z = -40 M wall
        gi = O.metric_tensor(O.F64, integ, q, True)
        # an incoming null covector heading into the wall (-z)
        p = np.array([rng.normal(), rng.normal(), -abs(rng.normal()) - 0.5, 0.0])
        a, b, c = gi[3, 3], 2 * gi[3, :3] @ p[:3], p[:3] @ gi[:3, :3] @ p[:3]
        p[3] = (-b + math.sqrt(b * b - 4 * a * c)) / (2 * a)   # p_t < 0: u.p < 0, as for camera rays
        at = np.concatenate([q, p])
        for code, u in ((5, (0.3, 0.6)), (5, (0.8, 0.1))):
            out, cnew = O.reflect_diffuse(O.F64, integ, at, -1.0, code, *u)
            po = out[4:]
            # null at the hit point, where it is built (the start point is then moved 1024
            # ulps off the wall with the momentum kept, bxdf_utilities.cuh:136-151)
            assert abs(po @ gi @ po) < 1e-12 * (po @ po)
            pup = gi @ po
            assert pup[2] > 0                                              # leaves the wall into the room
            assert cnew < 0


def test_schwarzschild_reflection_has_unit_local_energy():
    """The matter's four-velocity is the static observer's, u = (0, 0, 0, 1/sqrt(1 - 2/r)):
    the outgoing momentum has u.p = -1, i.e. p_t = -sqrt(1 - 2/r), and the new
    contraction is (u.p_out) z = -c / (u.p_in)."""
    integ = M.integrator(metric=M.SCHWARZSCHILD, precision=M.PREC_FP64, spin=0.0)
    rng = np.random.default_rng(6)
    for _ in range(30):
        q = np.array([rng.uniform(-30, 30), -20.0, rng.uniform(-30, 30), 2.0])   # on the y = -20 M floor
        r = np.linalg.norm(q[:3])
        gi = O.metric_tensor(O.F64, integ, q, True)
        p = np.array([rng.normal(), -abs(rng.normal()) - 0.5, rng.normal(), 0.0])
        a, b, c = gi[3, 3], 2 * gi[3, :3] @ p[:3], p[:3] @ gi[:3, :3] @ p[:3]
        p[3] = (-b + math.sqrt(b * b - 4 * a * c)) / (2 * a)   # p_t < 0: u.p < 0, as for camera rays
        contraction = -1.7
        out, cnew = O.reflect_diffuse(O.F64, integ, np.concatenate([q, p]), contraction, 1, *rng.uniform(size=2))
        ut = 1 / math.sqrt(1 - 2 / r)
        assert out[7] * ut == pytest.approx(-1.0, rel=1e-12)
        assert cnew == pytest.approx(-contraction / (ut * p[3]), rel=1e-12)


def box_scene(metric, precision, max_reflections=10, spin=0.0):
    integ, sc, spec, cam, spp = M.workload("box_led", W=40, H=34, spp=1, precision=precision)
    integ.metric_type = metric
    integ.spin = spin
    sc.max_reflections = max_reflections
    return integ, sc, cam


def rays_of(prec, integ, cam, seed, spp=1):
    n = cam.x_resolution * cam.y_resolution
    pix = np.tile(np.arange(n, dtype=np.uint32), spp)
    smp = np.repeat(np.arange(spp, dtype=np.uint32), n)
    return O.generate_rays(prec, integ, cam, seed, pix, smp)


@pytest.mark.parametrize("prec", [O.F32, O.F64])
def test_flat_space_paths_see_no_redshift(prec):
    integ, sc, cam = box_scene(M.MINKOWSKI, M.PREC_FP64)
    rays = rays_of(prec, integ, cam, 3)
    hits, steps, refl = O.trace_paths(prec, integ, sc, cam, 3, 0, rays)
    em = hits["cls"] == M.HIT_BOX_EMITTER
    assert em.mean() > 0.3 and refl[em].max() >= 4
    np.testing.assert_allclose(hits["g"][em], 1.0, rtol=1e-6 if prec == O.F32 else 1e-12)
    assert set(np.unique(hits["cls"])) <= {M.HIT_BOX_EMITTER, M.HIT_BOX_DIFFUSE}


def test_schwarzschild_paths_keep_the_static_redshift():
    """Static matter in a static spacetime: each reflection resets the local
    energy to 1 and carries the ratio so far, so the emitter's g depends only
    on the end points, sqrt(g_tt(emitter) / g_tt(camera)), after any number of
    reflections.  A dropped factor z, or an unnormalised outgoing energy,
    fails this."""
    integ, sc, cam = box_scene(M.SCHWARZSCHILD, M.PREC_FP64)
    integ.abs_tol, integ.rel_tol = 1e-8, 1e-8
    rays = rays_of(O.F64, integ, cam, 5)
    hits, steps, refl, at = O.trace_paths(O.F64, integ, sc, cam, 5, 0, rays, want_at=True)
    em = hits["cls"] == M.HIT_BOX_EMITTER
    assert em.sum() > 300 and (refl[em] >= 3).sum() > 50
    r_o = np.linalg.norm(rays[em, :3], axis=1)
    r_e = np.linalg.norm(at[em, :3], axis=1)
    expect = np.sqrt((1 - 2 / r_e) / (1 - 2 / r_o))
    np.testing.assert_allclose(hits["g"][em], expect, rtol=1e-6)   # g is stored in float32


def test_reflection_limit_and_step_count():
    integ, sc, cam = box_scene(M.KERR, M.PREC_FP64, spin=0.999)
    rays = rays_of(O.F64, integ, cam, 9)
    for limit in (0, 3):
        sc.max_reflections = limit
        hits, steps, refl = O.trace_paths(O.F64, integ, sc, cam, 9, 0, rays)
        assert refl.max() == limit
        # a path still on diffuse matter has used up its reflections
        assert np.all(refl[hits["cls"] == M.HIT_BOX_DIFFUSE] == limit)
        assert np.all(steps >= refl + 1)
    sc.max_reflections = 10
    h1, s1, r1 = O.trace_paths(O.F64, integ, sc, cam, 9, 0, rays)
    h2, s2, r2 = O.trace_paths(O.F64, integ, sc, cam, 9, 0, rays)
    assert np.array_equal(h1, h2) and np.array_equal(r1, r2)
