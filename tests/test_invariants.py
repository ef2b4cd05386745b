"""Pins of oracle/invariants.py (the constants of motion used for the
constraint-drift measurements): closed forms, the oracle's own metric, and
conservation along FP64 oracle geodesics."""
import math

import numpy as np
import pytest

import magik_amd as M
from oracle import invariants as I
from oracle import oracle as O


def random_states(rng, n, a):
    q = rng.uniform(-20, 20, size=(3, n))
    q[:, np.linalg.norm(q, axis=0) < 4] *= 3.0
    p = rng.normal(size=(4, n))
    return q, p


def test_hamiltonian_matches_the_oracle_metric():
    rng = np.random.default_rng(1)
    a = float(np.float32(-0.7))           # the integrator carries the spin as a float
    integ = M.integrator(metric=M.KERR, precision=M.PREC_FP64, spin=0.7, time_reversed=True)
    q, p = random_states(rng, 20, a)
    h = I.hamiltonian(q, p, a)
    for i in range(20):
        gi = O.metric_tensor(O.F64, integ, np.r_[q[:, i], 0.0], True)
        assert h[i] == pytest.approx(0.5 * p[:, i] @ gi @ p[:, i], rel=1e-12, abs=1e-12)


def test_carter_constant_reduces_to_total_angular_momentum_at_zero_spin():
    rng = np.random.default_rng(2)
    q, p = random_states(rng, 50, 0.0)
    Lvec = np.cross(q.T, p[:3].T).T
    expect = np.sum(Lvec ** 2, axis=0) - I.angular_momentum(q, p) ** 2
    np.testing.assert_allclose(I.carter_q(q, p, 0.0), expect, rtol=1e-10, atol=1e-9)


def test_carter_constant_on_the_equator_is_p_theta_squared():
    # in the equatorial plane cos(theta) = 0, so Q = p_theta^2 = (r p_y)^2
    a = 0.9
    q = np.array([[3.0, -5.0], [0.0, 0.0], [7.0, 2.0]])
    p = np.array([[0.3, -0.1], [0.7, 0.2], [-1.0, 0.4], [1.5, -2.0]])
    r = I.kerr_r(q[0], q[1], q[2], a)
    np.testing.assert_allclose(I.carter_q(q, p, a), (r * p[1]) ** 2, rtol=1e-12)


def trajectory(a_user, y0, n_steps=400, tol=1e-11):
    a_user = float(np.float32(a_user))
    integ = M.integrator(metric=M.KERR, precision=M.PREC_FP64, spin=a_user, time_reversed=False, abs_tol=tol, rel_tol=tol)
    ys = [np.array(y0, dtype=float)]
    y, h = ys[0], 1.0
    for _ in range(n_steps):
        y, h, _, _ = O.solve_step(O.F64, integ, y, h)
        ys.append(y)
        r = I.kerr_r(y[0], y[1], y[2], a_user)
        if r < 2.2 or r > 200:
            break
    return np.array(ys).T


def null_ic(a, q, pdir):
    a = float(np.float32(a))
    integ = M.integrator(metric=M.KERR, precision=M.PREC_FP64, spin=a, time_reversed=False)
    gi = O.metric_tensor(O.F64, integ, np.r_[q, 0.0], True)
    ps = np.asarray(pdir, dtype=float)
    A, B, Cc = gi[3, 3], 2 * (gi[3, :3] @ ps), ps @ gi[:3, :3] @ ps
    pt = (-B - math.sqrt(B * B - 4 * A * Cc)) / (2 * A)
    return np.r_[q, 0.0, ps, pt]


@pytest.mark.parametrize("a", [0.0, 0.5, 0.9])
def test_invariants_conserved_along_fp64_geodesics(a):
    a = float(np.float32(a))
    for q, pd in [((0.0, 3.0, 30.0), (0.05, -0.08, -1.0)), ((4.0, -6.0, 25.0), (-0.2, 0.3, -1.0)),
                  ((0.0, 8.0, 40.0), (0.13, -0.18, -1.0))]:
        ys = trajectory(a, null_ic(a, np.array(q), pd))
        assert ys.shape[1] > 20
        qq, pp = ys[:3], ys[4:8]
        h, dl, dq = I.drift(qq[:, :1], pp[:, :1], qq, pp, a)
        assert np.abs(h).max() < 1e-8
        assert np.abs(dl).max() < 1e-8
        assert np.abs(dq).max() < 1e-7
        # and Q is not conserved by a wrong formula: dropping the y term of p_theta breaks it
        r = I.kerr_r(qq[0], qq[1], qq[2], a)
        c = qq[1] / r
        s = np.sqrt(1 - c * c)
        wrong = ((c / s) * (qq[0] * pp[0] + qq[2] * pp[2])) ** 2 + c * c * (I.angular_momentum(qq, pp) ** 2 / s ** 2)
        assert np.ptp(wrong) > 1e-3 * np.abs(wrong).max()
