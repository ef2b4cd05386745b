"""The closed forms behind the split right-hand side of the mixed-precision
integrator (magik_amd/csrc/mgk_device.hpp, rhs), restated in float64 NumPy and
checked against the oracle's restatement of the reference right-hand side
(metric_kerr.cuh:38-105, metric_schwarzschild.cuh:8-72).  The identities used:
r^4 + a^2 y^2 = r^2 sigma, f = 2 r / sigma, dr/dx = x r / sigma,
dr/dy = y (r^2 + a^2) / (r sigma), and p . grad k in five terms.  The device
code itself is exercised end to end by the GPU parity tests."""
import numpy as np
import pytest

from oracle import oracle as O


def closed_form_rhs(q, p, a, kerr=True, mass_sign=1.0, drop_u_term=False):
    x, y, z = q
    px, py, pz, pt = p
    y_sq = y * y
    if kerr:
        a_sq = a * a
        w = (x * x + y_sq + z * z) - a_sq
        sig_sq = 4 * a_sq * y_sq + w * w
        inv_sigma = 1 / np.sqrt(sig_sq)
        r_sq = 0.5 * (w + sig_sq * inv_sigma)
    else:
        a_sq = 0.0
        r_sq = x * x + y_sq + z * z
    inv_r = 1 / np.sqrt(r_sq)
    r = r_sq * inv_r
    inv_rsq = inv_r * inv_r
    if not kerr:
        inv_sigma = inv_rsq
    d = r_sq + a_sq
    inv_d = 1 / d if kerr else inv_rsq
    drx, dry, drz = x * r * inv_sigma, y * d * inv_r * inv_sigma, z * r * inv_sigma
    inv_sig_sq = inv_sigma * inv_sigma
    if kerr:
        f = 2 * r * inv_sigma
        dfs = 2 * (3 * a_sq * y_sq - r_sq * r_sq) * inv_rsq * inv_sig_sq
        dfy_extra = 4 * a_sq * y * inv_r * inv_sig_sq
    else:
        f = mass_sign * 2 * inv_r
        dfs = -2 * mass_sign * inv_rsq
        dfy_extra = 0.0
    dfx, dfy, dfz = dfs * drx, dfs * dry - dfy_extra, dfs * drz
    A = x * px + z * pz
    B = x * pz - z * px
    ky = y * inv_r
    u = (r * A + a * B) * inv_d
    kx, kz = (r * x - a * z) * inv_d, (r * z + a * x) * inv_d
    ex, ez = (r * px + a * pz) * inv_d, (r * pz - a * px) * inv_d
    kTp = (u + py * ky) - pt
    Hc = (A - (0.0 if drop_u_term else 2 * r * u)) * inv_d - py * ky * inv_r
    bx, by, bz = drx * Hc + ex, dry * Hc + py * inv_r, drz * Hc + ez
    fk, h2 = f * kTp, 0.5 * kTp * kTp
    # oracle order: dq = (x, y, z, t), dp = (x, y, z, t)
    return np.array([px - fk * kx, py - fk * ky, pz - fk * kz, -pt + fk,
                     h2 * dfx + fk * bx, h2 * dfy + fk * by, h2 * dfz + fk * bz, 0.0])


def random_states(seed, n):
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        q = rng.uniform(-15, 15, 3)
        if np.linalg.norm(q) > 2.5:
            out.append((q, rng.normal(size=4)))
    return out


@pytest.mark.parametrize("a", [0.0, 0.5, -0.9, 0.99])
def test_kerr_closed_form_matches_reference_rhs(a):
    for q, p in random_states(int(abs(a) * 100) + 1, 200):
        ref = O.rhs(O.F64, 2, a, np.r_[q, 7.0, p])
        got = closed_form_rhs(q, p, a)
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12 * np.abs(ref).max())


@pytest.mark.parametrize("mass_sign", [1.0, -1.0])
def test_schwarzschild_closed_form_matches_reference_rhs(mass_sign):
    for q, p in random_states(7, 200):
        ref = O.rhs(O.F64, 1, mass_sign, np.r_[q, 7.0, p])
        got = closed_form_rhs(q, p, 0.0, kerr=False, mass_sign=mass_sign)
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12 * np.abs(ref).max())


def test_a_dropped_term_is_detected():
    bad = 0
    for q, p in random_states(3, 50):
        ref = O.rhs(O.F64, 2, 0.9, np.r_[q, 7.0, p])
        got = closed_form_rhs(q, p, 0.9, drop_u_term=True)
        bad += np.abs(got - ref).max() > 1e-6 * np.abs(ref).max()
    assert bad >= 45
