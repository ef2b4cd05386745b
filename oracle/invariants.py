"""Constants of motion of null geodesics in Kerr-Schild Cartesian coordinates
(TEST INFRASTRUCTURE ONLY: used by tests/ and bench.py to measure constraint
drift; never by the product path).

Coordinates as in the reference (metric_kerr.cuh:8-13, :192-205): spin along
+y, covariant momentum p = (p_x, p_y, p_z, p_t), inverse metric
g^{mu nu} = eta^{mu nu} - f l^mu l^nu with
    r^2 solving (x^2 + z^2) / (r^2 + a^2) + y^2 / r^2 = 1,
    f = 2 r^3 / (r^4 + a^2 y^2),
    k = ((r x - a z) / (r^2 + a^2), y / r, (r z + a x) / (r^2 + a^2)),
so that 2 H = |p|^2 - p_t^2 - f (k.p - p_t)^2.  `a` is the spin of the
equations (already negated for time-reversed tracing).

Conserved along every geodesic: H (= 0 for light), E = -p_t, the angular
momentum about the spin axis L = z p_x - x p_z, and Carter's constant
(Carter 1968)
    Q = p_theta^2 + cos^2(theta) (L^2 / sin^2(theta) - a^2 E^2),
with the Boyer-Lindquist theta of the point (cos theta = y / r) and
p_theta = p_i dx^i/dtheta at fixed r and azimuth.  In these coordinates
x and z scale with sin(theta) and y = r cos(theta), so
    p_theta = cot(theta) (x p_x + z p_z) - r sin(theta) p_y.
At a = 0 this gives Q = |x cross p|^2 - L^2 (the total angular momentum
squared minus its spin-axis part), which the tests check.
"""
from __future__ import annotations

import numpy as np


def kerr_r(x, y, z, a):
    w = x * x + y * y + z * z - a * a
    return np.sqrt(0.5 * (w + np.sqrt(w * w + 4.0 * a * a * y * y)))


def hamiltonian(q, p, a):
    """2 H = g^{mu nu} p_mu p_nu; q = (x, y, z[, t]), p = (p_x, p_y, p_z, p_t)."""
    x, y, z = q[0], q[1], q[2]
    px, py, pz, pt = p[0], p[1], p[2], p[3]
    r = kerr_r(x, y, z, a)
    f = 2.0 * r ** 3 / (r ** 4 + a * a * y * y)
    d = r * r + a * a
    kx, ky, kz = (r * x - a * z) / d, y / r, (r * z + a * x) / d
    kp = kx * px + ky * py + kz * pz - pt
    return 0.5 * (px * px + py * py + pz * pz - pt * pt - f * kp * kp)


def angular_momentum(q, p):
    """L about the spin (y) axis: z p_x - x p_z."""
    return q[2] * p[0] - q[0] * p[2]


def carter_q(q, p, a):
    x, y, z = q[0], q[1], q[2]
    px, py, pz, pt = p[0], p[1], p[2], p[3]
    r = kerr_r(x, y, z, a)
    c = y / r
    s2 = np.maximum(1.0 - c * c, 1e-300)
    s = np.sqrt(s2)
    p_theta = (c / s) * (x * px + z * pz) - r * s * py
    L = angular_momentum(q, p)
    E = -pt
    return p_theta * p_theta + c * c * (L * L / s2 - a * a * E * E)


def drift(q0, p0, q1, p1, a):
    """Relative drift of H, L and Q between two states (arrays over rays).
    H is scaled by |p|^2; L and Q by their magnitude scale |q||p| and (|q||p|)^2."""
    scale = np.linalg.norm(np.asarray(q0[:3]), axis=0) * np.linalg.norm(np.asarray(p0[:3]), axis=0)
    h1 = hamiltonian(q1, p1, a) / np.maximum(np.sum(np.asarray(p1[:3]) ** 2, axis=0), 1e-300)
    dl = (angular_momentum(q1, p1) - angular_momentum(q0, p0)) / scale
    dq = (carter_q(q1, p1, a) - carter_q(q0, p0, a)) / (scale * scale)
    return h1, dl, dq
