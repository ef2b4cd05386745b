// CPU oracle for the wavefront spectral geodesic tracer.
//
// TEST INFRASTRUCTURE ONLY.  This header restates, on the CPU, the reference
// renderer's per-ray arithmetic so the HIP path can be checked against it.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
// the library built from it (oracle/liboracle.so).  The product path never
// calls into it.
//
// Every function is templated on the real type R:
//   R = float  : the reference integrator as written (all FP32, reference
//                operation order, so that with -ffp-contract=off it is bit-exact
//                with the reference's own internal_math.h RKF45 driver);
//   R = double : the all-FP64 oracle that the mixed-precision GPU path and the
//                surrogate are judged against.
//
// Citations are to files of the reference repository (ErikHall1801/Magik).
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>

namespace oracle {

// ---------------------------------------------------------------------------
// Constants (internal_math.h:10-16).  Kept in float, as the reference does;
// the FP64 oracle promotes the same float values.
// ---------------------------------------------------------------------------
constexpr float kPi = 3.14159265358979323846f;
constexpr float kSpeedOfLight = 299792458.0f;
constexpr float kPlanck = 6.62607015E-34f;
constexpr float kBoltzmann = 1.380649E-23f;
constexpr float kGravitational = 6.67430E-11f;

template <class R> inline R sqrt_(R v) { return std::sqrt(v); }
// The step controller's x^0.2 (internal_math.h:2193).  The reference calls the
// platform powf, whose last bit is not specified; here the FP32 instance is the
// correctly rounded power (the double result rounded once to float), which the
// GPU's FP32 reference path computes the same way, so both sides take the same
// step sizes.
template <class R> inline R pow_(R a, R b) { return std::pow(a, b); }
template <> inline float pow_<float>(float a, float b) { return float(std::pow(double(a), double(b))); }
// Azimuths of the hit records: for FP32 the correctly rounded atan2 (the double
// result rounded once), the value the GPU's FP32 reference path also records.
template <class R> inline R atan2_(R y, R x) { return std::atan2(y, x); }
template <> inline float atan2_<float>(float y, float x) { return float(std::atan2(double(y), double(x))); }
template <class R> inline R abs_(R v) { return std::fabs(v); }
template <class R> inline R max_(R a, R b) { return std::fmax(a, b); }

// ---------------------------------------------------------------------------
// Vectors.  4-vectors keep t LAST ({x,y,z,t}, internal_math.h:909-915); the
// 4-vector dot product is Euclidean, ((x*x + y*y) + z*z) + t*t
// (internal_math.h:1094-1097): a contraction only because momenta are stored
// covariant and velocities contravariant.
// ---------------------------------------------------------------------------
template <class R> struct V3 { R x{}, y{}, z{}; };
template <class R> struct V4 { R x{}, y{}, z{}, t{}; };

template <class R> inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> inline V3<R> operator-(V3<R> a) { return {-a.x, -a.y, -a.z}; }
template <class R> inline V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> inline V3<R> operator*(R s, V3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class R> inline V3<R> operator/(V3<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }
template <class R> inline R dot3(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> inline R len2_3(V3<R> a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
template <class R> inline V3<R> norm3(V3<R> a) {
    R l = sqrt_(a.x * a.x + a.y * a.y + a.z * a.z);
    return {a.x / l, a.y / l, a.z / l};
}

template <class R> inline V4<R> operator+(V4<R> a, V4<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z, a.t + b.t}; }
template <class R> inline V4<R> operator-(V4<R> a, V4<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z, a.t - b.t}; }
template <class R> inline V4<R> operator*(R s, V4<R> a) { return {s * a.x, s * a.y, s * a.z, s * a.t}; }
template <class R> inline V4<R> operator*(V4<R> a, R s) { return {a.x * s, a.y * s, a.z * s, a.t * s}; }
template <class R> inline V4<R> operator/(V4<R> a, R s) { return {a.x / s, a.y / s, a.z / s, a.t / s}; }
template <class R> inline R dot4(V4<R> a, V4<R> b) { return ((a.x * b.x) + (a.y * b.y) + (a.z * b.z)) + (a.t * b.t); }

// Row-major 3x3 (internal_math.h:1611-1700).
template <class R> struct M3 { R m[9]; };
template <class R> inline V3<R> mul(const M3<R>& a, V3<R> v) {
    return {a.m[0] * v.x + a.m[1] * v.y + a.m[2] * v.z,
            a.m[3] * v.x + a.m[4] * v.y + a.m[5] * v.z,
            a.m[6] * v.x + a.m[7] * v.y + a.m[8] * v.z};
}
template <class R> inline M3<R> mul(const M3<R>& a, const M3<R>& b) {
    M3<R> r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            r.m[3 * i + j] = a.m[3 * i + 0] * b.m[0 + j] + a.m[3 * i + 1] * b.m[3 + j] + a.m[3 * i + 2] * b.m[6 + j];
    return r;
}
// rotate(alpha, beta, gamma) = m_x(gamma) * m_y(beta) * m_z(alpha), applied to
// the identity (internal_math.h:1784-1810).
template <class R> inline M3<R> rotation(R alpha, R beta, R gamma) {
    R ca = std::cos(alpha), sa = std::sin(alpha), cb = std::cos(beta), sb = std::sin(beta);
    R cg = std::cos(gamma), sg = std::sin(gamma);
    M3<R> mz{{ca, -sa, 0, sa, ca, 0, 0, 0, 1}};
    M3<R> my{{cb, 0, sb, 0, 1, 0, -sb, 0, cb}};
    M3<R> mx{{1, 0, 0, 0, cg, -sg, 0, sg, cg}};
    return mul(mul(mx, my), mz);
}

// ---------------------------------------------------------------------------
// Geodesic state: covariant momentum p_mu, position x^mu
// (relativity_utilities.cuh:10-13) and its arithmetic (:35-58).
// ---------------------------------------------------------------------------
template <class R> struct State { V4<R> q; V4<R> p; };
template <class R> inline State<R> operator+(State<R> a, State<R> b) { return {a.q + b.q, a.p + b.p}; }
template <class R> inline State<R> operator-(State<R> a, State<R> b) { return {a.q - b.q, a.p - b.p}; }
template <class R> inline State<R> operator*(R s, State<R> a) { return {s * a.q, s * a.p}; }

// RMS of err/(abs + rel*max(|y0|,|y1|)) over the 8 components, summed in the
// order t,x,y,z of position then momentum (relativity_utilities.cuh:15-29).
template <class R>
inline R rms_error(const State<R>& e, const State<R>& y0, const State<R>& y1, R abs_tol, R rel_tol) {
    auto c = [&](R ev, R a, R b) { return ev / (abs_tol + rel_tol * max_(abs_(a), abs_(b))); };
    R s0 = c(e.q.t, y0.q.t, y1.q.t), s1 = c(e.q.x, y0.q.x, y1.q.x), s2 = c(e.q.y, y0.q.y, y1.q.y), s3 = c(e.q.z, y0.q.z, y1.q.z);
    R s4 = c(e.p.t, y0.p.t, y1.p.t), s5 = c(e.p.x, y0.p.x, y1.p.x), s6 = c(e.p.y, y0.p.y, y1.p.y), s7 = c(e.p.z, y0.p.z, y1.p.z);
    R total = (s0 * s0 + s1 * s1 + s2 * s2 + s3 * s3 + s4 * s4 + s5 * s5 + s6 * s6 + s7 * s7) / R(8);
    return sqrt_(total);
}

// ---------------------------------------------------------------------------
// Metrics.  Kerr in Cartesian Kerr-Schild coordinates with the spin axis along
// +y: g_{mu nu} = eta + f l l, l = (kx, ky, kz, 1), f = 2 r^3/(r^4 + a^2 y^2).
// ---------------------------------------------------------------------------
enum Metric : int { kMinkowski = 0, kSchwarzschild = 1, kKerr = 2 };

// r(x,y,z) solving r^4 - (rho^2 - a^2) r^2 - a^2 y^2 = 0 (metric_kerr.cuh:8-13).
template <class R> inline R kerr_schild_r(R x, R y, R z, R a) {
    R Rv = (x * x + y * y + z * z - a * a);
    R r2 = R(0.5) * Rv + R(0.5) * sqrt_(Rv * Rv + R(4) * a * a * y * y);
    return sqrt_(r2);
}

// Hamiltonian right-hand side (dx^mu/dlambda, dp_mu/dlambda), dp_t = 0
// (metric_kerr.cuh:38-105).  Operation order follows the reference.
template <class R> inline State<R> kerr_rhs(const State<R>& s, R a) {
    R x = s.q.x, y = s.q.y, z = s.q.z;
    R pt = s.p.t, px = s.p.x, py = s.p.y, pz = s.p.z;
    R a_sq = a * a;
    R y_sq = y * y;
    R w_sq = (x * x + y * y + z * z) - a_sq;
    R sigma = sqrt_(R(4) * a_sq * y_sq + w_sq * w_sq);
    R r = sqrt_(R(0.5) * (w_sq + sigma));
    R r_sq = r * r, r_cu = r_sq * r, r_quad = r_sq * r_sq, r_six = r_cu * r_cu;
    R inv_r = R(1) / r;
    R inv_rsq = inv_r * inv_r;
    R inv_sigma = R(1) / sigma;
    R scale_xz = R(0.5) * inv_r * (R(1) + w_sq * inv_sigma);
    R scale_y = R(0.5) * inv_r * (R(1) + (w_sq + R(2) * a_sq) * inv_sigma);
    R drx = x * scale_xz, dry = y * scale_y, drz = z * scale_xz;
    R d = r_sq + a_sq;
    R inv_d = R(1) / d;
    R inv_dsq = inv_d * inv_d;
    R kx_num = r * x - a * z;
    R kz_num = r * z + a * x;
    R kx = kx_num * inv_d, kz = kz_num * inv_d, ky = y * inv_r;
    R f_den = r_quad + a_sq * y_sq;
    R inv_f_den_sq = R(1) / (f_den * f_den);
    R f = R(2) * r_cu / f_den;
    R dfdq_scale = R(2) * (R(3) * a_sq * r_sq * y_sq - r_six) * inv_f_den_sq;
    R dfx = dfdq_scale * drx;
    R dfy = dfdq_scale * dry - R(4) * a_sq * r_cu * y * inv_f_den_sq;
    R dfz = dfdq_scale * drz;
    R dkx_dx = ((x * drx + r) * d - R(2) * r * drx * kx_num) * inv_dsq;
    R dkx_dy = ((x * dry) * d - R(2) * r * dry * kx_num) * inv_dsq;
    R dkx_dz = ((x * drz - a) * d - R(2) * r * drz * kx_num) * inv_dsq;
    R dky_dx = -y * drx * inv_rsq;
    R dky_dy = (r - y * dry) * inv_rsq;
    R dky_dz = -y * drz * inv_rsq;
    R dkz_dx = ((a + z * drx) * d - R(2) * r * drx * kz_num) * inv_dsq;
    R dkz_dy = ((z * dry) * d - R(2) * r * dry * kz_num) * inv_dsq;
    R dkz_dz = ((r + z * drz) * d - R(2) * r * drz * kz_num) * inv_dsq;
    R kTp = px * kx + py * ky + pz * kz - pt;
    R f_kTp = f * kTp;
    R half_kTp_sq = R(0.5) * kTp * kTp;
    R bx = px * dkx_dx + py * dky_dx + pz * dkz_dx;
    R by = px * dkx_dy + py * dky_dy + pz * dkz_dy;
    R bz = px * dkx_dz + py * dky_dz + pz * dkz_dz;
    State<R> out;
    out.q = {px - f_kTp * kx, py - f_kTp * ky, pz - f_kTp * kz, -pt + f_kTp};
    out.p = {half_kTp_sq * dfx + f_kTp * bx, half_kTp_sq * dfy + f_kTp * by, half_kTp_sq * dfz + f_kTp * bz, R(0)};
    return out;
}

// Schwarzschild: the a = 0 Kerr expression with f and df scaled by mass_sign
// (metric_schwarzschild.cuh:8-72).
template <class R> inline State<R> schwarzschild_rhs(const State<R>& s, R mass_sign) {
    R x = s.q.x, y = s.q.y, z = s.q.z;
    R pt = s.p.t, px = s.p.x, py = s.p.y, pz = s.p.z;
    R w_sq = (x * x + y * y + z * z);
    R sigma = w_sq;
    R r = sqrt_(R(0.5) * (w_sq + sigma));
    R r_sq = r * r, r_cu = r_sq * r, r_quad = r_sq * r_sq, r_six = r_cu * r_cu;
    R inv_r = R(1) / r;
    R inv_rsq = inv_r * inv_r;
    R inv_sigma = R(1) / sigma;
    R scale_xz = R(0.5) * inv_r * (R(1) + w_sq * inv_sigma);
    R scale_y = R(0.5) * inv_r * (R(1) + w_sq * inv_sigma);
    R drx = x * scale_xz, dry = y * scale_y, drz = z * scale_xz;
    R d = r_sq;
    R inv_d = R(1) / d;
    R inv_dsq = inv_d * inv_d;
    R kx_num = r * x, kz_num = r * z;
    R kx = kx_num * inv_d, kz = kz_num * inv_d, ky = y * inv_r;
    R f_den = r_quad;
    R inv_f_den_sq = R(1) / (f_den * f_den);
    R f = mass_sign * R(2) * r_cu / f_den;
    R dfdq_scale = mass_sign * R(2) * -r_six * inv_f_den_sq;
    R dfx = dfdq_scale * drx, dfy = dfdq_scale * dry, dfz = dfdq_scale * drz;
    R dkx_dx = ((x * drx + r) * d - R(2) * r * drx * kx_num) * inv_dsq;
    R dkx_dy = ((x * dry) * d - R(2) * r * dry * kx_num) * inv_dsq;
    R dkx_dz = ((x * drz) * d - R(2) * r * drz * kx_num) * inv_dsq;
    R dky_dx = -y * drx * inv_rsq;
    R dky_dy = (r - y * dry) * inv_rsq;
    R dky_dz = -y * drz * inv_rsq;
    R dkz_dx = ((z * drx) * d - R(2) * r * drx * kz_num) * inv_dsq;
    R dkz_dy = ((z * dry) * d - R(2) * r * dry * kz_num) * inv_dsq;
    R dkz_dz = ((r + z * drz) * d - R(2) * r * drz * kz_num) * inv_dsq;
    R kTp = px * kx + py * ky + pz * kz - pt;
    R f_kTp = f * kTp;
    R half_kTp_sq = R(0.5) * kTp * kTp;
    R bx = px * dkx_dx + py * dky_dx + pz * dkz_dx;
    R by = px * dkx_dy + py * dky_dy + pz * dkz_dy;
    R bz = px * dkx_dz + py * dky_dz + pz * dkz_dz;
    State<R> out;
    out.q = {px - f_kTp * kx, py - f_kTp * ky, pz - f_kTp * kz, -pt + f_kTp};
    out.p = {half_kTp_sq * dfx + f_kTp * bx, half_kTp_sq * dfy + f_kTp * by, half_kTp_sq * dfz + f_kTp * bz, R(0)};
    return out;
}

// Minkowski: straight line, dx = (p_x, p_y, p_z, -p_t) (metric_minkowski.cuh:8-13).
template <class R> inline State<R> minkowski_rhs(const State<R>& s) {
    State<R> out;
    out.q = {s.p.x, s.p.y, s.p.z, -s.p.t};
    out.p = {R(0), R(0), R(0), R(0)};
    return out;
}

// Per-metric parameters (make_params, metric_kerr.cuh:125-147,
// metric_schwarzschild.cuh:92-114, metric_minkowski.cuh:30-42).  The spin is
// stored already negated when time_reversed is set (metric_kerr.cuh:129).
template <class R> struct MetricParams {
    int metric = kKerr;
    bool time_reversed = true;
    R mass_kg = R(6.7329546075E26f);
    R spin = R(0);              // signed spin used by the equations (already reversed)
    R mass_sign = R(1);         // Schwarzschild only
    R horizon_scalar = R(1.01f);
    R h_max = R(1000);
    R abs_tol = R(1e-3f), rel_tol = R(1e-10f);
    uint32_t n_max = 56;
};

template <class R>
inline MetricParams<R> make_params(int metric, bool time_reversed, float mass_kg, float spin, float horizon_scalar,
                                   float abs_tol, float rel_tol, uint32_t n_max) {
    MetricParams<R> p;
    p.metric = metric;
    p.time_reversed = time_reversed;
    if (metric == kSchwarzschild) {
        p.mass_kg = R(std::fabs(mass_kg));
        p.mass_sign = mass_kg > 0.0f ? R(1) : R(-1);
    } else {
        p.mass_kg = R(mass_kg);
    }
    p.spin = R(time_reversed ? -spin : spin);
    p.horizon_scalar = R(horizon_scalar);
    p.abs_tol = R(abs_tol);
    p.rel_tol = R(rel_tol);
    p.n_max = n_max;
    return p;
}

// length_scale = G M / c^2 in float, as the reference computes it
// (metric_kerr.cuh:154-157); 1 for Minkowski (metric_minkowski.cuh:44-47).
template <class R> inline R length_scale(const MetricParams<R>& p) {
    if (p.metric == kMinkowski) return R(1);
    // evaluated in float for every state type: the FP64 paths promote the float value
    const float m = float(p.mass_kg);
    return R(float(kGravitational) * m / (float(kSpeedOfLight) * float(kSpeedOfLight)));
}
template <class R> inline R initial_h(const MetricParams<R>& p) { return p.metric == kMinkowski ? R(1000) : R(1); }

template <class R> inline State<R> rhs(const State<R>& s, const MetricParams<R>& p) {
    if (p.metric == kKerr) return kerr_rhs(s, p.spin);
    if (p.metric == kSchwarzschild) return schwarzschild_rhs(s, p.mass_sign);
    return minkowski_rhs(s);
}

// in_domain: inside the (scaled) horizon.  Kerr metric_kerr.cuh:169-175,
// Schwarzschild metric_schwarzschild.cuh:136-141, Minkowski never.
template <class R> inline bool in_domain(V4<R> q, const MetricParams<R>& p) {
    if (p.metric == kKerr) {
        if (std::fabs(p.spin) > R(1)) return false;
        R r = kerr_schild_r(q.x, q.y, q.z, p.spin);
        if (((R(1) + sqrt_(R(1) - p.spin * p.spin)) * p.horizon_scalar) < r) return false;
        return true;
    }
    if (p.metric == kSchwarzschild) {
        R r = sqrt_(q.x * q.x + q.y * q.y + q.z * q.z);
        if (p.mass_sign < R(0)) return r < R(1e-3f);
        return r <= R(2) * p.horizon_scalar;
    }
    return false;
}

// Symmetric 4x4 tensor (relativity_utilities.cuh:83-89).
template <class R> struct Sym4 { R tt, tx, ty, tz, xx, xy, xz, yy, yz, zz; };

template <class R> inline Sym4<R> compute_g(V4<R> q, const MetricParams<R>& p, bool inverse) {
    if (p.metric == kMinkowski) return {R(-1), 0, 0, 0, R(1), 0, 0, R(1), 0, R(1)};
    R f, kx, ky, kz;
    if (p.metric == kKerr) {
        R a = p.spin;
        R r = kerr_schild_r(q.x, q.y, q.z, a);
        f = R(2) * r * r * r / (r * r * r * r + a * a * q.y * q.y);           // metric_kerr.cuh:15-18
        kx = (r * q.x - a * q.z) / (r * r + a * a);                           // :20-23
        ky = q.y / r;                                                         // :25-28
        kz = (r * q.z + a * q.x) / (r * r + a * a);                           // :30-33
    } else {
        R r = sqrt_(q.x * q.x + q.y * q.y + q.z * q.z);
        f = R(2) * p.mass_sign / r;                                           // metric_schwarzschild.cuh:143-153
        kx = q.x / r; ky = q.y / r; kz = q.z / r;
    }
    if (!inverse)                                                             // metric_kerr.cuh:177-190
        return {R(-1) + f, f * kx, f * ky, f * kz, R(1) + f * kx * kx, f * kx * ky, f * kx * kz,
                R(1) + f * ky * ky, f * ky * kz, R(1) + f * kz * kz};
    return {R(-1) - f, f * kx, f * ky, f * kz, R(1) - f * kx * kx, -f * kx * ky, -f * kx * kz,   // :192-205
            R(1) - f * ky * ky, -f * ky * kz, R(1) - f * kz * kz};
}

// Reference observer (metric_kerr.cuh:207-226: BL ZAMO-like u^t, u^phi about
// +y; metric_schwarzschild.cuh:167-171: static; Minkowski (0,0,0,1)).
template <class R> inline V4<R> reference_observer(V4<R> q, const MetricParams<R>& p) {
    if (p.metric == kKerr) {
        R a = p.spin;
        R r = kerr_schild_r(q.x, q.y, q.z, a);
        R r_sq = r * r, a_sq = a * a;
        R cos_theta = q.y / r;
        R sin2_theta = R(1) - cos_theta * cos_theta;
        R sigma = r_sq + a_sq * cos_theta * cos_theta;
        R delta = r_sq - R(2) * r + a_sq;
        R Lambda = (r_sq + a_sq) * (r_sq + a_sq) - a_sq * delta * sin2_theta;
        R u_t = sqrt_(Lambda / (delta * sigma));
        R u_phi = R(2) * a * r / sqrt_(Lambda * delta * sigma);
        return {q.z * u_phi, R(0), -q.x * u_phi, u_t};
    }
    if (p.metric == kSchwarzschild) {
        Sym4<R> g = compute_g(q, p, false);
        return {R(0), R(0), R(0), R(1) / sqrt_(-g.tt)};
    }
    return {R(0), R(0), R(0), R(1)};
}

// g(u, v) (relativity_utilities.cuh:122-130) and index lowering/raising (:132-139).
template <class R> inline R inner(const Sym4<R>& g, V4<R> u, V4<R> v) {
    return g.tt * u.t * v.t + g.tx * (u.t * v.x + u.x * v.t) + g.ty * (u.t * v.y + u.y * v.t) +
           g.tz * (u.t * v.z + u.z * v.t) + g.xx * u.x * v.x + g.xy * (u.x * v.y + u.y * v.x) +
           g.xz * (u.x * v.z + u.z * v.x) + g.yy * u.y * v.y + g.yz * (u.y * v.z + u.z * v.y) + g.zz * u.z * v.z;
}
template <class R> inline V4<R> contract(const Sym4<R>& g, V4<R> u) {
    return {g.tx * u.t + g.xx * u.x + g.xy * u.y + g.xz * u.z, g.ty * u.t + g.xy * u.x + g.yy * u.y + g.yz * u.z,
            g.tz * u.t + g.xz * u.x + g.yz * u.y + g.zz * u.z, g.tt * u.t + g.tx * u.x + g.ty * u.y + g.tz * u.z};
}

template <class R> struct Seeds { V4<R> X{R(1), 0, 0, 0}, Y{0, R(1), 0, 0}, Z{0, 0, R(1), 0}; };
template <class R> struct Tetrad { V4<R> t, x, y, z; };

// Gram-Schmidt with e_t = u, order z, x, y (relativity_utilities.cuh:141-156).
template <class R> inline Tetrad<R> build_tetrad(const Sym4<R>& g, const Seeds<R>& s, V4<R> u) {
    Tetrad<R> e;
    e.t = u;
    e.z = s.Z + inner(g, s.Z, u) * u;
    e.z = e.z / sqrt_(inner(g, e.z, e.z));
    e.x = s.X + inner(g, s.X, u) * u - inner(g, s.X, e.z) * e.z;
    e.x = e.x / sqrt_(inner(g, e.x, e.x));
    e.y = s.Y + inner(g, s.Y, u) * u - inner(g, s.Y, e.z) * e.z - inner(g, s.Y, e.x) * e.x;
    e.y = e.y / sqrt_(inner(g, e.y, e.y));
    return e;
}
template <class R> inline V4<R> local_to_global(const Tetrad<R>& e, V4<R> l) {   // :158-161
    return e.t * l.t + e.x * l.x + e.y * l.y + e.z * l.z;
}

// Four-velocity of an observer moving with proper velocity w (units of c)
// relative to the reference observer; w is negated for time-reversed tracing
// (relativity_utilities.cuh:163-173).
template <class R> inline V4<R> rest_frame_four_velocity(V4<R> q, const Seeds<R>& seeds, V3<R> w, const MetricParams<R>& p) {
    Sym4<R> g = compute_g(q, p, false);
    V4<R> u_ref = reference_observer(q, p);
    Tetrad<R> e = build_tetrad(g, seeds, u_ref);
    V3<R> wv = p.time_reversed ? -w : w;
    V4<R> local{wv.x, wv.y, wv.z, sqrt_(len2_3(wv) - R(-1))};
    return local_to_global(e, local);
}

// Projection of a 4-vector onto an observer's frame (relativity_utilities.cuh:175-184).
template <class R> inline V4<R> global_to_observer_rest_frame(V4<R> q, const Seeds<R>& seeds, V4<R> v, V4<R> u, const MetricParams<R>& p) {
    Sym4<R> g = compute_g(q, p, false);
    Tetrad<R> e = build_tetrad(g, seeds, u);
    return {inner(g, e.x, v), inner(g, e.y, v), inner(g, e.z, v), -inner(g, e.t, v)};
}

// Local spatial direction -> covariant global momentum (relativity_utilities.cuh:186-195).
template <class R>
inline V4<R> direction_to_momentum(V4<R> q, const Seeds<R>& seeds, V3<R> d, V4<R> u, R squared_norm, const MetricParams<R>& p) {
    Sym4<R> g = compute_g(q, p, false);
    Tetrad<R> e = build_tetrad(g, seeds, u);
    V4<R> local{d.x, d.y, d.z, sqrt_(len2_3(d) - squared_norm)};
    V4<R> p_up = local_to_global(e, local);
    return contract(g, p_up);
}

// Redshift factor with NaN/Inf mapped to 1e-10 (relativity_utilities.cuh:62-67).
template <class R> inline R redshift_factor(R initial_contraction, V4<R> u, V4<R> p) {
    R z = initial_contraction / dot4(u, p);
    if (std::isnan(z) || std::isinf(z)) z = R(1e-10f);
    return z;
}

// ---------------------------------------------------------------------------
// RKF45 with the reference's controller (internal_math.h:2160-2205): advance
// with the 5th-order weights, error = 5th - 4th, h_new = 0.9 h err^-0.2 capped
// at h_max, accept when err < 1, and after n_max rejections return the last y1.
// h_taken is the last tried h; *h becomes the next h.
// Stage sums are evaluated left to right exactly as written there.
// ---------------------------------------------------------------------------
// One trial: the 5th-order solution over th and its error estimate (stage sums
// left to right as internal_math.h:2167-2189 writes them).
template <class R, class F> inline State<R> rkf45_trial(F f, const State<R>& y0, R th, State<R>* err, State<R>* k5_out = nullptr) {
    State<R> k1, k2, k3, k4, k5, k6;
    k1 = f(y0);
    k2 = f(y0 + (R(1) / R(4) * th) * k1);
    k3 = f(y0 + (R(3) / R(32) * th) * k1 + (R(9) / R(32) * th) * k2);
    k4 = f(y0 + (R(1932) / R(2197) * th) * k1 - (R(7200) / R(2197) * th) * k2 + (R(7296) / R(2197) * th) * k3);
    k5 = f(y0 + (R(439) / R(216) * th) * k1 - (R(8) * th) * k2 + (R(3680) / R(513) * th) * k3 - (R(845) / R(4104) * th) * k4);
    k6 = f(y0 - (R(8) / R(27) * th) * k1 + (R(2) * th) * k2 - (R(3544) / R(2565) * th) * k3 + (R(1859) / R(4104) * th) * k4 -
           (R(11) / R(40) * th) * k5);
    *err = (R(-1) / R(360) * th) * k1 + (R(128) / R(4275) * th) * k3 + (R(2197) / R(75240) * th) * k4 - (R(1) / R(50) * th) * k5 -
           (R(2) / R(55) * th) * k6;
    if (k5_out) *k5_out = k5;
    return y0 + (R(16) / R(135) * th) * k1 + (R(6656) / R(12825) * th) * k3 + (R(28561) / R(56430) * th) * k4 -
           (R(9) / R(50) * th) * k5 + (R(2) / R(55) * th) * k6;
}

template <class R, class F>
inline State<R> rkf45_step(F f, const State<R>& y0, R* h, R* h_taken, R h_max, uint32_t n_max, R abs_tol, R rel_tol,
                           uint32_t* trials = nullptr) {
    State<R> y1{}, err;
    R h_new = 0, th = *h, rms = 0, rms_safe = 0;
    uint32_t n = 0;
    for (uint32_t i = 0; i < n_max; i++) {
        ++n;
        y1 = rkf45_trial(f, y0, th, &err);
        rms = rms_error(err, y0, y1, abs_tol, rel_tol);
        rms_safe = rms > R(1e-30f) ? rms : R(1e-30f);
        h_new = R(0.9f) * th * pow_(R(1) / rms_safe, R(0.2f));
        if (h_new > h_max) h_new = h_max;
        *h_taken = th;
        th = h_new;
        if (rms_safe < R(1)) break;
    }
    *h = th;
    if (trials) *trials = n;
    return y1;
}

template <class R>
inline State<R> solve_step(const State<R>& y0, const MetricParams<R>& p, R* h, R* h_taken, uint32_t* trials = nullptr) {
    if (p.metric == kMinkowski) {   // explicit Euler with the stored h (metric_minkowski.cuh:49-53)
        *h_taken = *h;
        if (trials) *trials = 1;
        return y0 + *h * minkowski_rhs(y0);
    }
    auto f = [&](const State<R>& s) { return rhs(s, p); };
    return rkf45_step(f, y0, h, h_taken, p.h_max, p.n_max, p.abs_tol, p.rel_tol, trials);
}

// ---------------------------------------------------------------------------
// Camera (pinhole "thin lens"): sensor point with jitter, direction toward the
// lens point (0,0,-f) (device_sensor_utilities.cuh:35-80, thin_lens.cuh:60-76).
// ju, jv are the two jitter uniforms in [0,1).
// ---------------------------------------------------------------------------
template <class R>
inline V3<R> sensor_position(R sensor_width_mm, uint32_t W, uint32_t H, uint32_t px, uint32_t py, R ju_u, R jv_u) {
    R inv_s = sensor_width_mm / sqrt_(R(float(W * W + H * H)));
    R width_mm = R(float(W)) * inv_s, height_mm = R(float(H)) * inv_s;
    R num_px = R(float(W * H));
    R area = (width_mm * height_mm) / R(1E6f);
    R side = sqrt_(area / num_px);
    R ju = (ju_u - R(0.5)) * side, jv = (jv_u - R(0.5)) * side;
    R u = -((R(float(px)) / R(float(W))) - R(0.5));
    R v = -(R(float(py)) / R(float(H)) - R(0.5));
    R ratio = R(float(H)) / R(float(W));
    if (ratio < R(1)) v *= ratio; else u /= ratio;
    return {((-u * sensor_width_mm) / R(1000)) + ju, ((-v * sensor_width_mm) / R(1000)) + jv, R(0)};
}

template <class R> inline void pinhole_ray(V3<R> sensor, R focal_mm, V3<R>* dir, V3<R>* orig) {
    R f_m = focal_mm / R(1000);
    V3<R> lens{R(0), R(0), -f_m};
    *orig = lens;
    *dir = norm3(lens - sensor);
}

// Initial conditions of the render path (kernel_generate_initial_conditions.cu:54-74):
// x = (pos/L, t = 1/(L/c)), u from the reference observer with the observer's
// proper velocity (identity seeds), p_mu from the camera-local direction with
// the camera's seed vectors, contraction u.p, h0.
template <class R> struct Camera {
    V3<R> position;            // SI metres
    M3<R> orientation;         // camera -> world (row-major)
    R focal_mm = R(12), sensor_width_mm = R(24);
    uint32_t W = 0, H = 0;
    V3<R> observer_velocity{};  // m/s
};

template <class R> struct RayIC { State<R> y; R contraction; R h; V4<R> u_obs; };

template <class R>
inline RayIC<R> initial_conditions(const Camera<R>& cam, const MetricParams<R>& p, V3<R> local_dir, V3<R> world_orig) {
    R L = length_scale(p);
    V4<R> q{world_orig.x / L, world_orig.y / L, world_orig.z / L, R(1) / (L / R(kSpeedOfLight))};
    Seeds<R> ident;
    V4<R> u = rest_frame_four_velocity(q, ident, cam.observer_velocity / R(kSpeedOfLight), p);
    Seeds<R> cs;
    V3<R> bx = mul(cam.orientation, V3<R>{R(1), 0, 0}), by = mul(cam.orientation, V3<R>{0, R(1), 0}),
          bz = mul(cam.orientation, V3<R>{0, 0, R(1)});
    cs.X = {bx.x, bx.y, bx.z, R(0)};
    cs.Y = {by.x, by.y, by.z, R(0)};
    cs.Z = {bz.x, bz.y, bz.z, R(0)};
    V4<R> pm = direction_to_momentum(q, cs, local_dir / L, u, R(0), p);
    RayIC<R> ic;
    ic.y = {q, pm};
    ic.contraction = dot4(u, pm);
    ic.h = initial_h(p);
    ic.u_obs = u;
    return ic;
}

// ---------------------------------------------------------------------------
// Scene traceables.  The AABB slab and celestial sphere restate
// traceables.cuh:17-86; the thin disk, emissive sphere and Keplerian emitter
// are this build's additions for the benchmark scenes.
// ---------------------------------------------------------------------------
enum HitClass : uint32_t {
    HIT_NONE = 0, HIT_HORIZON = 1, HIT_ESCAPE = 2, HIT_DISK = 3, HIT_SPHERE = 4, HIT_STEP_LIMIT = 5,
    HIT_BOX_DIFFUSE = 6, HIT_BOX_EMITTER = 7, HIT_MESH_DIFFUSE = 8, HIT_MESH_EMITTER = 9
};
enum Matter : uint32_t { MATTER_DIFFUSE = 1, MATTER_ILLUMINANT = 2 };   // kernel_bxdf.cu:126 slots 1 and 2

// Triangle meshes placed by affine instances.  The reference reserves mesh
// intersection (wavefront.cpp:94-100) without implementing it; this is the
// restatement of the library's design: two-sided triangles (Moller &
// Trumbore, JGT 1997) against the step's chord taken into each instance's
// object space.  No hierarchy: every instance and triangle is tested, which
// is what the library's top and bottom levels must reproduce.  The world ->
// object maps, normal maps and triangles are held as the library stores them
// (float), then widened to R.
template <class R> struct MeshSet {
    struct Tri { V3<R> v0, e1, e2, n; };   // object space; n the unit geometric normal
    struct Inst {
        uint32_t mesh, matter;
        R w2o[12];                         // world -> object, row-major 3 x 4
        R nm[9];                           // object -> world normals: the transposed linear part of w2o
        V3<R> lo, hi;                      // object-space bounds of the mesh (a cull that drops no hit)
    };
    std::vector<std::vector<Tri>> mesh;
    std::vector<Inst> inst;
};

template <class R> struct Scene {
    R celestial_radius_m = R(1000);   // measured from the camera origin, SI metres
    bool disk = false;
    bool disk_on_geodesic = false;    // locate the disk crossing on the geodesic, not on the step's chord
    R disk_r_in = R(6), disk_r_out = R(20), disk_T_in = R(1e4);
    bool sphere = false;
    V3<R> sphere_center{};            // dimensionless coordinates
    R sphere_radius = R(1), sphere_T = R(6500);
    uint32_t sphere_checker = 16;
    R background_scale = R(0), background_T = R(6500);
    uint32_t step_limit = 512;
    // axis-aligned boxes in SI metres (stage.h:19-22), tested in order
    struct Box { V3<R> mn, mx; uint32_t matter; };
    uint32_t n_box = 0;
    Box box[4];
    uint32_t max_reflections = 10;   // render_settings.h:17
    R led_scale = R(7e5f);           // matter_examples.cuh:70
    const MeshSet<R>* meshes = nullptr;
    bool hermite = false;            // Hermite chords for the meshes (else the straight chord)
};

template <class R> struct Hit {
    uint32_t cls = HIT_NONE;
    R g = R(0), a0 = R(0), a1 = R(0);
    uint32_t steps = 0, trials = 0;
    State<R> at{};   // box hits: the chord point of the hit and the step's end momentum
};

// Slab test of the chord against an AABB, hit parameter in chord units
// (traceables.cuh:17-74).  The normal (an axis, against the chord direction)
// comes back as a code: axis 0-2, plus 3 when it points along the negative axis.
template <class R>
inline bool aabb_hit(R* t_out, V3<R> dir, V3<R> pos, R inv_chord, V3<R> bmin, V3<R> bmax, int* normal_code = nullptr) {
    V3<R> b[2] = {bmin, bmax};
    V3<R> inv{R(1) / dir.x, R(1) / dir.y, R(1) / dir.z};
    int sx = inv.x < 0, sy = inv.y < 0, sz = inv.z < 0;
    int axis_min = 0, axis_max = 0;
    R tmin = (b[sx].x - pos.x) * inv.x, tmax = (b[1 - sx].x - pos.x) * inv.x;
    R tymin = (b[sy].y - pos.y) * inv.y, tymax = (b[1 - sy].y - pos.y) * inv.y;
    if ((tmin > tymax) || (tymin > tmax)) return false;
    if (tymin > tmin) { tmin = tymin; axis_min = 1; }
    if (tymax < tmax) { tmax = tymax; axis_max = 1; }
    R tzmin = (b[sz].z - pos.z) * inv.z, tzmax = (b[1 - sz].z - pos.z) * inv.z;
    if ((tmin > tzmax) || (tzmin > tmax)) return false;
    if (tzmin > tmin) { tmin = tzmin; axis_min = 2; }
    if (tzmax < tmax) { tmax = tzmax; axis_max = 2; }
    if (tmax < R(0)) return false;
    bool inside = tmin < R(0);
    R t = (inside ? tmax : tmin) * inv_chord;
    if (t > R(1)) return false;
    *t_out = t;
    if (normal_code) {
        const int axis = inside ? axis_max : axis_min;
        const int sgn[3] = {sx, sy, sz};
        R ns = sgn[axis] ? R(1) : R(-1);
        if (inside) ns = -ns;
        const R d = axis == 0 ? dir.x : (axis == 1 ? dir.y : dir.z);
        if (ns * d > R(0)) ns = -ns;   // against the chord direction
        *normal_code = axis + (ns < R(0) ? 3 : 0);
    }
    return true;
}

template <class R> inline V3<R> normal_of(int code) {
    const R s = code >= 3 ? R(-1) : R(1);
    const int axis = code % 3;
    return {axis == 0 ? s : R(0), axis == 1 ? s : R(0), axis == 2 ? s : R(0)};
}

// White LED spectrum: five Gaussians exp(-(x - mu)^2 / (2 w^2)) * scale
// (internal_math.h:2441-2445, :2501-2509).
template <class R> inline R white_led(R lambda_nm) {
    auto gd = [](R x, R mu, R w, R k) { return std::exp(R(-0.5) * (((x - mu) * (x - mu)) / (w * w))) * k; };
    R f0 = gd(lambda_nm, R(577.0), R(50.7937003968), R(157.0));
    R f1 = gd(lambda_nm, R(451.0), R(7.07106781187), R(200.0));
    R f2 = gd(lambda_nm, R(455.0), R(14.1421356237), R(50.0));
    R f3 = gd(lambda_nm, R(528.0), R(13.7840487521), R(20.0));
    R f4 = gd(lambda_nm, R(690.0), R(44.72135955), R(10.0));
    return f0 + f1 + f2 + f3 + f4;
}

// Self-intersection offset along the normal n (Wächter & Binder, "A Fast and
// Robust Method for Avoiding Self-Intersection", Ray Tracing Gems ch. 6, as
// used by bxdf_utilities.cuh:112-134): a move of int(1024 n) units in the last
// place of each float coordinate, or 2^-14 n near the origin.  Evaluated on the
// float coordinates; in double the same float displacement is added.
inline float offset_coord(float x, float n) {
    const float int_scale = 256.0f * 4.0f, float_scale = (1.0f / 65536.0f) * 4.0f, origin = 1.0f / 32.0f;
    const int of = int(int_scale * n);
    int32_t xi;
    std::memcpy(&xi, &x, 4);
    xi += (x < 0) ? -of : of;
    float p;
    std::memcpy(&p, &xi, 4);
    return std::fabs(x) < origin ? x + float_scale * n : p;
}
template <class R> inline V4<R> offset_geodesic(V4<R> q, V3<R> n) {
    const float fx = float(q.x), fy = float(q.y), fz = float(q.z);
    return {q.x + R(offset_coord(fx, float(n.x)) - fx), q.y + R(offset_coord(fy, float(n.y)) - fy),
            q.z + R(offset_coord(fz, float(n.z)) - fz), q.t};
}

// Reflection off diffuse box matter at rest (kernel_bxdf.cu:69-95 with
// bxdf_utilities.cuh:56-110, :136-151 and matter_examples.cuh:36-60,
// brdf.cuh:12-29).  `at` holds the hit point and the step's end momentum
// (covariant).  The rest-frame incoming direction and geometric normal come
// from the reference observer's tetrad; faces with an x normal reflect
// specularly, the others send (p + n) / |p + n| for p uniform on the sphere
// (cosine-weighted about the facing normal n).  Returns the new state (offset
// position, outgoing covariant momentum) and the new u.p contraction, which
// carries the redshift accumulated so far.  reflect_matter takes the global
// geometric normal ng and the specular choice; mesh matter is Lambertian.
template <class R>
inline void reflect_matter(const MetricParams<R>& p, const State<R>& at, R contraction, V3<R> ng, bool specular, R u0, R u1,
                           State<R>* out, R* contraction_out) {
    const Seeds<R> seeds;
    const V4<R> q = at.q;
    const V4<R> u = rest_frame_four_velocity(q, seeds, V3<R>{R(0), R(0), R(0)}, p);
    const Sym4<R> ginv = compute_g(q, p, true);
    const V4<R> prf = global_to_observer_rest_frame(q, seeds, contract(ginv, at.p), u, p);
    const V3<R> din = norm3(V3<R>{prf.x, prf.y, prf.z});
    const R z = redshift_factor(contraction, u, at.p);
    const V4<R> nrf = global_to_observer_rest_frame(q, seeds, contract(ginv, V4<R>{ng.x, ng.y, ng.z, R(0)}), u, p);
    V3<R> nr = norm3(V3<R>{nrf.x, nrf.y, nrf.z});
    if (dot3(nr, din) > R(0)) nr = -nr;
    V3<R> dout;
    if (specular) {
        dout = norm3(din - (R(2) * nr * dot3(din, nr)));
    } else {
        const R th = std::acos(R(1) - R(2) * u0), ph = R(2) * R(kPi) * u1;
        const V3<R> s{std::cos(ph) * std::sin(th), std::sin(ph) * std::sin(th), std::cos(th)};
        dout = norm3(s + nr);
    }
    const V4<R> pout = direction_to_momentum(q, seeds, dout, u, R(0), p);
    out->q = offset_geodesic(q, ng);
    out->p = pout;
    *contraction_out = dot4(u, pout) * z;
}

// Box matter: the face normal from its code, specular on faces with an x normal.
template <class R>
inline void reflect_diffuse(const MetricParams<R>& p, const State<R>& at, R contraction, int normal_code, R u0, R u1, State<R>* out,
                            R* contraction_out) {
    const V3<R> ng = normal_of<R>(normal_code);
    reflect_matter(p, at, contraction, ng, ng.x * ng.x > R(0), u0, u1, out, contraction_out);
}

// Thin equatorial disk y = 0, r_in <= r <= r_out: chord crossing of the plane.
template <class R> inline bool disk_hit(R* t_out, R* r_out, V4<R> q0, V4<R> q1, R a, R r_in, R r_max) {
    R y0 = q0.y, y1 = q1.y;
    if (!((y0 > R(0) && y1 <= R(0)) || (y0 < R(0) && y1 >= R(0)))) return false;
    R t = y0 / (y0 - y1);
    R x = q0.x + t * (q1.x - q0.x), z = q0.z + t * (q1.z - q0.z);
    R rho2 = x * x + z * z - a * a;
    if (rho2 <= R(0)) return false;
    R r = sqrt_(rho2);
    if (r < r_in || r > r_max) return false;
    *t_out = t;
    *r_out = r;
    return true;
}

// The disk crossing on the geodesic instead of on the step's chord (a scene
// option; the chord's crossing is off by O(h^2) of the step's curvature).  One
// RKF45 trial -- the 5th-order solution, no step control -- from the step's
// start over the chord's crossing parameter, t h, then one Newton step to the
// plane y = 0 along the tangent dx/dlambda of the trial's fifth stage, which
// Fehlberg's tableau evaluates at the trial's end (c5 = 1).  The remaining
// offset is second order in the first estimate's error.  False when the
// corrected point lies off the disk (the path then goes on).
template <class R, class F>
inline bool disk_on_geodesic(F f, const State<R>& y0, R t, R h, R a, R r_in, R r_max, R* r_out, V3<R>* x_out) {
    State<R> err, d;
    const State<R> yc = rkf45_trial(f, y0, t * h, &err, &d);
    const R dl = -yc.q.y / d.q.y;
    const R x = yc.q.x + dl * d.q.x, z = yc.q.z + dl * d.q.z;
    const R rho2 = x * x + z * z - a * a;
    if (!(rho2 > R(0))) return false;
    const R r = sqrt_(rho2);
    if (r < r_in || r > r_max) return false;
    *r_out = r;
    *x_out = V3<R>{x, R(0), z};
    return true;
}

// Segment-sphere: smallest root in [0,1] of |q0 + t (q1-q0) - c|^2 = R^2.
template <class R> inline bool sphere_hit(R* t_out, V4<R> q0, V4<R> q1, V3<R> c, R rad) {
    V3<R> o{q0.x - c.x, q0.y - c.y, q0.z - c.z};
    V3<R> d{q1.x - q0.x, q1.y - q0.y, q1.z - q0.z};
    R A = dot3(d, d), B = dot3(o, d), C = dot3(o, o) - rad * rad;
    if (A <= R(0)) return false;
    R disc = B * B - A * C;
    if (disc < R(0)) return false;
    R s = sqrt_(disc);
    R t = (-B - s) / A;
    if (t < R(0)) t = (-B + s) / A;
    if (t < R(0) || t > R(1)) return false;
    *t_out = t;
    return true;
}

// Keplerian circular-orbit four-velocity in Kerr-Schild Cartesian coordinates,
// co-rotating with the (signed) spin a of the equations; u = (z w, 0, -x w, u^t)
// with w = u^t * Omega, Omega = s/(r^{3/2} + |a|), s = sign(a) (+1 for a = 0),
// u^t = (r^{3/2} + |a|)/(r^{3/4} sqrt(r^{3/2} - 3 r^{1/2} + 2|a|)) (Bardeen,
// Press & Teukolsky 1972).  Zero-velocity (u^t from g_tt) for Minkowski.
template <class R> inline V4<R> kepler_velocity(V4<R> q, R r, const MetricParams<R>& p) {
    if (p.metric == kMinkowski) return {R(0), R(0), R(0), R(1)};
    R a = p.metric == kKerr ? p.spin : R(0);
    R s = a < R(0) ? R(-1) : R(1);
    R aa = std::fabs(a);
    R sr = sqrt_(r);
    R r32 = r * sr;
    R ut = (r32 + aa) / (sqrt_(r32) * sqrt_(r32 - R(3) * sr + R(2) * aa));
    R w = ut * s / (r32 + aa);
    return {q.z * w, R(0), -q.x * w, ut};
}

// Planck's law in W sr^-1 m^-3 with lambda in nm (internal_math.h:2511-2526).
template <class R> inline R planck(R lambda_nm, R T) {
    if (T <= R(0)) return R(0);
    R lm = lambda_nm * R(1e-9);
    const R t0 = R(kPlanck) * R(kSpeedOfLight);
    const R t1 = R(2) * t0 * R(kSpeedOfLight);
    R t2 = lm * lm * lm * lm * lm;
    R t3 = std::exp(t0 / (lm * R(kBoltzmann) * T)) - R(1);
    return (t1 / t2) * (R(1) / t3);
}

// Segment o + t d (d not normalised) against a two-sided triangle: the
// Moller-Trumbore solution of o + t d = v0 + u e1 + v e2, accepted for
// u, v >= 0, u + v <= 1 and 0 <= t < tmax.
template <class R> inline bool seg_tri(V3<R> o, V3<R> d, V3<R> v0, V3<R> e1, V3<R> e2, R tmax, R* t_out) {
    const V3<R> pv{d.y * e2.z - d.z * e2.y, d.z * e2.x - d.x * e2.z, d.x * e2.y - d.y * e2.x};
    const R det = dot3(e1, pv);
    if (det == R(0)) return false;
    const R inv = R(1) / det;
    const V3<R> tv{o.x - v0.x, o.y - v0.y, o.z - v0.z};
    const R u = dot3(tv, pv) * inv;
    if (u < R(0) || u > R(1)) return false;
    const V3<R> qv{tv.y * e1.z - tv.z * e1.y, tv.z * e1.x - tv.x * e1.z, tv.x * e1.y - tv.y * e1.x};
    const R v = dot3(d, qv) * inv;
    if (v < R(0) || u + v > R(1)) return false;
    const R t = dot3(e2, qv) * inv;
    if (!(t >= R(0) && t < tmax)) return false;
    *t_out = t;
    return true;
}

template <class R> inline V3<R> affine_point(const R m[12], V3<R> x) {
    return {m[0] * x.x + m[1] * x.y + m[2] * x.z + m[3], m[4] * x.x + m[5] * x.y + m[6] * x.z + m[7],
            m[8] * x.x + m[9] * x.y + m[10] * x.z + m[11]};
}
template <class R> inline V3<R> affine_dir(const R m[12], V3<R> x) {
    return {m[0] * x.x + m[1] * x.y + m[2] * x.z, m[4] * x.x + m[5] * x.y + m[6] * x.z, m[8] * x.x + m[9] * x.y + m[10] * x.z};
}

// Closest triangle on o + t d, t in [0, tb), over all instances; lowers tb.
template <class R>
inline bool mesh_segment(const MeshSet<R>& ms, V3<R> o, V3<R> d, R* tb, uint32_t* inst, uint32_t* tri) {
    bool found = false;
    for (uint32_t k = 0; k < ms.inst.size(); ++k) {
        const auto& I = ms.inst[k];
        const V3<R> oo = affine_point(I.w2o, o), dd = affine_dir(I.w2o, d);
        // the segment against the mesh's object-space bounds: drops only triangles it cannot meet
        R t0 = R(0), t1 = *tb;
        const R oc[3] = {oo.x, oo.y, oo.z}, dc[3] = {dd.x, dd.y, dd.z}, lc[3] = {I.lo.x, I.lo.y, I.lo.z}, hc[3] = {I.hi.x, I.hi.y, I.hi.z};
        bool miss = false;
        for (int c = 0; c < 3 && !miss; ++c) {
            const R pad = R(1e-4) * (hc[c] - lc[c]) + R(1e-6);
            if (dc[c] == R(0)) { miss = oc[c] < lc[c] - pad || oc[c] > hc[c] + pad; continue; }
            R a = (lc[c] - pad - oc[c]) / dc[c], b = (hc[c] + pad - oc[c]) / dc[c];
            if (a > b) std::swap(a, b);
            t0 = std::max(t0, a);
            t1 = std::min(t1, b);
            miss = t0 > t1;
        }
        if (miss) continue;
        const auto& M = ms.mesh[I.mesh];
        for (uint32_t j = 0; j < M.size(); ++j) {
            R t;
            if (seg_tri(oo, dd, M[j].v0, M[j].e1, M[j].e2, *tb, &t)) { *tb = t; *inst = k; *tri = j; found = true; }
        }
    }
    return found;
}

template <class R> inline V3<R> hermite3(V3<R> p0, V3<R> D0, V3<R> p1, V3<R> D1, R s) {
    const R s2 = s * s, s3 = s2 * s;
    const R a = R(2) * s3 - R(3) * s2 + R(1), b = s3 - R(2) * s2 + s, c = R(3) * s2 - R(2) * s3, e = s3 - s2;
    return {a * p0.x + b * D0.x + c * p1.x + e * D1.x, a * p0.y + b * D0.y + c * p1.y + e * D1.y, a * p0.z + b * D0.z + c * p1.z + e * D1.z};
}
template <class R> inline V3<R> hermite3_d(V3<R> p0, V3<R> D0, V3<R> p1, V3<R> D1, R s) {
    const R s2 = s * s;
    const R a = R(6) * s2 - R(6) * s, b = R(3) * s2 - R(4) * s + R(1), e = R(3) * s2 - R(2) * s;
    return {a * (p0.x - p1.x) + b * D0.x + e * D1.x, a * (p0.y - p1.y) + b * D0.y + e * D1.y, a * (p0.z - p1.z) + b * D0.z + e * D1.z};
}

constexpr int kHermiteSegments = 4, kHermiteFine = 8;
constexpr double kBaryEps = 1e-4;

// The first hit along n straight sub-chords through the curve's points at
// s_lo + j (s_hi - s_lo) / n, below s_max: the sub-chord's index or -1, the
// hit's curve parameter and its point on the sub-chord.
template <class R>
inline int curve_pass(const MeshSet<R>& ms, V3<R> p0, V3<R> D0, V3<R> p1, V3<R> D1, R s_lo, R s_hi, int n, R s_max, R* s_hit, V3<R>* x,
                      uint32_t* inst, uint32_t* tri) {
    const R ds = (s_hi - s_lo) / R(n);
    V3<R> q = s_lo == R(0) ? p0 : hermite3(p0, D0, p1, D1, s_lo);
    for (int j = 0; j < n; ++j) {
        const R tmax = std::min((s_max - s_lo) / ds - R(j), R(1));
        if (!(tmax > R(0))) break;
        const R s1 = s_lo + R(j + 1) * ds;
        const V3<R> e = s1 >= R(1) ? p1 : hermite3(p0, D0, p1, D1, s1);
        R tb = tmax;
        if (mesh_segment(ms, q, e - q, &tb, inst, tri)) {
            *s_hit = s_lo + (R(j) + tb) * ds;
            *x = q + tb * (e - q);
            return j;
        }
        q = e;
    }
    return -1;
}

// The chord of one step against the meshes, hits below s_max.  p0, p1 the end
// points and D0, D1 the end tangents times the step's parameter length (world
// SI).  Straight chord: one segment.  Hermite chord: kHermiteSegments straight
// sub-chords through points of the cubic, then kHermiteFine within the first
// that hits; three Newton steps on the hit triangle's plane (in object space)
// inside the fine sub-chord's parameter range move the hit onto the cubic,
// kept only where that crossing lies on the triangle (else the sub-chord's hit).
template <class R>
inline bool mesh_chord(const MeshSet<R>& ms, V3<R> p0, V3<R> p1, V3<R> D0, V3<R> D1, bool herm, R s_max, R* s_out, V3<R>* x_out,
                       uint32_t* inst_out, uint32_t* tri_out) {
    uint32_t inst = 0, tri = 0;
    if (!herm) {
        R tb = std::min(s_max, R(1));
        if (!mesh_segment(ms, p0, p1 - p0, &tb, &inst, &tri)) return false;
        *s_out = tb;
        *x_out = p0 + tb * (p1 - p0);
    } else {
        R s_hit = 0;
        V3<R> x{};
        const int k = curve_pass(ms, p0, D0, p1, D1, R(0), R(1), kHermiteSegments, s_max, &s_hit, &x, &inst, &tri);
        if (k < 0) return false;
        R lo = R(k) / R(kHermiteSegments), hi = R(k + 1) / R(kHermiteSegments);
        {
            R s2 = 0;
            V3<R> x2{};
            uint32_t i2 = 0, t2 = 0;
            const int j = curve_pass(ms, p0, D0, p1, D1, lo, hi, kHermiteFine, s_max, &s2, &x2, &i2, &t2);
            if (j >= 0) {
                const R ds = (hi - lo) / R(kHermiteFine);
                lo = lo + R(j) * ds;
                hi = lo + ds;
                s_hit = s2;
                x = x2;
                inst = i2;
                tri = t2;
            }
        }
        const auto& I = ms.inst[inst];
        const auto& T = ms.mesh[I.mesh][tri];
        const V3<R> n{T.e1.y * T.e2.z - T.e1.z * T.e2.y, T.e1.z * T.e2.x - T.e1.x * T.e2.z, T.e1.x * T.e2.y - T.e1.y * T.e2.x};
        R s = s_hit;
        for (int it = 0; it < 3; ++it) {
            const R f = dot3(n, affine_point(I.w2o, hermite3(p0, D0, p1, D1, s)) - T.v0);
            const R fd = dot3(n, affine_dir(I.w2o, hermite3_d(p0, D0, p1, D1, s)));
            if (fd == R(0)) break;
            s = std::min(std::max(s - f / fd, lo), hi);
        }
        const V3<R> xc = hermite3(p0, D0, p1, D1, s);
        const V3<R> w = affine_point(I.w2o, xc) - T.v0;
        const R d00 = dot3(T.e1, T.e1), d01 = dot3(T.e1, T.e2), d11 = dot3(T.e2, T.e2), d20 = dot3(w, T.e1), d21 = dot3(w, T.e2);
        const R den = d00 * d11 - d01 * d01;
        const R bu = (d11 * d20 - d01 * d21) / den, bv = (d00 * d21 - d01 * d20) / den;
        const R eps = R(kBaryEps);
        if (bu >= -eps && bv >= -eps && bu + bv <= R(1) + eps && s < s_max) {
            *s_out = s;
            *x_out = xc;
        } else {
            *s_out = s_hit;
            *x_out = x;
        }
    }
    *inst_out = inst;
    *tri_out = tri;
    return true;
}

// World-space unit normal of a mesh hit, turned against the coordinate
// velocity v of the arriving geodesic.
template <class R> inline V3<R> mesh_normal(const MeshSet<R>& ms, uint32_t inst, uint32_t tri, V3<R> v) {
    const auto& I = ms.inst[inst];
    const V3<R> no = ms.mesh[I.mesh][tri].n;
    V3<R> n = norm3(V3<R>{I.nm[0] * no.x + I.nm[1] * no.y + I.nm[2] * no.z, I.nm[3] * no.x + I.nm[4] * no.y + I.nm[5] * no.z,
                          I.nm[6] * no.x + I.nm[7] * no.y + I.nm[8] * no.z});
    if (dot3(n, v) > R(0)) n = -n;
    return n;
}

// Trace one ray from its initial conditions to a termination class.
// Per step: solve_step, then the chord tests of kernel_intersect.cu:55-138
// (closest of disk / sphere / celestial sphere at t = 1; else horizon),
// then the step limit of kernel_enforce_ray_limits.cu:37-64 (n_step > limit).
template <class R>
inline Hit<R> trace_ray(const RayIC<R>& ic, const MetricParams<R>& p, const Scene<R>& sc, V3<R> cam_origin, uint32_t n_step0 = 0) {
    Hit<R> hit;
    State<R> y0 = ic.y;
    R h = ic.h, h_taken = 0;
    R L = length_scale(p);
    uint32_t n_step = n_step0, n_trial = 0;
    for (;;) {
        uint32_t tr = 0;
        State<R> y1 = solve_step(y0, p, &h, &h_taken, &tr);
        n_trial += tr;
        ++n_step;
        R best = R(1e30f);
        uint32_t cls = HIT_NONE;
        R t, rr = 0, r_hit = 0;
        V3<R> x_disk{};
        bool disk_geo = false;
        // with the crossing on the geodesic the chord's radius only selects the
        // steps to refine, over radii widened by the chord's error (factors
        // 0.7 and 1.02 on the disk's radii); the geodesic point decides
        const bool geo = sc.disk_on_geodesic && p.metric != kMinkowski;
        const R chord_in = geo ? sc.disk_r_in * R(0.7f) : sc.disk_r_in, chord_out = geo ? sc.disk_r_out * R(1.02f) : sc.disk_r_out;
        if (sc.disk && disk_hit(&t, &rr, y0.q, y1.q, p.metric == kKerr ? p.spin : R(0), chord_in, chord_out)) {
            bool keep = t < best;
            if (keep && geo) {
                auto f = [&](const State<R>& s) { return rhs(s, p); };
                keep = disk_on_geodesic(f, y0, t, h_taken, p.metric == kKerr ? p.spin : R(0), sc.disk_r_in, sc.disk_r_out, &rr, &x_disk);
                disk_geo = keep;
            }
            if (keep) { best = t; cls = HIT_DISK; r_hit = rr; }
        }
        if (sc.sphere && sphere_hit(&t, y0.q, y1.q, sc.sphere_center, sc.sphere_radius)) {
            if (t < best) { best = t; cls = HIT_SPHERE; }
        }
        int box_normal = 0;
        if (sc.n_box > 0) {
            // kernel_intersect.cu:63-105: the chord in SI metres
            const V3<R> o = V3<R>{y0.q.x, y0.q.y, y0.q.z} * L, e = V3<R>{y1.q.x, y1.q.y, y1.q.z} * L;
            const V3<R> d = e - o;
            const R len = sqrt_(len2_3(d));
            const V3<R> dir = d / len;
            for (uint32_t b = 0; b < sc.n_box; ++b) {
                int nc = 0;
                if (aabb_hit(&t, dir, o, R(1) / len, sc.box[b].mn, sc.box[b].mx, &nc) && t < best) {
                    best = t;
                    cls = sc.box[b].matter == MATTER_DIFFUSE ? HIT_BOX_DIFFUSE : HIT_BOX_EMITTER;
                    box_normal = nc;
                }
            }
        }
        uint32_t m_inst = 0, m_tri = 0;
        V3<R> m_x{};
        if (sc.meshes) {
            const V3<R> p0 = V3<R>{y0.q.x, y0.q.y, y0.q.z} * L, p1 = V3<R>{y1.q.x, y1.q.y, y1.q.z} * L;
            V3<R> D0{}, D1{};
            if (sc.hermite) {
                // end tangents dx/dlambda = g^{mu nu} p_nu times the step's parameter length, in metres
                const V4<R> v0 = contract(compute_g(y0.q, p, true), y0.p), v1 = contract(compute_g(y1.q, p, true), y1.p);
                const R k = R(float(h_taken)) * L;
                D0 = V3<R>{v0.x, v0.y, v0.z} * k;
                D1 = V3<R>{v1.x, v1.y, v1.z} * k;
            }
            R s;
            if (mesh_chord(*sc.meshes, p0, p1, D0, D1, sc.hermite, std::min(best, R(1)), &s, &m_x, &m_inst, &m_tri)) {
                best = s;
                cls = sc.meshes->inst[m_inst].matter == MATTER_DIFFUSE ? HIT_MESH_DIFFUSE : HIT_MESH_EMITTER;
            }
        }
        {
            V3<R> x1 = V3<R>{y1.q.x, y1.q.y, y1.q.z} * L - cam_origin;
            if (!(len2_3(x1) < sc.celestial_radius_m * sc.celestial_radius_m)) {
                if (R(1) < best) { best = R(1); cls = HIT_ESCAPE; }
            }
        }
        if (cls == HIT_NONE && in_domain(y1.q, p)) cls = HIT_HORIZON;
        if (cls == HIT_NONE && n_step > sc.step_limit) cls = HIT_STEP_LIMIT;
        if (cls != HIT_NONE) {
            hit.cls = cls;
            hit.steps = n_step;
            hit.trials = n_trial;
            if (cls == HIT_DISK || cls == HIT_SPHERE) {
                V4<R> qh = y0.q + best * (y1.q - y0.q);
                V4<R> ph = y0.p + best * (y1.p - y0.p);
                if (cls == HIT_DISK) {
                    // circular emitter: u.p = u^t p_t + u^phi p_phi with the conserved p_t and
                    // p_phi = z p_x - x p_z of the step's end point
                    V4<R> u = kepler_velocity(V4<R>{R(0), R(0), R(1), R(0)}, r_hit, p);   // (w, 0, 0, u^t)
                    R lphi = y1.q.z * y1.p.x - y1.q.x * y1.p.z;
                    R z = ic.contraction / (u.t * y1.p.t + u.x * lphi);
                    hit.g = (std::isnan(z) || std::isinf(z)) ? R(1e-10f) : z;
                } else {
                    hit.g = redshift_factor(ic.contraction, reference_observer(qh, p), ph);
                }
                if (cls == HIT_DISK) {
                    hit.a0 = r_hit;
                    hit.a1 = disk_geo ? atan2_(x_disk.z, x_disk.x) : atan2_(qh.z, qh.x);
                } else {
                    V3<R> n = norm3(V3<R>{qh.x - sc.sphere_center.x, qh.y - sc.sphere_center.y, qh.z - sc.sphere_center.z});
                    hit.a0 = n.y;
                    hit.a1 = atan2_(n.z, n.x);
                }
            } else if (cls == HIT_BOX_DIFFUSE || cls == HIT_BOX_EMITTER) {
                // the hit point on the chord, the momentum of the step's end (kernel_intersect.cu:129-131)
                hit.at.q = y0.q + best * (y1.q - y0.q);
                hit.at.p = y1.p;
                hit.a0 = R(box_normal);
                hit.a1 = best;
                if (cls == HIT_BOX_EMITTER) hit.g = redshift_factor(ic.contraction, reference_observer(hit.at.q, p), y1.p);
            } else if (cls == HIT_MESH_DIFFUSE || cls == HIT_MESH_EMITTER) {
                // the hit point on the chord or the curve, the momentum of the step's end
                hit.at.q = V4<R>{m_x.x / L, m_x.y / L, m_x.z / L, y0.q.t + best * (y1.q.t - y0.q.t)};
                hit.at.p = y1.p;
                hit.a0 = R(m_tri);
                hit.a1 = R(m_inst);
                if (cls == HIT_MESH_EMITTER) hit.g = redshift_factor(ic.contraction, reference_observer(hit.at.q, p), y1.p);
            } else if (cls == HIT_ESCAPE) {
                V3<R> d = norm3(V3<R>{y1.q.x - y0.q.x, y1.q.y - y0.q.y, y1.q.z - y0.q.z});
                hit.a0 = d.y;
                hit.a1 = atan2_(d.z, d.x);
                hit.g = redshift_factor(ic.contraction, V4<R>{R(0), R(0), R(0), R(1)}, y1.p);
            }
            return hit;
        }
        y0 = y1;
    }
}

// A whole path: trace, and at diffuse box matter reflect and trace again, up to
// max_reflections reflections (kernel_enforce_ray_limits.cu:48-50); the step
// count runs on across reflections.  Reflection k (0-based) draws its two
// uniforms from Philox block 1 + k of the path.  The step size restarts from
// the initial conditions' h after each reflection.
template <class R>
inline Hit<R> trace_path(const RayIC<R>& ic, const MetricParams<R>& p, const Scene<R>& sc, V3<R> cam_origin, uint64_t seed,
                         uint32_t cam_id, uint32_t pixel, uint32_t sample, uint32_t* n_reflect = nullptr);

// ---------------------------------------------------------------------------
// Spectral pipeline.
// ---------------------------------------------------------------------------
// Phi((x - mean)/sigma) (internal_math.h:2019-2024).
template <class R> inline R norm_gaussian_cdf(R mean, R x, R sigma) {
    return (R(1) / R(2)) * (R(1) + std::erf((x - mean) / (sigma * sqrt_(R(2)))));
}

// Sensor irradiance of one sample (device_color_utilities.cuh:22-29).
template <class R>
inline R radiance_to_irradiance(R L, R throughput, R sensor_w_mm, R sensor_h_mm, R iso, R fps, R shutter, uint32_t W, uint32_t H) {
    R area = (sensor_w_mm * sensor_h_mm) / R(1e6f);
    R n_pixel = R(float(W * H));
    R num = L * throughput * (area / n_pixel) * iso * R(2) * R(kPi);
    R den = (R(360) * fps) / shutter;
    return num / den;
}

// Linear interpolation in an equidistant table (color_matching_functions.cuh:35-44).
inline void sample_cmf(const float* tab /*n x 3*/, uint32_t n, float wmin, float wmax, float w, float out[3]) {
    float spacing = (wmax - wmin) / float(n - 1);
    float fi = std::fmin(std::fmax((w - wmin) / spacing, 0.0f), float(n - 1));
    uint32_t lo = uint32_t(std::fmin(float(uint32_t(fi)), float(n - 2)));
    float t = fi - float(lo);
    for (int c = 0; c < 3; ++c) out[c] = (tab[3 * lo + c] * (1.0f - t)) + (tab[3 * (lo + 1) + c] * t);
}

// Bin-averaged CMF weights normalised so sum_i W_i.y * bin_width = 1
// (host_color_utilities.cuh:145-178), in float as the reference computes them.
inline void bin_weights(float* W /*n_bin x 3*/, uint32_t n_bin, float bmin, float bmax, const float* tab, uint32_t n_tab,
                        float tmin, float tmax) {
    float sum_y = 0.0f;
    float bin_range = (bmax - bmin) / float(n_bin);
    float tab_range = (tmax - tmin) / float(n_tab - 1);
    uint32_t sub = uint32_t(std::ceil(bin_range / tab_range));
    float sub_w = bin_range / float(sub);
    for (uint32_t i = 0; i < n_bin; i++) {
        float w[3] = {0, 0, 0};
        float x0 = bmin + float(i) * bin_range;
        for (uint32_t j = 0; j < sub; j++) {
            float c[3];
            sample_cmf(tab, n_tab, tmin, tmax, x0 + (float(j) + 0.5f) * sub_w, c);
            for (int k = 0; k < 3; ++k) w[k] = w[k] + c[k] * sub_w;
        }
        for (int k = 0; k < 3; ++k) W[3 * i + k] = w[k];
        sum_y += w[1];
    }
    float kk = 1.0f / sum_y, inv_bw = 1.0f / bin_range;
    for (uint32_t i = 0; i < n_bin * 3; i++) W[i] = W[i] * (kk * inv_bw);
}

// The wavelength sampler's table (host_color_utilities.cuh:188-233, with the
// 5% floor of color_pipeline.cuh:36): s = x + y + z resampled at n_tab points
// across [lo, hi]; cdf[i + 1] = cdf[i] + ((s_i + s_{i+1}) / 2 + 0.05 mean(s)) h
// in float, normalised; inverted at u = k / 255 by one walk up the CDF with
// linear interpolation inside the bin.
inline void cmf_inverse_cdf(const float* tab, uint32_t n_tab, float tmin, float tmax, float lo, float hi, float* inv /*256*/) {
    const uint32_t n = n_tab - 1;
    std::vector<float> s(n_tab), cdf(n + 1, 0.0f);
    const float h = (hi - lo) / float(n);
    for (uint32_t i = 0; i < n_tab; ++i) {
        float c[3];
        sample_cmf(tab, n_tab, tmin, tmax, lo + float(i) * h, c);
        s[i] = c[0] + c[1] + c[2];
    }
    float mean = 0.0f;
    for (uint32_t i = 0; i < n_tab; ++i) mean += s[i];
    mean /= float(n_tab);
    const float fl = 0.05f * mean;
    for (uint32_t i = 0; i < n; ++i) cdf[i + 1] = cdf[i] + (0.5f * (s[i] + s[i + 1]) + fl) * h;
    const float total = cdf[n];
    for (uint32_t j = 0; j <= n; ++j) cdf[j] /= total;
    uint32_t b = 0;
    for (uint32_t k = 0; k < 256; ++k) {
        const float u = float(k) / 255.0f;
        while (b < n - 1 && cdf[b + 1] < u) ++b;
        const float t = (u - cdf[b]) / std::fmax(cdf[b + 1] - cdf[b], 1e-12f);
        inv[k] = lo + (float(b) + t) * h;
    }
}

}  // namespace oracle

namespace oracle {

// ---------------------------------------------------------------------------
// Philox4x32-10 counter-based generator (Salmon, Moraes, Dror & Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11): 10 rounds of
// (hi, lo) = mulhilo(M, c) with Weyl-sequence key bumps.
// ---------------------------------------------------------------------------
inline void philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = uint64_t(0xD2511F53u) * c0;
        uint64_t p1 = uint64_t(0xCD9E8D57u) * c2;
        uint32_t hi0 = uint32_t(p0 >> 32), lo0 = uint32_t(p0);
        uint32_t hi1 = uint32_t(p1 >> 32), lo1 = uint32_t(p1);
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// 24-bit uniform in [0,1).
inline float u01(uint32_t v) { return float(v >> 8) * (1.0f / 16777216.0f); }

// Per-path uniforms keyed by (seed, camera) with counter (pixel, sample, block, 0):
// block 0 -> {jitter u, jitter v, hero-wavelength u, spare}.
inline void path_uniforms(uint64_t seed, uint32_t camera, uint32_t pixel, uint32_t sample, uint32_t block, float u[4]) {
    uint32_t ctr[4] = {pixel, sample, block, 0u};
    uint32_t key[2] = {uint32_t(seed) ^ (camera * 0x9E3779B9u), uint32_t(seed >> 32) + camera};
    uint32_t o[4];
    philox4x32_10(ctr, key, o);
    for (int i = 0; i < 4; ++i) u[i] = u01(o[i]);
}

template <class R>
inline Hit<R> trace_path(const RayIC<R>& ic, const MetricParams<R>& p, const Scene<R>& sc, V3<R> cam_origin, uint64_t seed,
                         uint32_t cam_id, uint32_t pixel, uint32_t sample, uint32_t* n_reflect) {
    RayIC<R> cur = ic;
    uint32_t steps = 0, trials = 0;
    for (uint32_t k = 0;; ++k) {
        Hit<R> h = trace_ray(cur, p, sc, cam_origin, steps);
        steps = h.steps;
        trials += h.trials;
        if ((h.cls != HIT_BOX_DIFFUSE && h.cls != HIT_MESH_DIFFUSE) || k >= sc.max_reflections) {
            h.trials = trials;
            if (n_reflect) *n_reflect = k;
            return h;
        }
        float u[4];
        path_uniforms(seed, cam_id, pixel, sample, 1 + k, u);
        if (h.cls == HIT_BOX_DIFFUSE) {
            reflect_diffuse(p, h.at, cur.contraction, int(h.a0), R(u[0]), R(u[1]), &cur.y, &cur.contraction);
        } else {
            const V4<R> v = contract(compute_g(h.at.q, p, true), h.at.p);
            const V3<R> ng = mesh_normal(*sc.meshes, uint32_t(h.a1), uint32_t(h.a0), V3<R>{v.x, v.y, v.z});
            reflect_matter(p, h.at, cur.contraction, ng, false, R(u[0]), R(u[1]), &cur.y, &cur.contraction);
        }
        cur.h = ic.h;
    }
}

// Star catalogue: n stars, direction uniform on the sphere (cos(theta) about +y
// and azimuth atan2(z, x)), magnitude uniform in [0, 8].
struct Star { float dx, dy, dz, brightness; };
inline std::vector<Star> star_catalogue(uint64_t seed, uint32_t n, float scale) {
    std::vector<Star> s(n);
    uint32_t key[2] = {uint32_t(seed), uint32_t(seed >> 32) ^ 0x53544152u};
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t ctr[4] = {i, 0u, 0u, 0x73746172u}, o[4];
        philox4x32_10(ctr, key, o);
        float c = 2.0f * u01(o[0]) - 1.0f;
        float ph = 2.0f * kPi * u01(o[1]) - kPi;
        float m = 8.0f * u01(o[2]);
        float sn = std::sqrt(std::fmax(0.0f, 1.0f - c * c));
        s[i] = {sn * std::cos(ph), c, sn * std::sin(ph), scale * std::pow(10.0f, -0.4f * m)};
    }
    return s;
}

// Sum of the brightness of every star whose disc (angular radius rho) holds d.
// Brute force over a catalogue pre-sorted by dy: only stars with
// |dy - d.y| <= rho can contain d.
struct StarIndex {
    std::vector<Star> sorted;
    float rho = 0, cos_rho = 1;
    void build(std::vector<Star> s, float rho_rad) {
        std::sort(s.begin(), s.end(), [](const Star& a, const Star& b) { return a.dy < b.dy; });
        sorted = std::move(s);
        rho = rho_rad;
        cos_rho = std::cos(rho_rad);
    }
    double lookup(double x, double y, double z) const {
        double acc = 0;
        auto lo = std::lower_bound(sorted.begin(), sorted.end(), float(y - rho - 1e-6),
                                   [](const Star& a, float v) { return a.dy < v; });
        const double ch = 2.0 * std::sin(0.5 * double(rho)), ch2 = ch * ch;
        for (auto it = lo; it != sorted.end() && it->dy <= y + rho + 1e-6; ++it) {
            double ex = double(it->dx) - x, ey = double(it->dy) - y, ez = double(it->dz) - z;
            if (ex * ex + ey * ey + ez * ez <= ch2) acc += it->brightness;
        }
        return acc;
    }
};

}  // namespace oracle
