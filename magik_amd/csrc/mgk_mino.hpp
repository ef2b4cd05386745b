// Disk crossing of a Kerr null geodesic from its constants of motion, by the
// Mino-time integrals of the separated radial and polar motion (Carter 1968;
// Mino 2003; root structure as in Gralla & Lupsasca, PRD 101, 044032, 2020).
//
// With E-normalised constants lambda = L / E and eta = Q / E^2 and the Mino
// time d tau = |E| d lambda_affine / Sigma, a photon's Boyer-Lindquist r and
// u = cos(theta) obey
//     (dr/dtau)^2 = R(r) = r^4 + A r^2 + B r + C,
//         A = a^2 - eta - lambda^2, B = 2 (eta + (lambda - a)^2), C = -a^2 eta,
//     (du/dtau)^2 = U(u) = eta - (eta + lambda^2 - a^2) u^2 - a^2 u^4
//                        = (u_+ - u^2) (a^2 u^2 + A_m),
// so the n-th crossing of the equator happens at a Mino time G_n fixed by the
// polar motion alone, and the radius there solves  int dr / sqrt(R) = G_n
// along the radial branch (inbound, or outbound past the turning point r_4).
// Kerr-Schild r and theta are Boyer-Lindquist's, and E = -p_t, L = z p_x -
// x p_z (spin along +y) and Carter's Q are the same in both coordinates.
//
// Every integral is a 16-node Gauss-Legendre rule over a smooth integrand:
// the radial ones in s = 1/r (P(s) = s^4 R(1/s) = 1 + A s^2 + B s^3 + C s^4),
// with s = s_4 - w^2 taking out the square-root singularity at a turning
// point, and the polar ones in u = sqrt(u_+) - w^2 likewise.  The crossing
// radius is found by a bracketed Newton iteration seeded by the caller's
// estimate.  Near-critical rays (the outer pair of R's roots within
// kMinoMinGap of each other outside the horizon), rays that cross the equator
// more than kMinoMaxCrossings times and radii that do not converge within
// kMinoMaxNewton steps are left unresolved.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace mgk {

constexpr int kMinoMaxCrossings = 4;
constexpr int kMinoMaxNewton = 12;   // Newton steps on a crossing's radius (bisection when one leaves its bracket)
constexpr double kMinoMinGap = 0.02;  // relative separation of the outer root pair below which a ray is unresolved

// Gauss-Legendre nodes and weights on [0, 1] (16 points; the Legendre
// polynomial's roots mapped from [-1, 1], weights halved).
struct Gl16 {
    static constexpr double x[16] = {0.005299532504175031, 0.0277124884633837, 0.06718439880608412, 0.1222977958224985,
                                     0.19106187779867811, 0.2709916111713863, 0.35919822461037054, 0.4524937450811813,
                                     0.5475062549188188, 0.6408017753896295, 0.7290083888286136, 0.8089381222013219,
                                     0.8777022041775016, 0.9328156011939159, 0.9722875115366163, 0.994700467495825};
    static constexpr double w[16] = {0.013576229705877019, 0.031126761969323853, 0.047579255841246296, 0.062314485627767015,
                                     0.07479799440828838, 0.08457825969750131, 0.0913017075224618, 0.09472530522753429,
                                     0.09472530522753429, 0.0913017075224618, 0.08457825969750131, 0.07479799440828838,
                                     0.062314485627767015, 0.047579255841246296, 0.031126761969323853, 0.013576229705877019};
};

// 1/sqrt(x) of the quadrature nodes: on the device the v_rsq_f64 instruction
// alone, far cheaper than a correctly rounded square root and division.
// Measured on MI355X over 2^20 arguments spread over [2^-8, 2^8]: at most
// 5.2e-8 relative error, 7.7e-9 on average, 3e-9 mean bias; the integrals,
// and a disk crossing's radius, are then good to ~5e-8 relative, the
// precision (2^-24) in which a hit's radius is stored.  A Newton step on
// top (~1e-15) cost 4 FP64 instructions per node and moved no radiance.
__host__ __device__ inline double mino_rsqrt(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rsq(x);
#else
    return 1.0 / sqrt(x);
#endif
}

enum MinoKind : int { kMinoUnresolved = 0, kMinoCaptured = 1, kMinoEscapes = 2, kMinoDisk = 3 };

struct MinoResult {
    int kind;        // MinoKind
    int crossing;    // which crossing of the equator meets the disk (1, 2, ...)
    int iters;       // Newton iterations of the crossing's radius
    double r;        // the crossing's radius (kMinoDisk)
};

// Constants of motion and the Mino-time directions of a Kerr-Schild state
// (q = x, y, z; p = p_x, p_y, p_z, p_t covariant; a the spin of the equations).
struct MinoRay {
    double lam, eta, a;
    double r, u;     // Boyer-Lindquist r and cos(theta) of the state
    double s;        // 1 / r
    int sr, su;      // signs of dr/dtau and du/dtau
};

__host__ __device__ inline MinoRay mino_ray(const double q[3], const double p[4], double a) {
    MinoRay m;
    const double x = q[0], y = q[1], z = q[2];
    const double rho2 = x * x + y * y + z * z;
    const double w = rho2 - a * a;
    const double sig = sqrt(w * w + 4.0 * a * a * y * y);
    const double r = sqrt(0.5 * (w + sig));
    const double d = r * r + a * a;
    // (six divisions, as reciprocals, where the formulas have eleven)
    const double inv_r = 1.0 / r, inv_d = 1.0 / d;
    // dx^i/dlambda = g^{i nu} p_nu = p_i - f (k.p - p_t) k_i with the Kerr-Schild f and k
    const double f = 2.0 * r * r * r / (r * r * r * r + a * a * y * y);
    const double kx = (r * x - a * z) * inv_d, ky = y * inv_r, kz = (r * z + a * x) * inv_d;
    const double kp = kx * p[0] + ky * p[1] + kz * p[2] - p[3];
    const double vx = p[0] - f * kp * kx, vy = p[1] - f * kp * ky, vz = p[2] - f * kp * kz;
    // r(x, y, z) of the quartic r^4 - (rho^2 - a^2) r^2 - a^2 y^2 = 0: grad r = (x r, y d / r, z r) / sig
    const double rdot = (x * r * vx + y * d * inv_r * vy + z * r * vz) / sig;
    const double u = y * inv_r;
    const double udot = (vy * r - y * rdot) * (inv_r * inv_r);
    const double E = -p[3], inv_E = 1.0 / E;
    const double L = z * p[0] - x * p[2];
    // Carter's constant with p_theta = cot(theta) (x p_x + z p_z) - r sin(theta) p_y
    const double s2 = fmax(1.0 - u * u, 1e-300), s = sqrt(s2), inv_s = 1.0 / s;
    const double pth = (u * inv_s) * (x * p[0] + z * p[2]) - r * s * p[1];
    const double Q = pth * pth + u * u * (L * L * (inv_s * inv_s) - a * a * E * E);
    m.lam = L * inv_E;
    m.eta = Q * (inv_E * inv_E);
    m.a = a;
    m.r = r;
    m.s = inv_r;
    m.u = u;
    m.sr = rdot < 0.0 ? -1 : 1;
    m.su = udot < 0.0 ? -1 : 1;
    return m;
}

namespace mino_detail {

// cos(acos(c) / 3) in float: on the device acosf and the hardware cosine (the
// angle lies in [0, pi/3]), ~1e-7.
__host__ __device__ inline float cos_third_acos(float c) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __cosf(acosf(c) * (1.0f / 3.0f));
#else
    return std::cos(std::acos(c) * (1.0f / 3.0f));
#endif
}

// Largest real root of the depressed cubic t^3 + p t + q = 0.  Three real
// roots: the trigonometric form in float, then one Newton step on the cubic in
// FP64 (the error e -> ~e^2; linear only where the two largest roots merge),
// ~60 instructions in place of ~250 for acos and cos in FP64.
__host__ __device__ inline double cubic_max_root(double p, double q) {
    const double disc = 0.25 * q * q + p * p * p * (1.0 / 27.0);
    if (disc > 0.0) {
        const double sd = sqrt(disc);
        return cbrt(-0.5 * q + sd) + cbrt(-0.5 * q - sd);
    }
    if (p >= 0.0) return cbrt(-q);   // p = 0 = disc: a triple root
    const double rp = sqrt(p * (-1.0 / 3.0));
    const double c = fmin(1.0, fmax(-1.0, (1.5 * q) / (p * rp)));
    double t = 2.0 * rp * double(cos_third_acos(float(c)));
    const double d = 3.0 * t * t + p;
    if (d != 0.0) t -= ((t * t + p) * t + q) / d;
    return t;
}

// Radial potential in s = 1/r: P(s) = 1 + A s^2 + B s^3 + C s^4.
struct Radial {
    double A, B, C;
    __host__ __device__ double P(double s) const { return 1.0 + s * s * (A + s * (B + s * C)); }
    __host__ __device__ double R(double r) const { return r * r * r * r + A * r * r + B * r + C; }
    __host__ __device__ double dR(double r) const { return 4.0 * r * r * r + 2.0 * A * r + B; }
};

// The outer root pair of R(r) (Ferrari: R = product of two quadratics through
// the positive root m of 8 m^3 + 8 A m^2 + (2 A^2 - 8 C) m - B^2 = 0): r_3,4 =
// z -/+ sqrt(-A/2 - z^2 - B/(4z)), z = sqrt(m/2).  Real pair: r3 <= r4, the
// largest root polished by Newton steps.  Complex pair: z +/- i w.
struct RootPair {
    bool ok, real;
    double r3, r4;   // real pair
    double z, w;     // complex pair
};
__host__ __device__ inline RootPair radial_pair(const Radial& R) {
    RootPair rp{false, false, 0.0, 0.0, 0.0, 0.0};
    const double A = R.A, B = R.B, C = R.C;
    // m^3 + A m^2 + (A^2/4 - C) m - B^2/8 = 0, m = t - A/3
    // (divisions by constants as products with their reciprocals)
    const double p = A * A * (-1.0 / 12.0) - C;
    const double q = A * A * A * (-1.0 / 108.0) + A * C * (1.0 / 3.0) - B * B * 0.125;
    const double m = cubic_max_root(p, q) - A * (1.0 / 3.0);
    if (!(m > 0.0)) return rp;
    const double z = sqrt(0.5 * m);
    const double rad = -0.5 * A - z * z - B / (4.0 * z);
    rp.ok = true;
    rp.z = z;
    if (!(rad >= 0.0)) {
        rp.w = sqrt(-rad);
        return rp;
    }
    rp.real = true;
    const double sr = sqrt(rad);
    double r = z + sr;
    for (int k = 0; k < 3; ++k) {
        const double d = R.dR(r);
        if (!(d != 0.0)) break;
        r -= R.R(r) / d;
    }
    rp.r4 = r;
    rp.r3 = z - sr;
    return rp;
}

}  // namespace mino_detail

// The disk and horizon radii of mino_disk with their reciprocals (per scene;
// a kernel sets them up once instead of dividing per path).
struct MinoDisk {
    double r_horizon, r_in, r_out;
    double s_h, s_in, s_out;   // 1 / r_horizon, 1 / r_in, 1 / r_out
};
__host__ __device__ inline MinoDisk mino_disk_radii(double r_horizon, double r_in, double r_out) {
    return MinoDisk{r_horizon, r_in, r_out, 1.0 / r_horizon, 1.0 / r_in, 1.0 / r_out};
}

// Disk crossing (or capture / escape) of the ray: the first crossing of the
// equator within [r_in, r_out].  r_seed > 0 asks for the crossing's radius and
// is the estimate the Newton iteration starts from when it lies on the
// crossing's branch; r_seed <= 0 asks for the class only (res.r stays 0).
// dk.r_horizon is the outer horizon radius 1 + sqrt(1 - a^2).  Only rays that start
// inward (sr < 0) and outside the disk plane (u != 0) are resolved.
__host__ __device__ inline MinoResult mino_disk(const MinoRay& ray, const MinoDisk& dk, double r_seed) {
    using namespace mino_detail;
    const double r_horizon = dk.r_horizon;
    MinoResult res{kMinoUnresolved, 0, 0, 0.0};
    const double a = ray.a, lam = ray.lam, eta = ray.eta;
    if (ray.sr >= 0 || ray.u == 0.0) return res;
    // ---- polar motion: the Mino times of the equator crossings
    const double Ap = a * a;                       // U = (u_+ - u^2)(Ap u^2 + Am)
    const double Bq = eta + lam * lam - a * a, Cq = eta;
    const Radial rad{a * a - eta - lam * lam, 2.0 * (eta + (lam - a) * (lam - a)), -a * a * eta};
    const double s_obs = ray.s, s_h = dk.s_h;
    const RootPair pr = radial_pair(rad);
    if (!pr.ok) return res;
    const double r4 = pr.r4;
    const bool turning = pr.real && r4 > r_horizon && r4 < ray.r;
    // a near-double root outside the horizon (the photon orbits of a ray close to
    // the critical curve) makes the radial integrands nearly singular: left unresolved
    if (pr.real && pr.r3 > r_horizon && pr.r4 - pr.r3 < kMinoMinGap * pr.r4) return res;
    if (!pr.real && pr.z > r_horizon && pr.z < ray.r && pr.w < kMinoMinGap * pr.z) return res;
    if (!(eta > 0.0)) {
        // no equator crossing: a plunging ray is captured, the others are left
        if (!turning) res.kind = kMinoCaptured;
        return res;
    }
    const double D = sqrt(Bq * Bq + 4.0 * Ap * Cq);
    const double up = Bq >= 0.0 ? 2.0 * Cq / (Bq + D) : (D - Bq) / (2.0 * Ap);
    const double Am = Bq >= 0.0 ? 0.5 * (Bq + D) : 2.0 * Ap * Cq / (D - Bq);
    const double su = sqrt(up);
    const double uo = fabs(ray.u);
    if (!(uo < su)) return res;
    // tail(u) = int_u^{sqrt(u_+)} du / sqrt(U), with u = sqrt(u_+) - w^2
    auto tail = [&](double u) {
        const double W = sqrt(su - u);
        double acc = 0.0;
#pragma unroll 4
        for (int k = 0; k < 16; ++k) {
            const double w = W * Gl16::x[k], uu = su - w * w;
            acc += Gl16::w[k] * mino_rsqrt((su + uu) * (Ap * uu * uu + Am));
        }
        return 2.0 * W * acc;
    };
    // the half period, tail(0) = int_0^{sqrt(u_+)} du / sqrt(U): with u = sqrt(u_+) sin(phi)
    // it is int_0^{pi/2} dphi / sqrt(A_m cos^2 + (A_m + a^2 u_+) sin^2) = pi / (2 AGM(sqrt(A_m),
    // sqrt(A_m + a^2 u_+))), the arithmetic-geometric mean converging quadratically (a
    // few square roots in place of 16 nodes)
    double ga = sqrt(Am), gb = sqrt(Am + Ap * up);
    for (int k = 0; k < 12 && fabs(ga - gb) > 1e-15 * ga; ++k) {
        const double gm = 0.5 * (ga + gb);
        gb = sqrt(ga * gb);
        ga = gm;
    }
    const double G_half = 3.141592653589793 / (ga + gb);
    const bool toward = (ray.su < 0) == (ray.u > 0.0);   // moving toward u = 0
    const double t_obs = tail(uo);
    const double G1 = toward ? G_half - t_obs : G_half + t_obs;

    // ---- radial motion
    const double s_in = dk.s_in, s_out = dk.s_out;
    const double s_seed = r_seed > 0.0 ? 1.0 / r_seed : -1.0;
    double s4 = 0.0;
    double b0 = 0.0, b1 = 0.0, b2 = 0.0, b3 = 0.0;   // Q(s) = P(s) / (s4 - s) = -(b3 s^3 + b2 s^2 + b1 s + b0)
    if (turning) {
        s4 = 1.0 / r4;
        b3 = rad.C;
        b2 = rad.B + s4 * b3;
        b1 = rad.A + s4 * b2;
        b0 = s4 * b1;
    }
    auto Qs = [&](double s) { return -(((b3 * s + b2) * s + b1) * s + b0); };
    // K(W) = int_{s4 - W^2}^{s4} ds / sqrt(P) (turning rays)
    auto K = [&](double W) {
        double acc = 0.0;
#pragma unroll 4
        for (int k = 0; k < 16; ++k) {
            const double w = W * Gl16::x[k];
            acc += Gl16::w[k] * mino_rsqrt(Qs(s4 - w * w));
        }
        return 2.0 * W * acc;
    };
    // J(s) = int_{s_obs}^{s} ds / sqrt(P) (plunging rays)
    auto J = [&](double s) {
        double acc = 0.0;
        const double ds = s - s_obs;
#pragma unroll 4
        for (int k = 0; k < 16; ++k) acc += Gl16::w[k] * mino_rsqrt(rad.P(s_obs + ds * Gl16::x[k]));
        return ds * acc;
    };
    // The crossings are taken in order.  With a seed (the caller wants the
    // radius) each one's radius is solved for on its branch (bracketed Newton,
    // seeded by the estimate when that lies on the branch) and a crossing
    // within [r_in, r_out] is the disk hit; Newton stops once a step is below
    // 1e-4 relative: the error left after it is of the order of its square
    // (the integrals are smooth in W and s; at most 4e-7 relative in radius
    // on the a = 0.9 and a = 0.99 disk frames against a converged solve).  Without one only the class is
    // wanted: a crossing meets the disk when its Mino time lies between those
    // of the disk's edges on its branch, and no radius is solved for.
    const bool want_r = s_seed > 0.0;
    if (turning) {
        const double W_obs = sqrt(s4 - s_obs), W_inf = sqrt(s4);
        const double K_obs = K(W_obs);
        double K_inf = -1.0;   // taken when a crossing lies past the turning point
        // the disk's window on each branch, as W = sqrt(s4 - s) ranges over it
        // (inbound: s in [max(s_out, s_obs), min(s_in, s4)]; outbound: s in
        // [s_out, min(s_in, s4)]); K is increasing in W.  Each end is
        // integrated when a crossing first needs it (K >= 0; -1: not yet).
        double Ka_in = -1.0, Kb = -1.0, Ka_out = -1.0;
        const double s_lo_in = fmax(s_out, s_obs), s_hi = fmin(s_in, s4);
        const bool win_in = s_lo_in < s_hi, win_out = s_out < s_hi;
        auto kb = [&]() { return Kb >= 0.0 ? Kb : (Kb = K(sqrt(s4 - s_hi))); };
        auto ka_in = [&]() { return Ka_in >= 0.0 ? Ka_in : (Ka_in = K(sqrt(s4 - s_lo_in))); };
        auto ka_out = [&]() { return Ka_out >= 0.0 ? Ka_out : (Ka_out = K(sqrt(s4 - s_out))); };
        for (int n = 1; n <= kMinoMaxCrossings; ++n) {
            const double G = G1 + 2.0 * double(n - 1) * G_half;
            // past the turning point the ray has escaped unless G < K_obs +
            // K_inf.  With a seed the test waits until the radius solve on the
            // outbound branch finds no crossing: one it finds lies before the
            // escape, and most disk paths then never need K_inf.
            auto escaped = [&]() {
                if (K_inf < 0.0) K_inf = K(W_inf);
                return G >= K_obs + K_inf;
            };
            if (!want_r && G > K_obs && escaped()) {
                res.kind = kMinoEscapes;
                return res;
            }
            if (!want_r) {
                const bool hit = G <= K_obs ? win_in && G <= K_obs - kb() && G >= K_obs - ka_in()
                                            : win_out && G >= K_obs + kb() && G <= K_obs + ka_out();
                if (hit) {
                    res.kind = kMinoDisk;
                    res.crossing = n;
                    return res;
                }
                continue;
            }
            // inbound (before the turning point): K(W) = K_obs - G, W in [0, W_obs];
            // outbound: K(W) = G - K_obs, W in [0, W_inf]
            const bool inbound = G <= K_obs;
            const double target = inbound ? K_obs - G : G - K_obs;
            double Wlo = 0.0, Whi = inbound ? W_obs : W_inf;
            // a crossing beyond the disk's radial range on either side cannot be
            // the hit: the branch's radii at its ends bound it
            if (s4 < s_out) {                                // the whole branch lies outside r_out
                if (!inbound && escaped()) {
                    res.kind = kMinoEscapes;
                    return res;
                }
                continue;
            }
            double W = 0.5 * (Wlo + Whi);
            if (s_seed > 0.0 && s_seed < s4 && s_seed > s4 - Whi * Whi) W = sqrt(s4 - s_seed);
            bool done = false, newton = false;
            for (int it = 0; it < kMinoMaxNewton && !done; ++it) {
                const double f = K(W) - target;
                if (f > 0.0) Whi = W; else Wlo = W;
                double Wn = W - 0.5 * f * sqrt(Qs(s4 - W * W));
                newton = Wn >= Wlo && Wn <= Whi;
                if (!newton) Wn = 0.5 * (Wlo + Whi);
                done = fabs(Wn - W) <= 1e-4 * W + 1e-14;
                W = Wn;
                res.iters += 1;
            }
            // no crossing found on the outbound branch (bisection steps toward
            // W_inf, where K stays below the target): the escape test decides
            if (!inbound && !(done && newton) && escaped()) {
                res.kind = kMinoEscapes;
                return res;
            }
            if (!done) return res;
            const double s = s4 - W * W;
            if (s < s_out || s > s_in) continue;             // this crossing misses the disk
            res.kind = kMinoDisk;
            res.crossing = n;
            res.r = 1.0 / s;
            return res;
        }
        return res;
    }
    // plunging: every crossing is inbound, until the horizon
    const double J_h = J(s_h);
    const double s_lo = fmax(s_out, s_obs), s_hi = fmin(s_in, s_h);
    const bool win = s_lo < s_hi;
    double Ja = 0.0, Jb = 0.0;   // the disk's window, integrated when a crossing before the horizon needs it
    bool have_window = false;
    for (int n = 1; n <= kMinoMaxCrossings; ++n) {
        const double G = G1 + 2.0 * double(n - 1) * G_half;
        if (G >= J_h) {
            res.kind = kMinoCaptured;
            return res;
        }
        if (!want_r) {
            if (win && !have_window) {
                Ja = J(s_lo);
                Jb = J(s_hi);
                have_window = true;
            }
            if (win && G >= Ja && G <= Jb) {
                res.kind = kMinoDisk;
                res.crossing = n;
                return res;
            }
            continue;
        }
        double lo = s_obs, hi = s_h;
        double s = 0.5 * (lo + hi);
        if (s_seed > lo && s_seed < hi) s = s_seed;
        bool done = false;
        for (int it = 0; it < kMinoMaxNewton && !done; ++it) {
            const double f = J(s) - G;
            if (f > 0.0) hi = s; else lo = s;
            double sn = s - f * sqrt(rad.P(s));
            if (!(sn >= lo && sn <= hi)) sn = 0.5 * (lo + hi);
            done = fabs(sn - s) <= 1e-4 * s;
            s = sn;
            res.iters += 1;
        }
        if (!done) return res;
        if (s < s_out || s > s_in) continue;                 // this crossing misses the disk
        res.kind = kMinoDisk;
        res.crossing = n;
        res.r = 1.0 / s;
        return res;
    }
    return res;
}

__host__ __device__ inline MinoResult mino_disk(const MinoRay& ray, double r_horizon, double r_in, double r_out, double r_seed) {
    return mino_disk(ray, mino_disk_radii(r_horizon, r_in, r_out), r_seed);
}

}  // namespace mgk
