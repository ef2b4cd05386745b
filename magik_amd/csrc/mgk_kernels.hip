// HIP kernels of the wavefront spectral geodesic tracer for gfx950.
//
//   k_generate_rays  : pinhole ray + observer tetrad + null covariant momentum
//                      per (pixel, sample), written SoA to HBM
//                      (kernel_generate_thin_lens_ray.cu:51-77 fused with
//                      kernel_generate_initial_conditions.cu:54-74).
//   k_trace          : persistent RKF45 integrator.  Each wave64 pulls chunks
//                      of ray indices from one device counter and refills idle
//                      lanes by ballot/mbcnt, so lanes whose geodesics end early
//                      take new rays instead of idling (replaces
//                      kernel_advance_geodesic / kernel_intersect /
//                      kernel_enforce_ray_limits and their host-driven queue
//                      loop, wavefront.cpp:90-117).  One RKF45 trial per loop
//                      iteration, so rejected trials in one lane do not stall
//                      the others.
//   k_shade          : per pixel, all samples: emission, g^5 boost, sensor
//                      irradiance, Gaussian redistribution into spectral bins
//                      (in LDS), fold to CIE XYZ (kernel_bxdf.cu emit path,
//                      kernel_redistribute_radiance.cu:55-92,
//                      kernel_splat_to_renderbuffer.cu:37-57).
//   k_geodesic_steps : magik_trace_geodesic's per-step records for one ray.
#include "mgk_device.hpp"
#include "mgk_internal.hpp"
#include "mgk_mesh.hpp"
#include "mgk_raygen.hpp"

namespace mgk {

// ---------------------------------------------------------------------------
// Small state helpers with an explicit, left-to-right evaluation order.
// ---------------------------------------------------------------------------
template <class S> __device__ __forceinline__ St<S> add(const St<S>& a, const St<S>& b) {
#pragma clang fp contract(off)
    return {a.qx + b.qx, a.qy + b.qy, a.qz + b.qz, a.qt + b.qt, a.px + b.px, a.py + b.py, a.pz + b.pz, a.pt + b.pt};
}
template <class S> __device__ __forceinline__ St<S> sub(const St<S>& a, const St<S>& b) {
#pragma clang fp contract(off)
    return {a.qx - b.qx, a.qy - b.qy, a.qz - b.qz, a.qt - b.qt, a.px - b.px, a.py - b.py, a.pz - b.pz, a.pt - b.pt};
}
template <class S> __device__ __forceinline__ St<S> mul(S s, const St<S>& a) {
#pragma clang fp contract(off)
    return {s * a.qx, s * a.qy, s * a.qz, s * a.qt, s * a.px, s * a.py, s * a.pz, s * a.pt};
}

// RMS norm of relativity_utilities.cuh:15-29.
template <class S>
__device__ __forceinline__ S rms_norm(const St<S>& e, const St<S>& y0, const St<S>& y1, S at, S rt) {
#pragma clang fp contract(off)
    auto c = [&](S ev, S a, S b) { return ev / (at + rt * fmax(fabs(a), fabs(b))); };
    S s0 = c(e.qt, y0.qt, y1.qt), s1 = c(e.qx, y0.qx, y1.qx), s2 = c(e.qy, y0.qy, y1.qy), s3 = c(e.qz, y0.qz, y1.qz);
    S s4 = c(e.pt, y0.pt, y1.pt), s5 = c(e.px, y0.px, y1.px), s6 = c(e.py, y0.py, y1.py), s7 = c(e.pz, y0.pz, y1.pz);
    S total = (s0 * s0 + s1 * s1 + s2 * s2 + s3 * s3 + s4 * s4 + s5 * s5 + s6 * s6 + s7 * s7) / S(8);
    return sqrt_<false>(total);
}

// The same norm for the split (mixed / near-zone) trials.  The FP64 error
// components and scales are formed first and their ratios taken in FP32 with
// a hardware reciprocal.  In units u = 2^-24 each ratio is within 2.5 u of
// the exact one (two roundings to float, the 1 ulp reciprocal, the product),
// each square within 5.5 u, the sum of the eight within 9 u and the root
// within 5.5 u (3.3e-7) of the FP64 norm of the same components.  Within
// kNormBand of 1 the norm is taken again in FP64, so every accept/reject
// decision is the FP64 norm's; the step-size factor keeps the FP32 value.
constexpr float kNormBand = 0x1p-14f;   // 6.1e-5, ~185x the FP32 form's error bound

__device__ __forceinline__ double rms_norm_split(const St<double>& e, const St<double>& y0, const St<double>& y1, double at, double rt) {
    auto c = [&](double ev, double a, double b) {
        const float sc = float(at + rt * fmax(fabs(a), fabs(b)));
        return float(ev) * __builtin_amdgcn_rcpf(sc);
    };
    const float s0 = c(e.qt, y0.qt, y1.qt), s1 = c(e.qx, y0.qx, y1.qx), s2 = c(e.qy, y0.qy, y1.qy), s3 = c(e.qz, y0.qz, y1.qz);
    const float s4 = c(e.pt, y0.pt, y1.pt), s5 = c(e.px, y0.px, y1.px), s6 = c(e.py, y0.py, y1.py), s7 = c(e.pz, y0.pz, y1.pz);
    const float total = (s0 * s0 + s1 * s1 + s2 * s2 + s3 * s3 + s4 * s4 + s5 * s5 + s6 * s6 + s7 * s7) * 0.125f;
    const float n32 = __builtin_amdgcn_sqrtf(total);
    if (fabsf(n32 - 1.0f) < kNormBand) {
        auto d = [&](double ev, double a, double b) {
            const double q = ev / (at + rt * fmax(fabs(a), fabs(b)));
            return q * q;
        };
        const double t = d(e.qt, y0.qt, y1.qt) + d(e.qx, y0.qx, y1.qx) + d(e.qy, y0.qy, y1.qy) + d(e.qz, y0.qz, y1.qz) +
                         d(e.pt, y0.pt, y1.pt) + d(e.px, y0.px, y1.px) + d(e.py, y0.py, y1.py) + d(e.pz, y0.pz, y1.pz);
        return sqrt(t * 0.125);
    }
    return double(n32);
}

// x^0.2 of the step controller: the FP32 instance is the correctly rounded
// power (the double result rounded once), the value the CPU oracle also takes,
// so reference-parity trajectories do not part on the last bit of a powf.
__device__ __forceinline__ float powr(float a, float b) { return float(pow(double(a), double(b))); }
__device__ __forceinline__ double powr(double a, double b) { return pow(a, b); }

// One RKF45 trial in the reference's arithmetic (internal_math.h:2167-2199),
// contraction off so the FP32 instance rounds as the reference does.
template <int METRIC, class S>
__device__ __forceinline__ St<S> trial_reference(const St<S>& y0, S th, S a, S ms, S at, S rt, S h_max, S& h_new, bool& acc,
                                                 St<S>* k5_out = nullptr) {
#pragma clang fp contract(off)
    St<S> k1 = rhs_ref<METRIC, S>(y0, a, ms);
    St<S> k2 = rhs_ref<METRIC, S>(add(y0, mul(S(1) / S(4) * th, k1)), a, ms);
    St<S> k3 = rhs_ref<METRIC, S>(add(add(y0, mul(S(3) / S(32) * th, k1)), mul(S(9) / S(32) * th, k2)), a, ms);
    St<S> k4 = rhs_ref<METRIC, S>(
        add(sub(add(y0, mul(S(1932) / S(2197) * th, k1)), mul(S(7200) / S(2197) * th, k2)), mul(S(7296) / S(2197) * th, k3)), a, ms);
    St<S> k5 = rhs_ref<METRIC, S>(
        sub(add(sub(add(y0, mul(S(439) / S(216) * th, k1)), mul(S(8) * th, k2)), mul(S(3680) / S(513) * th, k3)),
            mul(S(845) / S(4104) * th, k4)),
        a, ms);
    St<S> k6 = rhs_ref<METRIC, S>(
        sub(add(sub(add(sub(y0, mul(S(8) / S(27) * th, k1)), mul(S(2) * th, k2)), mul(S(3544) / S(2565) * th, k3)),
                mul(S(1859) / S(4104) * th, k4)),
            mul(S(11) / S(40) * th, k5)),
        a, ms);
    if (k5_out) *k5_out = k5;
    St<S> y1 = add(sub(add(add(add(y0, mul(S(16) / S(135) * th, k1)), mul(S(6656) / S(12825) * th, k3)),
                           mul(S(28561) / S(56430) * th, k4)),
                       mul(S(9) / S(50) * th, k5)),
                   mul(S(2) / S(55) * th, k6));
    St<S> err = sub(sub(add(add(mul(S(-1) / S(360) * th, k1), mul(S(128) / S(4275) * th, k3)), mul(S(2197) / S(75240) * th, k4)),
                        mul(S(1) / S(50) * th, k5)),
                    mul(S(2) / S(55) * th, k6));
    S rms = rms_norm(err, y0, y1, at, rt);
    S rms_safe = rms > S(1e-30f) ? rms : S(1e-30f);
    h_new = S(0.9f) * th * powr(S(1) / rms_safe, S(0.2f));
    if (h_new > h_max) h_new = h_max;
    acc = rms_safe < S(1);
    return y1;
}

// Fehlberg tableau, and its 5th-order (b) and error (e = 4th - 5th) weights.
constexpr double A21 = 1.0 / 4, A31 = 3.0 / 32, A32 = 9.0 / 32;
constexpr double A41 = 1932.0 / 2197, A42 = -7200.0 / 2197, A43 = 7296.0 / 2197;
constexpr double A51 = 439.0 / 216, A52 = -8.0, A53 = 3680.0 / 513, A54 = -845.0 / 4104;
constexpr double A61 = -8.0 / 27, A62 = 2.0, A63 = -3544.0 / 2565, A64 = 1859.0 / 4104, A65 = -11.0 / 40;
constexpr double B1 = 16.0 / 135, B3 = 6656.0 / 12825, B4 = 28561.0 / 56430, B5 = -9.0 / 50, B6 = 2.0 / 55;
constexpr double E1 = -1.0 / 360, E3 = 128.0 / 4275, E4 = 2197.0 / 75240, E5 = -1.0 / 50, E6 = -2.0 / 55;
// beta_j = sum_i b_i a_ij and eps_j = sum_i e_i a_ij: the weights of the stage
// momentum derivatives inside the position update once the stage momenta
// p_i = p0 + h sum_j a_ij dp_j are substituted exactly.
constexpr double BT1 = B3 * A31 + B4 * A41 + B5 * A51 + B6 * A61;
constexpr double BT2 = B3 * A32 + B4 * A42 + B5 * A52 + B6 * A62;
constexpr double BT3 = B4 * A43 + B5 * A53 + B6 * A63;
constexpr double BT4 = B5 * A54 + B6 * A64;
constexpr double BT5 = B6 * A65;
constexpr double ET1 = E3 * A31 + E4 * A41 + E5 * A51 + E6 * A61;
constexpr double ET2 = E3 * A32 + E4 * A42 + E5 * A52 + E6 * A62;
constexpr double ET3 = E4 * A43 + E5 * A53 + E6 * A63;
constexpr double ET4 = E5 * A54 + E6 * A64;
constexpr double ET5 = E6 * A65;

// One RKF45 trial of the mixed-precision path.  The six right-hand sides and
// the stage states run in SS (float in the bulk, double in the near zone).
// The update and the error estimate are assembled in FP64 from the split
// right-hand side dq = P(p) - c: with the stage momenta substituted exactly,
//   y1_q  = y0_q + h P(p0) + h^2 sum beta_j P(dp_j) - h sum b_i c_i
//   err_q =                 h^2 sum eps_j  P(dp_j) - h sum e_i c_i
// so the large, cancelling P(p) terms never pass through FP32 and the error
// estimate carries no FP32 rounding noise of |p|.
constexpr double kA[6][5] = {{0, 0, 0, 0, 0},
                              {A21, 0, 0, 0, 0},
                              {A31, A32, 0, 0, 0},
                              {A41, A42, A43, 0, 0},
                              {A51, A52, A53, A54, 0},
                              {A61, A62, A63, A64, A65}};
constexpr double kB[6] = {B1, 0, B3, B4, B5, B6};
constexpr double kE[6] = {E1, 0, E3, E4, E5, E6};
constexpr double kBT[6] = {BT1, BT2, BT3, BT4, BT5, 0};
constexpr double kET[6] = {ET1, ET2, ET3, ET4, ET5, 0};

template <int METRIC, class SS, bool FAST>
__device__ __forceinline__ void trial_split(const St<double>& y0, double h, float a, float ms, const TraceParams& P, St<double>& y1,
                                            double& rms_out, SS* tangent5 = nullptr) {
    const SS qx = SS(y0.qx), qy = SS(y0.qy), qz = SS(y0.qz);
    const SS px = SS(y0.px), py = SS(y0.py), pz = SS(y0.pz), pt = SS(y0.pt);
    const SS hs = SS(h), as = SS(a), mss = SS(ms);
    // Stage sums are folded forward as soon as a stage is evaluated, so only the
    // pending partial states of the later stages stay live (not all six k's).
    SS Sq[6][3], Sp[6][3];
#pragma unroll
    for (int j = 0; j < 6; ++j)
#pragma unroll
        for (int c = 0; c < 3; ++c) Sq[j][c] = Sp[j][c] = SS(0);
    // Weighted sums of the stage values in the stage precision: none of them
    // carries the large, cancelling P(p) term (that enters the FP64 assembly
    // as h P(p0) only), so their rounding stays far below the tolerance.
    SS bp[3] = {0, 0, 0}, ep[3] = {0, 0, 0}, bt[3] = {0, 0, 0}, et[3] = {0, 0, 0};
    SS cb[4] = {0, 0, 0, 0}, ce[4] = {0, 0, 0, 0};
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        const SS sqx = qx + Sq[i][0], sqy = qy + Sq[i][1], sqz = qz + Sq[i][2];
        const SS spx = px + Sp[i][0], spy = py + Sp[i][1], spz = pz + Sp[i][2];
        const Rhs<SS> r = rhs<METRIC, SS, FAST>(sqx, sqy, sqz, spx, spy, spz, pt, as, mss);
        const SS kx = spx - r.cx, ky = spy - r.cy, kz = spz - r.cz;
        if (i == 4 && tangent5) {   // dx/dlambda of the fifth stage, evaluated at the trial's end (c5 = 1)
            tangent5[0] = kx; tangent5[1] = ky; tangent5[2] = kz;
        }
#pragma unroll
        for (int j = i + 1; j < 6; ++j) {
            const SS c = SS(kA[j][i]) * hs;
            Sq[j][0] += c * kx; Sq[j][1] += c * ky; Sq[j][2] += c * kz;
            Sp[j][0] += c * r.dpx; Sp[j][1] += c * r.dpy; Sp[j][2] += c * r.dpz;
        }
        if (kB[i] != 0.0) {
            const SS b = SS(kB[i]), ee = SS(kE[i]);
            bp[0] += b * r.dpx; bp[1] += b * r.dpy; bp[2] += b * r.dpz;
            ep[0] += ee * r.dpx; ep[1] += ee * r.dpy; ep[2] += ee * r.dpz;
            cb[0] += b * r.cx; cb[1] += b * r.cy; cb[2] += b * r.cz; cb[3] += b * r.ct;
            ce[0] += ee * r.cx; ce[1] += ee * r.cy; ce[2] += ee * r.cz; ce[3] += ee * r.ct;
        }
        if (kBT[i] != 0.0) {
            const SS bt_i = SS(kBT[i]), et_i = SS(kET[i]);
            bt[0] += bt_i * r.dpx; bt[1] += bt_i * r.dpy; bt[2] += bt_i * r.dpz;
            et[0] += et_i * r.dpx; et[1] += et_i * r.dpy; et[2] += et_i * r.dpz;
        }
    }
    // FP64 assembly (see above).
    const double hh = h * h;
    y1.px = y0.px + h * double(bp[0]);
    y1.py = y0.py + h * double(bp[1]);
    y1.pz = y0.pz + h * double(bp[2]);
    y1.pt = y0.pt;
    y1.qx = y0.qx + h * y0.px + (hh * double(bt[0]) - h * double(cb[0]));
    y1.qy = y0.qy + h * y0.py + (hh * double(bt[1]) - h * double(cb[1]));
    y1.qz = y0.qz + h * y0.pz + (hh * double(bt[2]) - h * double(cb[2]));
    y1.qt = y0.qt - h * y0.pt - h * double(cb[3]);
    St<double> e;
    e.px = h * double(ep[0]); e.py = h * double(ep[1]); e.pz = h * double(ep[2]); e.pt = 0.0;
    e.qx = hh * double(et[0]) - h * double(ce[0]); e.qy = hh * double(et[1]) - h * double(ce[1]);
    e.qz = hh * double(et[2]) - h * double(ce[2]); e.qt = -h * double(ce[3]);
    rms_out = rms_norm_split(e, y0, y1, double(P.abs_tol), double(P.rel_tol));
}

// Null-constraint renormalisation (FP64): rescale the spatial momentum by the
// root s ~ 1 of H(s) = 1/2 g^{mu nu} p_mu p_nu = 0 with p_t (the conserved
// energy) held fixed.  With g^{mu nu} = eta - f L L, L = (k, -1):
//   (A - f B^2) s^2 + 2 f B p_t s - (1 + f) p_t^2 = 0,  A = |p|^2, B = k.p.
template <int METRIC>
__device__ __forceinline__ void renormalise_null(St<double>& y, double a, double ms) {
    if constexpr (METRIC == MINKOWSKI) {
        double A = y.px * y.px + y.py * y.py + y.pz * y.pz;
        double s = fabs(y.pt) / sqrt(A);
        if (fabs(s - 1.0) < 1e-3) { y.px *= s; y.py *= s; y.pz *= s; }
    } else {
        // divisions and square roots through the Newton-refined hardware
        // seeds (a few ulp): the scale s only has to restore H = 0 far below
        // the integration tolerance
        double f, kx, ky, kz;
        if constexpr (METRIC == KERR) {
            const double R = y.qx * y.qx + y.qy * y.qy + y.qz * y.qz - a * a;
            const double r = sqrt_<true>(0.5 * R + 0.5 * sqrt_<true>(R * R + 4.0 * a * a * y.qy * y.qy));
            const double r2 = r * r;
            f = 2.0 * r2 * r * rcp<true>(r2 * r2 + a * a * y.qy * y.qy);
            const double id = rcp<true>(r2 + a * a);
            kx = (r * y.qx - a * y.qz) * id;
            ky = y.qy * rcp<true>(r);
            kz = (r * y.qz + a * y.qx) * id;
        } else {
            const double ir = rcp<true>(sqrt_<true>(y.qx * y.qx + y.qy * y.qy + y.qz * y.qz));
            f = 2.0 * ms * ir;
            kx = y.qx * ir; ky = y.qy * ir; kz = y.qz * ir;
        }
        double A = y.px * y.px + y.py * y.py + y.pz * y.pz;
        double B = kx * y.px + ky * y.py + kz * y.pz;
        double al = A - f * B * B, be = 2.0 * f * B * y.pt, ga = -(1.0 + f) * y.pt * y.pt;
        double disc = be * be - 4.0 * al * ga;
        if (!(disc >= 0.0) || al == 0.0) return;
        double qq = -0.5 * (be + copysign(sqrt_<true>(disc), be));
        double s1 = qq * rcp<true>(al), s2 = qq != 0.0 ? ga * rcp<true>(qq) : s1;
        double s = fabs(s1 - 1.0) < fabs(s2 - 1.0) ? s1 : s2;
        if (fabs(s - 1.0) < 1e-3) { y.px *= s; y.py *= s; y.pz *= s; }
    }
}

// Constants of motion of a Kerr (or positive-mass Schwarzschild) null ray in
// Kerr-Schild coordinates with the spin along +y: the axial angular momentum
// L = z p_x - x p_z and Carter's constant
//   Q = p_theta^2 + cos^2(theta) (L^2 / sin^2(theta) - a^2 E^2),
// cos(theta) = y / r, p_theta = cot(theta) (x p_x + z p_z) - r sin(theta) p_y
// (x and z scale with sin(theta) at fixed r and azimuth, y = r cos(theta)),
// E = -p_t.  Carter (1968).
struct Invariants {
    double r, f, kx, ky, kz, c, s2;
    double L, Q, pth;
};

template <int METRIC>
__device__ __forceinline__ bool invariants(const St<double>& y, double a, Invariants& v) {
    const double a2 = METRIC == KERR ? a * a : 0.0;
    const double R = y.qx * y.qx + y.qy * y.qy + y.qz * y.qz - a2;
    v.r = sqrt_<true>(0.5 * R + 0.5 * sqrt_<true>(R * R + 4.0 * a2 * y.qy * y.qy));
    const double r2 = v.r * v.r, ir = rcp<true>(v.r);
    v.f = 2.0 * r2 * v.r * rcp<true>(r2 * r2 + a2 * y.qy * y.qy);
    const double id = rcp<true>(r2 + a2);
    v.kx = (v.r * y.qx - a * y.qz) * id;
    v.ky = y.qy * ir;
    v.kz = (v.r * y.qz + a * y.qx) * id;
    v.c = y.qy * ir;
    v.s2 = 1.0 - v.c * v.c;
    if (!(v.s2 > 1e-6)) return false;    // near the spin axis theta is singular
    const double s = sqrt_<true>(v.s2), cot = v.c * rcp<true>(s);
    v.L = y.qz * y.px - y.qx * y.pz;
    v.pth = cot * (y.qx * y.px + y.qz * y.pz) - v.r * s * y.py;
    v.Q = v.pth * v.pth + v.c * v.c * (v.L * v.L * rcp<true>(v.s2) - a2 * y.pt * y.pt);
    return true;
}

// Renormalisation of the constants of motion after a completed step (FP64):
// one minimum-norm Newton step on the spatial momentum (position and
// E = -p_t held) back onto H = 0, L = L0, Q = Q0, i.e. p -= J^T (J J^T)^-1 c
// with c = (H, L - L0, Q - Q0) and J their gradients in p.  Where the three
// gradients are nearly dependent (turning points, the spin axis) only the
// null constraint is restored.
template <int METRIC>
__device__ __forceinline__ void renormalise_constants(St<double>& y, double a, double ms, double L0, double Q0) {
    if constexpr (METRIC == MINKOWSKI) {
        renormalise_null<METRIC>(y, a, ms);
    } else {
        if (METRIC == SCHWARZSCHILD && ms < 0.0) { renormalise_null<METRIC>(y, a, ms); return; }
        Invariants v;
        if (!invariants<METRIC>(y, a, v)) { renormalise_null<METRIC>(y, a, ms); return; }
        const double f = METRIC == KERR ? v.f : 2.0 * ms * rcp<true>(v.r);
        const double kp = v.kx * y.px + v.ky * y.py + v.kz * y.pz - y.pt;
        const double H = 0.5 * (y.px * y.px + y.py * y.py + y.pz * y.pz - y.pt * y.pt - f * kp * kp);
        const double gH[3] = {y.px - f * kp * v.kx, y.py - f * kp * v.ky, y.pz - f * kp * v.kz};
        const double gL[3] = {y.qz, 0.0, -y.qx};
        const double s = sqrt_<true>(v.s2), cot = v.c * rcp<true>(s);
        const double wl = 2.0 * v.c * v.c * v.L * rcp<true>(v.s2);
        const double gQ[3] = {2.0 * v.pth * cot * y.qx + wl * gL[0], -2.0 * v.pth * v.r * s,
                              2.0 * v.pth * cot * y.qz + wl * gL[2]};
        auto dot = [](const double* u, const double* w) { return u[0] * w[0] + u[1] * w[1] + u[2] * w[2]; };
        const double g00 = dot(gH, gH), g01 = dot(gH, gL), g02 = dot(gH, gQ);
        const double g11 = dot(gL, gL), g12 = dot(gL, gQ), g22 = dot(gQ, gQ);
        const double c00 = g11 * g22 - g12 * g12, c01 = g02 * g12 - g01 * g22, c02 = g01 * g12 - g02 * g11;
        const double det = g00 * c00 + g01 * c01 + g02 * c02;
        // rows within ~0.01 rad of dependence amplify the residuals: null step only
        if (!(det > 1e-4 * g00 * g11 * g22)) { renormalise_null<METRIC>(y, a, ms); return; }
        const double c11 = g00 * g22 - g02 * g02, c12 = g01 * g02 - g00 * g12, c22 = g00 * g11 - g01 * g01;
        const double e0 = H, e1 = v.L - L0, e2 = v.Q - Q0;
        const double id = rcp<true>(det);
        const double l0 = (c00 * e0 + c01 * e1 + c02 * e2) * id;
        const double l1 = (c01 * e0 + c11 * e1 + c12 * e2) * id;
        const double l2 = (c02 * e0 + c12 * e1 + c22 * e2) * id;
        const double dx = l0 * gH[0] + l1 * gL[0] + l2 * gQ[0];
        const double dy = l0 * gH[1] + l1 * gL[1] + l2 * gQ[1];
        const double dz = l0 * gH[2] + l1 * gL[2] + l2 * gQ[2];
        // a correction beyond ~1e-5 |p| is no integration error (those are ~tol = 1e-7):
        // a degenerate solve, keep the null step only
        const double p2 = y.px * y.px + y.py * y.py + y.pz * y.pz;
        if (!(dx * dx + dy * dy + dz * dz < 1e-10 * p2)) { renormalise_null<METRIC>(y, a, ms); return; }
        y.px -= dx; y.py -= dy; y.pz -= dz;
    }
}

// Per-ray spacetime constants: the spin of the equations, the scaled horizon
// radius, the FP64-zone bound and the centre of the escape sphere.  Uniform
// over a launch for a single camera; per ray for batched cameras and for the
// surrogate's training rays (spin stored as ray component 10).
template <class S> struct Lane {
    S a, hr;
    int in_range;
    double zone_rho2;
    S cx, cy, cz;
    S disk_r_in;
};

// Prograde ISCO radius (Bardeen, Press & Teukolsky 1972), units of M.
__device__ inline float isco_radius_f(float a) {
    a = fabsf(a);
    float z1 = 1.0f + cbrtf(1.0f - a * a) * (cbrtf(1.0f + a) + cbrtf(1.0f - a));
    float z2 = sqrtf(3.0f * a * a + z1 * z1);
    return 3.0f + z2 - sqrtf((3.0f - z1) * (3.0f + z1 + 2.0f * z2));
}
__host__ __device__ inline double isco_radius(double a) {
    a = fabs(a);
    double z1 = 1.0 + cbrt(1.0 - a * a) * (cbrt(1.0 + a) + cbrt(1.0 - a));
    double z2 = sqrt(3.0 * a * a + z1 * z1);
    return 3.0 + z2 - sqrt((3.0 - z1) * (3.0 + z1 + 2.0 * z2));
}

// Outer (retrograde) photon-orbit radius 2 (1 + cos(2/3 acos|a|)) (Bardeen,
// Press & Teukolsky 1972).
__device__ __forceinline__ double photon_orbit_outer(double a) {
    return 2.0 * (1.0 + cos((2.0 / 3.0) * acos(fmin(1.0, fabs(a)))));
}

template <int METRIC, class S> __device__ __forceinline__ void lane_constants(const TraceParams& P, S a, Lane<S>& ln) {
    constexpr bool F32 = sizeof(S) == 4;
    ln.a = a;
    ln.disk_r_in = (P.per_ray && P.disk_isco) ? S(isco_radius_f(float(a))) : S(P.disk_r_in);
    if (!P.per_ray) {
        ln.hr = F32 ? S(P.horizon_r_f) : S(P.horizon_r_d);
        ln.in_range = P.spin_in_range;
        ln.zone_rho2 = P.zone_rho2;
        return;
    }
    if constexpr (METRIC == KERR) {
        ln.in_range = fabs(a) > S(1) ? 0 : 1;
        ln.hr = (S(1) + sqrt_<false>(S(1) - a * a)) * S(P.horizon_scalar);
        const float af = float(a);
        const float rp = 1.0f + sqrtf(fmaxf(0.0f, 1.0f - af * af));
        const float rph = 2.0f * (1.0f + cosf((2.0f / 3.0f) * acosf(fminf(1.0f, fabsf(af)))));
        const double zr = double(fmaxf(1.5f * rph, rp + 0.1f)) * 1.0001;   // a hair wider than the host's FP64 bound
        ln.zone_rho2 = zr * zr + double(af) * double(af);
    } else {
        ln.in_range = 1;
        ln.hr = F32 ? S(P.horizon_r_f) : S(P.horizon_r_d);
        ln.zone_rho2 = P.zone_rho2;
    }
}

// Azimuth of a hit record.  The FP32 reference path takes the correctly rounded
// atan2 (the double result rounded once), as the CPU oracle does, so its
// records match the oracle's bit for bit; the other paths take FP32 atan2 of
// the rounded arguments (the record is FP32 either way): an FP64 atan2 in the
// zone kernels alone would cost their integration loops ~45 VGPRs.
template <class S> __device__ __forceinline__ float azimuth(S z, S x) {
    if constexpr (sizeof(S) == 4) return float(atan2(double(z), double(x)));
    else return atan2f(float(z), float(x));
}

// Hit record of a terminated path (g, a0, a1).
template <int METRIC, class S>
__device__ __forceinline__ void hit_record(const St<S>& y0, const St<S>& y1, uint32_t cls, S best, S r_hit, S contraction,
                                        const TraceParams& P, S a, int box_n, const S hq[3], uint32_t m_tri, uint32_t m_inst,
                                        mgk_hit_t& rec) {
#pragma clang fp contract(off)
    rec.cls = cls;
    rec.g = 0.0f; rec.a0 = 0.0f; rec.a1 = 0.0f;
    if (cls >= MGK_HIT_BOX_DIFFUSE) {
        // boxes: a0 the face normal's code, a1 the chord parameter (kernel_intersect.cu:125-132);
        // meshes: a0 the triangle's index in its mesh, a1 the instance
        const bool mesh = cls >= MGK_HIT_MESH_DIFFUSE;
        rec.a0 = mesh ? float(m_tri) : float(box_n);
        rec.a1 = mesh ? float(m_inst) : float(best);
        if (cls == MGK_HIT_BOX_EMITTER || cls == MGK_HIT_MESH_EMITTER) {
            // matter at rest: the reference observer at the hit point (boxes: on the chord), the
            // momentum of the step's end
            S hx, hy, hz;
            if (mesh) {
                hx = hq[0]; hy = hq[1]; hz = hq[2];
            } else {
                hx = y0.qx + best * (y1.qx - y0.qx); hy = y0.qy + best * (y1.qy - y0.qy); hz = y0.qz + best * (y1.qz - y0.qz);
            }
            const V4<S> u = reference_observer<METRIC, S>(hx, hy, hz, a, S(P.mass_sign));
            S g = contraction / dot4(u, V4<S>{y1.px, y1.py, y1.pz, y1.pt});
            if (isnan(g) || isinf(g)) g = S(1e-10f);
            rec.g = float(g);
        }
    } else if (cls == MGK_HIT_DISK || cls == MGK_HIT_SPHERE) {
        V4<S> ph{y0.px + best * (y1.px - y0.px), y0.py + best * (y1.py - y0.py), y0.pz + best * (y1.pz - y0.pz),
                 y0.pt + best * (y1.pt - y0.pt)};
        S hx = y0.qx + best * (y1.qx - y0.qx), hy = y0.qy + best * (y1.qy - y0.qy), hz = y0.qz + best * (y1.qz - y0.qz);
        S g;
        if (cls == MGK_HIT_DISK) {
            // u.p of the circular emitter from the conserved p_t and p_phi = z p_x - x p_z of the
            // step's end point: exact along the geodesic, free of chord interpolation error
            S ut, w;
            kepler_ut_w<METRIC, S>(r_hit, a, ut, w);
            S lphi = y1.qz * y1.px - y1.qx * y1.pz;
            g = contraction / (ut * y1.pt + w * lphi);
        } else {
            V4<S> u = reference_observer<METRIC, S>(hx, hy, hz, a, S(P.mass_sign));
            g = contraction / dot4(u, ph);
        }
        if (isnan(g) || isinf(g)) g = S(1e-10f);
        rec.g = float(g);
        if (cls == MGK_HIT_DISK) {
            rec.a0 = float(r_hit);
            rec.a1 = azimuth<S>(hz, hx);
        } else {
            S nx = hx - S(P.sphere_c[0]), ny = hy - S(P.sphere_c[1]), nz = hz - S(P.sphere_c[2]);
            S l = sqrt_<false>(nx * nx + ny * ny + nz * nz);
            rec.a0 = float(ny / l);
            rec.a1 = azimuth<S>(nz / l, nx / l);
        }
    } else if (cls == MGK_HIT_ESCAPE) {
        S dx = y1.qx - y0.qx, dy = y1.qy - y0.qy, dz = y1.qz - y0.qz;
        S l = sqrt_<false>(dx * dx + dy * dy + dz * dz);
        rec.a0 = float(dy / l);
        rec.a1 = azimuth<S>(dz / l, dx / l);
        S g = contraction / y1.pt;
        if (isnan(g) || isinf(g)) g = S(1e-10f);
        rec.g = float(g);
    }
}

// Chord tests of one completed step (kernel_intersect.cu:55-138 extended with
// the disk and the emissive sphere), then the horizon (in_domain) and the
// step limit (kernel_enforce_ray_limits.cu:48, n_step > limit).  Returns the
// class and fills the hit record.
// FAST_TESTS (the mixed-precision zone kernels): the per-step escape-sphere
// and horizon threshold tests run in FP32; only rays within ~1e-7 of a
// threshold can see a different step, and the hit itself is still recorded
// from the FP64 state.
// Slab test of a chord against an axis-aligned box (traceables.cuh:17-74): o
// the chord's start, d its unit direction, inv_len 1 / its length (SI
// metres).  The hit parameter is in chord units; the normal, along the face's
// axis and against the chord, comes back as a code (axis, + 3 if negative).
template <class T>
__device__ __forceinline__ bool box_chord(T ox, T oy, T oz, T dx, T dy, T dz, T inv_len, const float* bmin, const float* bmax, T& t_out,
                                          int& code) {
#pragma clang fp contract(off)
    const T ix = T(1) / dx, iy = T(1) / dy, iz = T(1) / dz;
    const bool sx = ix < T(0), sy = iy < T(0), sz = iz < T(0);
    T tmin = (T(sx ? bmax[0] : bmin[0]) - ox) * ix, tmax = (T(sx ? bmin[0] : bmax[0]) - ox) * ix;
    const T tymin = (T(sy ? bmax[1] : bmin[1]) - oy) * iy, tymax = (T(sy ? bmin[1] : bmax[1]) - oy) * iy;
    if ((tmin > tymax) || (tymin > tmax)) return false;
    int amin = 0, amax = 0;
    if (tymin > tmin) { tmin = tymin; amin = 1; }
    if (tymax < tmax) { tmax = tymax; amax = 1; }
    const T tzmin = (T(sz ? bmax[2] : bmin[2]) - oz) * iz, tzmax = (T(sz ? bmin[2] : bmax[2]) - oz) * iz;
    if ((tmin > tzmax) || (tzmin > tmax)) return false;
    if (tzmin > tmin) { tmin = tzmin; amin = 2; }
    if (tzmax < tmax) { tmax = tzmax; amax = 2; }
    if (tmax < T(0)) return false;
    const bool inside = tmin < T(0);
    const T t = (inside ? tmax : tmin) * inv_len;
    if (t > T(1)) return false;
    const int axis = inside ? amax : amin;
    const bool neg_axis = axis == 0 ? sx : (axis == 1 ? sy : sz);
    T ns = neg_axis ? T(1) : T(-1);
    if (inside) ns = -ns;
    const T dcomp = axis == 0 ? dx : (axis == 1 ? dy : dz);
    if (ns * dcomp > T(0)) ns = -ns;
    t_out = t;
    code = axis + (ns < T(0) ? 3 : 0);
    return true;
}

// The disk crossing on the geodesic (MGK_SCENE_DISK_ON_GEODESIC): from the
// step's start y0, one RKF45 trial over the chord's crossing parameter t h --
// the 5th-order solution, no step control -- then one Newton step to the plane
// y = 0 along the tangent dx/dlambda of the trial's fifth stage, which
// Fehlberg's tableau evaluates at the trial's end (c5 = 1).  The remaining
// offset is second order in the chord estimate's error.  yc and the tangent
// (dx, dy, dz) come from the caller's trial; false when the corrected point
// lies off the disk (the path then goes on).
template <class S>
__device__ __forceinline__ bool disk_point(const St<S>& yc, S dx, S dy, S dz, S a, S r_in, S r_max, S& r, S& x, S& z) {
#pragma clang fp contract(off)
    const S dl = -yc.qy / dy;
    x = yc.qx + dl * dx;
    z = yc.qz + dl * dz;
    const S rho2 = x * x + z * z - a * a;
    if (!(rho2 > S(0))) return false;
    r = sqrt_<false>(rho2);
    return !(r < r_in || r > r_max);
}

// Factors on the disk's radii that widen the chord's crossing test when the
// crossing is located on the geodesic (MGK_SCENE_DISK_ON_GEODESIC).
constexpr float kDiskChordInner = 0.7f, kDiskChordOuter = 1.02f;

// A refine functor that does nothing: scenes without the on-geodesic disk.
struct NoDiskRefine {
    template <class S> __device__ bool operator()(S, S&, S&, S&) const { return true; }
};

template <int METRIC, class S, bool FAST_TESTS = false, bool MESH = false, class DiskRefine = NoDiskRefine>
__device__ __forceinline__ uint32_t classify(const St<S>& y0, const St<S>& y1, uint32_t n_step, S contraction, const TraceParams& P,
                                             const Lane<S>& ln, float dl, mgk_hit_t& rec, S& t_hit, S hq[3],
                                             const DiskRefine& refine = DiskRefine{}) {
#pragma clang fp contract(off)
    const S a = METRIC == KERR ? ln.a : S(0);
    S best = S(1e30f);
    uint32_t cls = MGK_HIT_NONE;
    S r_hit = S(0);
    int box_n = 0;
    uint32_t m_tri = 0, m_inst = 0;
    bool disk_geo = false;
    S disk_x = S(0), disk_z = S(0);
    if (P.flags & MGK_SCENE_DISK) {
        S ya = y0.qy, yb = y1.qy;
        if ((ya > S(0) && yb <= S(0)) || (ya < S(0) && yb >= S(0))) {
            S t = ya / (ya - yb);
            S x = y0.qx + t * (y1.qx - y0.qx), z = y0.qz + t * (y1.qz - y0.qz);
            S rho2 = x * x + z * z - a * a;
            if (rho2 > S(0)) {
                S r = sqrt_<false>(rho2);
                // with the crossing located on the geodesic, the chord's radius
                // only selects the steps to refine: it may lie off the disk by
                // the chord's O(h^2) error while the geodesic's lies on it
                // (steps near the inner edge at high spin are ~20 % of r), so
                // the test is widened and the geodesic point decides
                const bool geo = METRIC != MINKOWSKI && (P.flags & MGK_SCENE_DISK_ON_GEODESIC);
                const S lo = geo ? ln.disk_r_in * S(kDiskChordInner) : ln.disk_r_in;
                const S hi = geo ? S(P.disk_r_out) * S(kDiskChordOuter) : S(P.disk_r_out);
                if (!(r < lo || r > hi) && t < best) {
                    bool keep = true;
                    if (geo) {
                        keep = refine(t, r, disk_x, disk_z);
                        disk_geo = keep;
                    }
                    if (keep) { best = t; cls = MGK_HIT_DISK; r_hit = r; }
                }
            }
        }
    }
    if (P.flags & MGK_SCENE_SPHERE) {
        S ox = y0.qx - S(P.sphere_c[0]), oy = y0.qy - S(P.sphere_c[1]), oz = y0.qz - S(P.sphere_c[2]);
        S dx = y1.qx - y0.qx, dy = y1.qy - y0.qy, dz = y1.qz - y0.qz;
        S A = dx * dx + dy * dy + dz * dz, B = ox * dx + oy * dy + oz * dz;
        S Cc = ox * ox + oy * oy + oz * oz - S(P.sphere_r) * S(P.sphere_r);
        S disc = B * B - A * Cc;
        if (A > S(0) && disc >= S(0)) {
            S sd = sqrt_<false>(disc);
            S t = (-B - sd) / A;
            if (t < S(0)) t = (-B + sd) / A;
            if (t >= S(0) && t <= S(1) && t < best) { best = t; cls = MGK_HIT_SPHERE; }
        }
    }
    if (P.flags & MGK_SCENE_BOXES) {
        // the chord in SI metres (kernel_intersect.cu:63-70); FP32 in the zone kernels
        using T = typename std::conditional<FAST_TESTS, float, S>::type;
        const T L = T(P.length_scale);
        const T ox = T(y0.qx) * L, oy = T(y0.qy) * L, oz = T(y0.qz) * L;
        T dx = T(y1.qx) * L - ox, dy = T(y1.qy) * L - oy, dz = T(y1.qz) * L - oz;
        const T len = sqrt_<false>(dx * dx + dy * dy + dz * dz);
        dx = dx / len; dy = dy / len; dz = dz / len;
        const T inv_len = T(1) / len;
        for (uint32_t b = 0; b < P.n_box; ++b) {
            T t;
            int code;
            if (box_chord<T>(ox, oy, oz, dx, dy, dz, inv_len, P.box_min[b], P.box_max[b], t, code) && S(t) < best) {
                best = S(t);
                cls = P.box_matter[b] == MGK_MATTER_DIFFUSE ? MGK_HIT_BOX_DIFFUSE : MGK_HIT_BOX_EMITTER;
                box_n = code;
            }
        }
    }
    if constexpr (MESH) {
        // the chord in SI metres; FP32 in the zone kernels
        using T = typename std::conditional<FAST_TESTS, float, S>::type;
        const T L = T(P.length_scale);
        const T p0[3] = {T(y0.qx) * L, T(y0.qy) * L, T(y0.qz) * L}, p1[3] = {T(y1.qx) * L, T(y1.qy) * L, T(y1.qz) * L};
        T D0[3] = {T(0), T(0), T(0)}, D1[3] = {T(0), T(0), T(0)};
        const bool herm = P.chord == MGK_CHORD_HERMITE;
        if (herm) {
            // end tangents dx/dlambda = g^{mu nu} p_nu, times the step's parameter length, in metres
            const T ms = T(P.mass_sign), at = T(a), k = T(dl) * L;
            const V4<T> v0 = lower(metric_g<METRIC, T>(T(y0.qx), T(y0.qy), T(y0.qz), at, ms, true),
                                   V4<T>{T(y0.px), T(y0.py), T(y0.pz), T(y0.pt)});
            const V4<T> v1 = lower(metric_g<METRIC, T>(T(y1.qx), T(y1.qy), T(y1.qz), at, ms, true),
                                   V4<T>{T(y1.px), T(y1.py), T(y1.pz), T(y1.pt)});
            D0[0] = v0.x * k; D0[1] = v0.y * k; D0[2] = v0.z * k;
            D1[0] = v1.x * k; D1[1] = v1.y * k; D1[2] = v1.z * k;
        }
        MeshHit<T> mh;
        if (mesh_chord<T>(P, p0, p1, D0, D1, herm, T(min_(best, S(1))), mh)) {
            best = S(mh.s);
            cls = mh.matter == MGK_MATTER_DIFFUSE ? MGK_HIT_MESH_DIFFUSE : MGK_HIT_MESH_EMITTER;
            m_tri = mh.tri;
            m_inst = mh.inst;
            const S iL = S(1) / S(P.length_scale);
            hq[0] = S(mh.x[0]) * iL; hq[1] = S(mh.x[1]) * iL; hq[2] = S(mh.x[2]) * iL;
        }
    }
    if constexpr (FAST_TESTS) {
        const float L = P.length_scale;
        const float ex = float(y1.qx) * L - float(ln.cx), ey = float(y1.qy) * L - float(ln.cy), ez = float(y1.qz) * L - float(ln.cz);
        const float d2 = ex * ex + ey * ey + ez * ez;
        if (!(d2 < P.cel_r2_f) && S(1) < best) { best = S(1); cls = MGK_HIT_ESCAPE; }
    } else {
        const S L = S(P.length_scale);
        S ex = y1.qx * L - ln.cx, ey = y1.qy * L - ln.cy, ez = y1.qz * L - ln.cz;
        S d2 = ex * ex + ey * ey + ez * ez;
        S r2 = sizeof(S) == 4 ? S(P.cel_r2_f) : S(P.cel_r2_d);
        if (!(d2 < r2) && S(1) < best) { best = S(1); cls = MGK_HIT_ESCAPE; }
    }
    if (cls == MGK_HIT_NONE) {
        if constexpr (METRIC == KERR) {
            if (ln.in_range) {
                if constexpr (FAST_TESTS) {
                    const float af = float(a), yf = float(y1.qy);
                    const float R = float(y1.qx) * float(y1.qx) + yf * yf + float(y1.qz) * float(y1.qz) - af * af;
                    const float r2 = 0.5f * R + 0.5f * __builtin_amdgcn_sqrtf(R * R + 4.0f * af * af * yf * yf);
                    if (!(float(ln.hr) * float(ln.hr) < r2)) cls = MGK_HIT_HORIZON;
                } else {
                    S r = kerr_schild_r(y1.qx, y1.qy, y1.qz, a);
                    if (!(ln.hr < r)) cls = MGK_HIT_HORIZON;
                }
            }
        } else if constexpr (METRIC == SCHWARZSCHILD) {
            S r = sqrt_<false>(y1.qx * y1.qx + y1.qy * y1.qy + y1.qz * y1.qz);
            if (P.mass_sign < 0.0f) { if (r < S(1e-3f)) cls = MGK_HIT_HORIZON; }
            else if (r <= ln.hr) cls = MGK_HIT_HORIZON;
        }
    }
    if (cls == MGK_HIT_NONE && n_step > P.step_limit) cls = MGK_HIT_STEP_LIMIT;
    if (cls == MGK_HIT_NONE) return cls;
    hit_record<METRIC, S>(y0, y1, cls, best, r_hit, contraction, P, a, box_n, hq, m_tri, m_inst, rec);
    if (cls == MGK_HIT_DISK && disk_geo) rec.a1 = azimuth<S>(disk_z, disk_x);
    t_hit = best;
    return cls;
}

template <class S> __device__ __forceinline__ void write_final(const TraceParams& P, uint64_t ray, const St<S>& y) {
    double* f = P.final_state;
    const uint64_t sd = P.stride;
    f[0 * sd + ray] = double(y.qx); f[1 * sd + ray] = double(y.qy); f[2 * sd + ray] = double(y.qz); f[3 * sd + ray] = double(y.qt);
    f[4 * sd + ray] = double(y.px); f[5 * sd + ray] = double(y.py); f[6 * sd + ray] = double(y.pz); f[7 * sd + ray] = double(y.pt);
}

// Terminal state of a path: the step's end, or for box and mesh matter the hit
// point (on the chord, or on the Hermite curve) with the step's end momentum
// (kernel_intersect.cu:129-131).
template <class S>
__device__ __forceinline__ void write_terminal(const TraceParams& P, uint64_t ray, uint32_t cls, const St<S>& y0, const St<S>& y1, S t,
                                               const S hq[3]) {
#pragma clang fp contract(off)
    if (cls < MGK_HIT_BOX_DIFFUSE) {
        write_final(P, ray, y1);
        return;
    }
    St<S> h = y1;
    if (cls >= MGK_HIT_MESH_DIFFUSE) {
        h.qx = hq[0];
        h.qy = hq[1];
        h.qz = hq[2];
    } else {
        h.qx = y0.qx + t * (y1.qx - y0.qx);
        h.qy = y0.qy + t * (y1.qy - y0.qy);
        h.qz = y0.qz + t * (y1.qz - y0.qz);
    }
    h.qt = y0.qt + t * (y1.qt - y0.qt);
    write_final(P, ray, h);
}


template <class S> __device__ __forceinline__ S load_c(const void* rays, uint64_t stride, int c, uint64_t i) {
    return reinterpret_cast<const S*>(rays)[uint64_t(c) * stride + i];
}

// Ray indices a wave takes from the device work counter per refill: one
// wave's width, so a wave that starts on a chunk holds exactly its rays.
// Larger chunks (128: the 1080p Kerr integration 2.4 % faster on one MI355X)
// let a wave's later refills mix rays in a run-time order, and the mixed
// path's wave-uniform renormalisation then makes a ray's last bits depend on
// the launch it was traced in.
constexpr uint32_t kChunk = 64;

// Takes the next 64-ray chunk for a wave from the work counters.  Chunk g of
// [0, n) belongs to counter g % kHeads, which hands its chunks out in order;
// the counters sit on separate 128-byte lines, so the waves of a pass spread
// their atomics over kHeads addresses instead of one.  A wave starts on the
// counter of its workgroup's XCD (workgroups go to the XCDs round-robin) and
// moves on when that counter's chunks are gone; after every counter has come
// up empty it is done.  Wave-uniform; false when no chunk is left.
__device__ __forceinline__ bool take_chunk(unsigned long long* heads, uint64_t n, uint32_t lane, uint32_t& hc, uint32_t& tried,
                                           uint64_t& chunk_next, uint64_t& chunk_end) {
    for (;;) {
        unsigned long long k = 0;
        if (lane == 0) k = atomicAdd(heads + uint64_t(kHeadStride) * hc, 1ull);
        k = __shfl(k, 0);
        const uint64_t base = (uint64_t(hc) + uint64_t(kHeads) * k) * kChunk;
        if (base < n) {
            chunk_next = base;
            chunk_end = base + kChunk < n ? base + kChunk : n;
            return true;
        }
        if (++tried >= kHeads) return false;
        hc = hc + 1 == kHeads ? 0u : hc + 1;
    }
}
constexpr int kRefillMin = 8;     // zone kernels: idle lanes that trigger a refill
// Zone kernels: the FP64 renormalisation of H, L and Q runs on the steps that
// complete in every kRenormEvery-th iteration of the wave.  The schedule is the
// wave's, not the lane's, so the whole wave skips the block on the other
// iterations; a lane is renormalised about every kRenormEvery steps.  On the
// 1080p Kerr frame the integration takes 44.1 ms renormalising every step,
// 38.0 every 4th iteration, 37.1 every 8th and 36.3 every 16th, all at the same
// radiance error (6.33e-4).  At 16 the mean drift of H / p^2 (2.0e-7) equals
// the all-FP64 path's own (1.9e-7), and that of L and Q stays below it.
constexpr uint32_t kRenormEvery = 16;

// Persistent integrator.  T is the storage / state type: float for the FP32
// reference path, double for the mixed and FP64 paths.
// MESH: instantiation for scenes with triangle meshes (the mesh test stays out
// of the register allocation of the others).
template <int METRIC, int PREC, bool MESH>
__global__ void __launch_bounds__(256) k_trace(TraceParams P) {
    using S = typename std::conditional<PREC == MGK_PRECISION_FP32_REFERENCE, float, double>::type;
    const uint32_t lane = __lane_id();
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));

    St<S> y{};
    S h = S(0), contraction = S(0);
    Lane<S> ln;
    lane_constants<METRIC, S>(P, S(P.spin), ln);
    ln.cx = S(P.cam[0]); ln.cy = S(P.cam[1]); ln.cz = S(P.cam[2]);
    bool active = false;
    uint64_t ray = 0;
    uint32_t n_step = 0, n_sub = 0;
    uint64_t steps_acc = 0, trials_acc = 0, fp64_acc = 0;

    uint64_t chunk_next = 0, chunk_end = 0;   // wave-uniform
    uint32_t head_c = blockIdx.x % kHeads, head_tried = 0;   // wave-uniform: work counter in use, counters found empty
    bool exhausted = false;                   // wave-uniform

    for (;;) {
        // ---- refill idle lanes (wave-uniform control flow) ----
        uint64_t idle = __ballot(!active);
        while (idle != 0 && !exhausted) {
            if (chunk_next >= chunk_end && !take_chunk(P.head, P.n, lane, head_c, head_tried, chunk_next, chunk_end)) {
                exhausted = true;
                break;
            }
            uint32_t avail = uint32_t(chunk_end - chunk_next);
            uint32_t cnt = uint32_t(__popcll(idle));
            uint32_t take = cnt < avail ? cnt : avail;
            if (!active) {
                uint32_t rank = uint32_t(__popcll(idle & lt_mask));
                if (rank < take) {
                    uint64_t slot = chunk_next + rank;
                    ray = P.queue ? uint64_t(P.queue[slot]) : slot;
                    y.qx = load_c<S>(P.rays, P.stride, 0, ray);
                    y.qy = load_c<S>(P.rays, P.stride, 1, ray);
                    y.qz = load_c<S>(P.rays, P.stride, 2, ray);
                    y.qt = load_c<S>(P.rays, P.stride, 3, ray);
                    y.px = load_c<S>(P.rays, P.stride, 4, ray);
                    y.py = load_c<S>(P.rays, P.stride, 5, ray);
                    y.pz = load_c<S>(P.rays, P.stride, 6, ray);
                    y.pt = load_c<S>(P.rays, P.stride, 7, ray);
                    h = load_c<S>(P.rays, P.stride, 8, ray);
                    contraction = load_c<S>(P.rays, P.stride, 9, ray);
                    if (P.per_ray) {
                        lane_constants<METRIC, S>(P, METRIC == KERR ? load_c<S>(P.rays, P.stride, 10, ray) : S(0), ln);
                        const S L = S(P.length_scale);
                        ln.cx = y.qx * L; ln.cy = y.qy * L; ln.cz = y.qz * L;   // escape sphere about the ray's origin
                    }
                    active = true;
                    n_step = P.steps_in ? P.steps_in[ray] : 0u;
                    n_sub = 0;
                }
            }
            chunk_next += take;
            idle = __ballot(!active);
        }
        if (__ballot(active) == 0) break;
        if (!active) continue;

        // ---- one trial ----
        St<S> y1;
        const S h_step = h;          // parameter length of this trial's step
        const float dl = float(h);   // ... in float (Hermite chords)
        bool step_done;
        if constexpr (METRIC == MINKOWSKI) {
#pragma clang fp contract(off)
            // explicit Euler with the stored h (metric_minkowski.cuh:49-53), each
            // product and sum rounded on its own as the reference's vector
            // operators do
            y1 = y;
            y1.qx = y.qx + h * y.px; y1.qy = y.qy + h * y.py; y1.qz = y.qz + h * y.pz; y1.qt = y.qt + h * -y.pt;
            step_done = true;
            ++trials_acc;
        } else if constexpr (PREC == MGK_PRECISION_MIXED) {
            double rms;
            double r2 = y.qx * y.qx + y.qy * y.qy + y.qz * y.qz;
            // rho^2 = r^2 + a^2 sin^2(theta): zone_rho2 = zone_r^2 + a^2 covers r < zone_r.  The choice is
            // made per wave: if any active lane is in the zone, the whole wave runs the FP64 trial, so a
            // wave never executes both trial bodies.
            bool near = r2 < ln.zone_rho2;
            if (__ballot(near) != 0) {
                trial_split<METRIC, double, true>(y, h, float(ln.a), P.mass_sign, P, y1, rms);
                ++fp64_acc;
            } else {
                trial_split<METRIC, float, true>(y, h, float(ln.a), P.mass_sign, P, y1, rms);
            }
            ++trials_acc;
            double rs = rms > 1e-30 ? rms : 1e-30;
            double fac = double(__expf(-0.2f * __logf(float(rs))));
            double hn = 0.9 * h * fac;
            if (hn > double(P.h_max)) hn = double(P.h_max);
            h = hn;
            ++n_sub;
            step_done = (rs < 1.0) || (n_sub >= P.n_max);
        } else {
            S hn;
            bool acc;
            y1 = trial_reference<METRIC, S>(y, h, ln.a, S(P.mass_sign), S(P.abs_tol), S(P.rel_tol), S(P.h_max), hn, acc);
            ++trials_acc;
            if (PREC == MGK_PRECISION_FP64) ++fp64_acc;
            h = hn;
            ++n_sub;
            step_done = acc || (n_sub >= P.n_max);
        }
        if (!step_done) continue;

        // ---- step complete: renormalise, classify ----
        ++n_step;
        n_sub = 0;
        ++steps_acc;
        if constexpr (PREC == MGK_PRECISION_MIXED) renormalise_null<METRIC>(y1, double(ln.a), double(P.mass_sign));
        mgk_hit_t rec;
        S t_hit = S(0);
        S hq[3];
        // the disk crossing on the geodesic: a trial of the kernel's own arithmetic
        auto refine = [&](S t, S& r, S& x, S& z) -> bool {
#pragma clang fp contract(off)
            // (mixed precision runs the zone kernels; this kernel's mixed instance traces flat space only)
            if constexpr (METRIC == MINKOWSKI || PREC == MGK_PRECISION_MIXED) {
                return true;
            } else {
                S hn;
                bool acc;
                St<S> d;   // the fifth stage, evaluated at the trial's end (c5 = 1): the tangent there
                const St<S> yc = trial_reference<METRIC, S>(y, t * h_step, ln.a, S(P.mass_sign), S(P.abs_tol), S(P.rel_tol),
                                                            S(P.h_max), hn, acc, &d);
                return disk_point<S>(yc, d.qx, d.qy, d.qz, ln.a, ln.disk_r_in, S(P.disk_r_out), r, x, z);
            }
        };
        uint32_t cls = classify<METRIC, S, false, MESH>(y, y1, n_step, contraction, P, ln, dl, rec, t_hit, hq, refine);
        if (cls != MGK_HIT_NONE) {
            P.hits[ray] = rec;
            if (P.steps_out) P.steps_out[ray] = n_step;
            if (P.final_state) write_terminal(P, ray, cls, y, y1, t_hit, hq);
            active = false;
        } else {
            y = y1;
        }
    }

    // ---- statistics: wave reduction, one atomic per counter per wave ----
    for (int off = 32; off > 0; off >>= 1) {
        steps_acc += __shfl_xor(steps_acc, off);
        trials_acc += __shfl_xor(trials_acc, off);
        fp64_acc += __shfl_xor(fp64_acc, off);
    }
    if (lane == 0) {
        atomicAdd(&P.stats[0], (unsigned long long)steps_acc);
        atomicAdd(&P.stats[1], (unsigned long long)trials_acc);
        atomicAdd(&P.stats[2], (unsigned long long)fp64_acc);
    }
}

// ---------------------------------------------------------------------------
// Mixed precision, split by zone.  The far-zone kernel runs FP32 trials, the
// near-zone kernel FP64 trials; a ray whose completed step crosses the zone
// boundary is parked (state to HBM) and appended to the other kernel's queue,
// compacted per wave with ballot + mbcnt and one atomic per wave.  The host
// alternates the two until both queues are empty.  A ray's arithmetic thus
// depends on its own position only, never on the rays sharing its wave.
// ---------------------------------------------------------------------------
// Zone-kernel instantiations by scene feature: kZoneMesh (triangle meshes),
// kZoneDiskGeo (the disk crossing on the geodesic).  Each keeps its code out of
// the others' register allocation.
constexpr int kZoneMesh = 1, kZoneDiskGeo = 2;

template <int METRIC, bool NEAR, int FEAT>
__global__ void __launch_bounds__(256, NEAR ? 2 : 3) k_trace_zone(TraceParams P, ZoneParams Z) {
    using S = double;
    constexpr bool MESH = (FEAT & kZoneMesh) != 0, DISKGEO = (FEAT & kZoneDiskGeo) != 0;
    const uint32_t lane = __lane_id();
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));

    St<S> y{};
    S h = S(0), contraction = S(0);
    S L0 = S(0), Q0 = S(0);   // the ray's axial angular momentum and Carter constant
    float h_acc = 0.0f, lnerr_acc = 0.0f;   // last accepted trial: h and ln(error); h_acc = 0: none yet
    Lane<S> ln;
    lane_constants<METRIC, S>(P, S(P.spin), ln);
    ln.cx = S(P.cam[0]); ln.cy = S(P.cam[1]); ln.cz = S(P.cam[2]);
    bool active = false, handoff = false;
    uint64_t ray = 0;
    uint32_t n_step = 0, n_sub = 0;
    // the disk crossing on the geodesic (MGK_SCENE_DISK_ON_GEODESIC): the next
    // trial is the sub-step to the chord's crossing
    bool disk_sub = false;
    double h_resume = 0.0;        // the controller's step, resumed after the sub-step
    uint64_t steps_acc = 0, trials_acc = 0, fp64_acc = 0;
    uint64_t chunk_next = 0, chunk_end = 0;
    uint32_t head_c = blockIdx.x % kHeads, head_tried = 0;   // wave-uniform: work counter in use, counters found empty
    bool exhausted = false;
    uint64_t ob_base = 0;          // the wave's current block of out-queue slots
    uint32_t ob_used = kOutBlock;  // slots of it taken (kOutBlock: no block yet)
    // hysteresis: enter the near zone at rho^2 < zone, leave it at rho^2 > 1.21 zone
    auto in_zone = [&](const St<S>& q) {
        const S r2 = q.qx * q.qx + q.qy * q.qy + q.qz * q.qz;
        return NEAR ? !(r2 > 1.21 * ln.zone_rho2) : (r2 < ln.zone_rho2);
    };

    uint32_t wave_it = 0;          // iterations of the wave (uniform)
    for (;;) {
        ++wave_it;
        // ---- park the rays that crossed the zone boundary (wave-uniform) ----
        const uint64_t ho = __ballot(handoff);
        if (ho != 0) {
            // slots first from the wave's current block, the rest from a new one
            const uint32_t k = uint32_t(__popcll(ho));
            const uint32_t avail = kOutBlock - ob_used;
            unsigned long long nb = 0;
            if (k > avail) {
                const int leader = __ffsll((unsigned long long)ho) - 1;
                if (int(lane) == leader) nb = atomicAdd(Z.out_count, (unsigned long long)kOutBlock);
                nb = __shfl(nb, leader);
            }
            if (handoff) {
                // parked at the out-queue slot: the next pass reads it back contiguously
                const uint32_t rank = uint32_t(__popcll(ho & lt_mask));
                const uint64_t pos = rank < avail ? ob_base + ob_used + rank : uint64_t(nb) + (rank - avail);
                Z.out_queue[pos] = uint32_t(ray);
                double* st = Z.state_out;
                const uint64_t sd = Z.cap;
                st[0 * sd + pos] = y.qx; st[1 * sd + pos] = y.qy; st[2 * sd + pos] = y.qz; st[3 * sd + pos] = y.qt;
                st[4 * sd + pos] = y.px; st[5 * sd + pos] = y.py; st[6 * sd + pos] = y.pz; st[7 * sd + pos] = y.pt;
                st[8 * sd + pos] = h;
                st[9 * sd + pos] = contraction; st[10 * sd + pos] = L0; st[11 * sd + pos] = Q0;
                st[12 * sd + pos] = double(h_acc); st[13 * sd + pos] = double(lnerr_acc);
                Z.steps_out[pos] = n_step;
                handoff = false;
                active = false;
            }
            if (k > avail) {
                ob_base = nb;
                ob_used = k - avail;
            } else {
                ob_used += k;
            }
        }
        // ---- refill idle lanes, once at least kRefillMin are idle: the refill
        // code runs only on the idle lanes, so refilling one or two at a time
        // spends most of its issue slots on masked lanes ----
        uint64_t idle = __ballot(!active);
        if (__popcll(idle) < kRefillMin && idle != ~0ull) idle = 0;
        while (idle != 0 && !exhausted) {
            if (chunk_next >= chunk_end && !take_chunk(P.head, Z.n_in, lane, head_c, head_tried, chunk_next, chunk_end)) {
                exhausted = true;
                break;
            }
            const uint32_t avail = uint32_t(chunk_end - chunk_next);
            const uint32_t cnt = uint32_t(__popcll(idle));
            const uint32_t take = cnt < avail ? cnt : avail;
            if (!active) {
                const uint32_t rank = uint32_t(__popcll(idle & lt_mask));
                const uint64_t slot = chunk_next + rank;
                const uint32_t qv = (rank < take && Z.in_queue) ? Z.in_queue[slot] : 0u;
                if (rank < take && qv != kHole) {
                    ray = Z.in_queue ? uint64_t(qv) : slot;
                    // initial conditions by ray id; parked states by queue slot
                    const double* src = Z.from_ic ? reinterpret_cast<const double*>(P.rays) : Z.state_in;
                    const uint64_t sd = Z.from_ic ? P.stride : Z.cap, at = Z.from_ic ? ray : slot;
                    y.qx = src[0 * sd + at]; y.qy = src[1 * sd + at]; y.qz = src[2 * sd + at];
                    y.qt = src[3 * sd + at]; y.px = src[4 * sd + at]; y.py = src[5 * sd + at];
                    y.pz = src[6 * sd + at]; y.pt = src[7 * sd + at]; h = src[8 * sd + at];
                    n_step = Z.from_ic ? (P.steps_in ? P.steps_in[ray] : 0u) : Z.steps_in[slot];
                    contraction = src[9 * sd + at];
                    if (P.per_ray) {
                        lane_constants<METRIC, S>(P, METRIC == KERR ? load_c<S>(P.rays, P.stride, 10, ray) : S(0), ln);
                        const S L = S(P.length_scale);
                        const double* ic = reinterpret_cast<const double*>(P.rays);
                        ln.cx = ic[0 * P.stride + ray] * L; ln.cy = ic[1 * P.stride + ray] * L; ln.cz = ic[2 * P.stride + ray] * L;
                    }
                    if (Z.from_ic) {
                        // constants of motion of the initial conditions (NaN: not defined on the spin axis)
                        Invariants v;
                        const bool ok = METRIC != MINKOWSKI && invariants<METRIC>(y, double(ln.a), v);
                        L0 = ok ? v.L : S(__builtin_nan(""));
                        Q0 = ok ? v.Q : S(__builtin_nan(""));
                        h_acc = 0.0f;
                        lnerr_acc = 0.0f;
                    } else {
                        L0 = src[10 * sd + at];
                        Q0 = src[11 * sd + at];
                        h_acc = float(src[12 * sd + at]);
                        lnerr_acc = float(src[13 * sd + at]);
                    }
                    n_sub = 0;
                    active = true;
                    handoff = !Z.no_handoff && (in_zone(y) != NEAR);
                }
            }
            chunk_next += take;
            idle = __ballot(!active);
        }
        if (__ballot(active) == 0) break;
        if (!active || handoff) continue;

        // ---- one trial ----
        using SS = typename std::conditional<NEAR, double, float>::type;
        St<S> y1;
        double rms;
        const double h_step = h;     // parameter length of this trial's step
        const float dl = float(h);   // ... in float (Hermite chords)
        SS tan5[3];                  // dx/dlambda of the fifth stage (at the trial's end)
        trial_split<METRIC, SS, true>(y, h, float(ln.a), P.mass_sign, P, y1, rms, DISKGEO ? tan5 : nullptr);
        if constexpr (NEAR) ++fp64_acc;
        ++trials_acc;
        if (DISKGEO && disk_sub) {
            // the sub-step to the chord's disk crossing, taken without step control,
            // then one Newton step along its end tangent to the plane: on the disk the
            // path ends there; off it, the path goes on from the sub-step's end
            disk_sub = false;
            double r, x, z;
            if (disk_point<double>(y1, double(tan5[0]), double(tan5[1]), double(tan5[2]), double(ln.a), double(ln.disk_r_in),
                                   double(P.disk_r_out), r, x, z)) {
                mgk_hit_t rec;
                S hq[3];
                hit_record<METRIC, S>(y1, y1, MGK_HIT_DISK, S(0), r, contraction, P, ln.a, 0, hq, 0u, 0u, rec);
                rec.a1 = azimuth<S>(z, x);
                P.hits[ray] = rec;
                if (P.steps_out) P.steps_out[ray] = n_step;
                if (P.final_state) write_final(P, ray, y1);
                active = false;
            } else {
                y = y1;
                h = h_resume;
            }
            continue;
        }
        const double rs = rms > 1e-30 ? rms : 1e-30;
        // Step factor: the reference's 0.9 err^-1/5 (internal_math.h:2190-2195).
        // After an accepted trial it is also bounded by Gustafsson's
        // predictive factor 0.9 (h / h_acc) (err_acc / err^2)^1/5 (Gustafsson,
        // ACM TOMS 20 (1994); Hairer & Wanner, Solving ODEs II, IV.8), which
        // extrapolates the growth of the error constant from the last two
        // accepted trials.  A ray falling toward the hole meets an error
        // constant that grows several-fold per step; the plain factor then
        // over-proposes and every step is rejected once.  The factor is
        // limited to [0.2, 5], and to 1 right after a rejection (Hairer,
        // Norsett & Wanner, Solving ODEs I, II.4).  The FP32-reference kernel
        // keeps the reference's controller.
        const float lr = __logf(float(rs));
        const bool accept = (rs < 1.0) || (n_sub + 1 >= P.n_max);
        float fac = 0.9f * __expf(-0.2f * lr);
        if (accept && h_acc > 0.0f) fac = fminf(fac, 0.9f * (float(h) / h_acc) * __expf(0.2f * lnerr_acc - 0.4f * lr));
        fac = fmaxf(fminf(fac, n_sub > 0 ? 1.0f : 5.0f), 0.2f);
        if (accept) {
            h_acc = float(h);
            lnerr_acc = fmaxf(lr, -4.6051702f);   // ln 1e-2: a tiny error says little about the trend
        }
        double hn = h * double(fac);
        if (hn > double(P.h_max)) hn = double(P.h_max);
        h = hn;
        ++n_sub;
        if (!accept) continue;

        // ---- step complete ----
        ++n_step;
        n_sub = 0;
        ++steps_acc;
        if (wave_it % kRenormEvery == 0u) {
            if (isnan(Q0)) renormalise_null<METRIC>(y1, double(ln.a), double(P.mass_sign));
            else renormalise_constants<METRIC>(y1, double(ln.a), double(P.mass_sign), L0, Q0);
        }
        mgk_hit_t rec;
        S t_hit = S(0);
        S hq[3];
        // the disk crossing on the geodesic: classify only notes the chord's
        // crossing; the sub-step to it runs through this loop's own trial
        double disk_req = -1.0;
        auto note = [&](S t, S&, S&, S&) -> bool {
            disk_req = t;
            return false;
        };
        uint32_t cls;
        if constexpr (DISKGEO) cls = classify<METRIC, S, true, MESH>(y, y1, n_step, contraction, P, ln, dl, rec, t_hit, hq, note);
        else cls = classify<METRIC, S, true, MESH>(y, y1, n_step, contraction, P, ln, dl, rec, t_hit, hq);
        if (DISKGEO && disk_req >= 0.0 && (cls == MGK_HIT_NONE || t_hit > disk_req)) {
            h_resume = h;
            h = disk_req * h_step;
            disk_sub = true;          // y stays: the sub-step starts at the step's start
        } else if (cls != MGK_HIT_NONE) {
            P.hits[ray] = rec;
            if (P.steps_out) P.steps_out[ray] = n_step;
            if (P.final_state) write_terminal(P, ray, cls, y, y1, t_hit, hq);
            active = false;
        } else {
            y = y1;
            handoff = !Z.no_handoff && (in_zone(y) != NEAR);
        }
    }

    // the unused tail of the wave's last block: holes for the next pass
    for (uint32_t i = ob_used + lane; i < kOutBlock; i += 64) Z.out_queue[ob_base + i] = kHole;

    for (int off = 32; off > 0; off >>= 1) {
        steps_acc += __shfl_xor(steps_acc, off);
        trials_acc += __shfl_xor(trials_acc, off);
        fp64_acc += __shfl_xor(fp64_acc, off);
    }
    if (lane == 0) {
        atomicAdd(&P.stats[0], (unsigned long long)steps_acc);
        atomicAdd(&P.stats[1], (unsigned long long)trials_acc);
        atomicAdd(&P.stats[2], (unsigned long long)fp64_acc);
    }
}

// ---------------------------------------------------------------------------
// Ray generation + initial conditions.
// ---------------------------------------------------------------------------
template <int METRIC, class S>
__device__ __forceinline__ void gen_one(const RayGenParams& P, uint64_t i, uint64_t o) {
#pragma clang fp contract(off)
    const uint32_t npix = P.W * P.H;
    const uint32_t pixel = uint32_t(i % npix);
    const uint32_t sample = P.sample_begin + uint32_t(i / npix);
    const uint32_t px = pixel % P.W, py = pixel / P.W;
    float u[4];
    path_uniforms(P.seed, P.camera_id, pixel, sample, 0, u);
    // sensor point (device_sensor_utilities.cuh:35-80)
    S sx, sy;
    sensor_point<S>(P.W, P.H, P.sensor_width_mm, px, py, u[0], u[1], sx, sy);
    // pinhole (thin_lens.cuh:60-76): origin at the lens point, direction lens - sensor
    S f_m = S(P.focal_mm) / S(1000);
    S dx = S(0) - sx, dy = S(0) - sy, dz = -f_m - S(0);
    S dl = sqrt_<false>(dx * dx + dy * dy + dz * dz);
    dx = dx / dl; dy = dy / dl; dz = dz / dl;
    const float* m = P.orientation;
    S ox = (S(m[0]) * S(0) + S(m[1]) * S(0) + S(m[2]) * -f_m) + S(P.position[0]);
    S oy = (S(m[3]) * S(0) + S(m[4]) * S(0) + S(m[5]) * -f_m) + S(P.position[1]);
    S oz = (S(m[6]) * S(0) + S(m[7]) * S(0) + S(m[8]) * -f_m) + S(P.position[2]);
    // initial conditions (kernel_generate_initial_conditions.cu:54-74)
    const S L = S(P.length_scale), a = S(P.spin), ms = S(P.mass_sign);
    S qx = ox / L, qy = oy / L, qz = oz / L, qt = S(1) / (L / S(kC));
    Sym4<S> g = metric_g<METRIC, S>(qx, qy, qz, a, ms, false);
    V4<S> uref = reference_observer<METRIC, S>(qx, qy, qz, a, ms);
    S wx = S(P.observer_velocity[0]) / S(kC), wy = S(P.observer_velocity[1]) / S(kC), wz = S(P.observer_velocity[2]) / S(kC);
    if (P.time_reversed) { wx = -wx; wy = -wy; wz = -wz; }
    V4<S> X{S(1), S(0), S(0), S(0)}, Y{S(0), S(1), S(0), S(0)}, Z{S(0), S(0), S(1), S(0)};
    V4<S> uobs = tetrad_apply(g, X, Y, Z, uref, V4<S>{wx, wy, wz, sqrt_<false>((wx * wx + wy * wy + wz * wz) - S(-1))});
    V4<S> CX{S(m[0]), S(m[3]), S(m[6]), S(0)}, CY{S(m[1]), S(m[4]), S(m[7]), S(0)}, CZ{S(m[2]), S(m[5]), S(m[8]), S(0)};
    S lx = dx / L, ly = dy / L, lz = dz / L;
    V4<S> pup = tetrad_apply(g, CX, CY, CZ, uobs, V4<S>{lx, ly, lz, sqrt_<false>((lx * lx + ly * ly + lz * lz) - S(0))});
    V4<S> pm = lower(g, pup);
    S* out = reinterpret_cast<S*>(P.rays);
    const uint64_t st = P.stride;
    out[0 * st + o] = qx; out[1 * st + o] = qy; out[2 * st + o] = qz; out[3 * st + o] = qt;
    out[4 * st + o] = pm.x; out[5 * st + o] = pm.y; out[6 * st + o] = pm.z; out[7 * st + o] = pm.t;
    out[8 * st + o] = METRIC == MINKOWSKI ? S(1000) : S(1);
    out[9 * st + o] = dot4(uobs, pm);
    if (P.per_ray_spin) out[10 * st + o] = a;   // the spin of the equations, per ray (batched cameras)
}

// FP32 reference arithmetic: every ray builds its own initial conditions in
// the reference's order.
template <int METRIC> __global__ void __launch_bounds__(256) k_generate_rays_ref(RayGenParams P) {
    const uint64_t n = uint64_t(P.W) * P.H * P.spp;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
        gen_one<METRIC, float>(P, i, i);
}

// FP64: the per-camera part of gen_one, once per camera.  tetrad_apply is
// linear in its local components, so the camera tetrad lowered with g is the
// matrix M of p_mu = M l, and u_obs . p = (u_obs^T M) l.
template <int METRIC> __device__ void ic_camera(const RayGenParams& P, IcCamera& C) {
#pragma clang fp contract(off)
    using S = double;
    const S f_m = S(P.focal_mm) / S(1000);
    const float* m = P.orientation;
    const S ox = (S(m[0]) * S(0) + S(m[1]) * S(0) + S(m[2]) * -f_m) + S(P.position[0]);
    const S oy = (S(m[3]) * S(0) + S(m[4]) * S(0) + S(m[5]) * -f_m) + S(P.position[1]);
    const S oz = (S(m[6]) * S(0) + S(m[7]) * S(0) + S(m[8]) * -f_m) + S(P.position[2]);
    const S L = S(P.length_scale), a = S(P.spin), ms = S(P.mass_sign);
    const S qx = ox / L, qy = oy / L, qz = oz / L;
    C.q[0] = qx; C.q[1] = qy; C.q[2] = qz; C.q[3] = S(1) / (L / S(kC));
    const Sym4<S> g = metric_g<METRIC, S>(qx, qy, qz, a, ms, false);
    const V4<S> uref = reference_observer<METRIC, S>(qx, qy, qz, a, ms);
    S wx = S(P.observer_velocity[0]) / S(kC), wy = S(P.observer_velocity[1]) / S(kC), wz = S(P.observer_velocity[2]) / S(kC);
    if (P.time_reversed) { wx = -wx; wy = -wy; wz = -wz; }
    const V4<S> X{S(1), S(0), S(0), S(0)}, Y{S(0), S(1), S(0), S(0)}, Z{S(0), S(0), S(1), S(0)};
    const V4<S> uobs = tetrad_apply(g, X, Y, Z, uref, V4<S>{wx, wy, wz, sqrt_<false>((wx * wx + wy * wy + wz * wz) - S(-1))});
    const V4<S> CX{S(m[0]), S(m[3]), S(m[6]), S(0)}, CY{S(m[1]), S(m[4]), S(m[7]), S(0)}, CZ{S(m[2]), S(m[5]), S(m[8]), S(0)};
    for (int k = 0; k < 4; ++k) {
        const V4<S> e{S(k == 0), S(k == 1), S(k == 2), S(k == 3)};
        const V4<S> pm = lower(g, tetrad_apply(g, CX, CY, CZ, uobs, e));
        C.M[0][k] = pm.x; C.M[1][k] = pm.y; C.M[2][k] = pm.z; C.M[3][k] = pm.t;
        C.w[k] = dot4(uobs, pm);
    }
}

__device__ void ic_camera_any(const RayGenParams& P, IcCamera& C) {
    if (P.metric == KERR) ic_camera<KERR>(P, C);
    else if (P.metric == SCHWARZSCHILD) ic_camera<SCHWARZSCHILD>(P, C);
    else ic_camera<MINKOWSKI>(P, C);
}

__global__ void k_ic_camera(RayGenParams P, IcCamera* out) {
    if (blockIdx.x == 0 && threadIdx.x == 0) ic_camera_any(P, *out);
}

__global__ void __launch_bounds__(64) k_ic_cameras(const RayGenParams* cams, uint32_t n_cam, IcCamera* out) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n_cam) ic_camera_any(cams[c], out[c]);
}

// The per-ray part (ray_ic_fp64, mgk_raygen.hpp), stored SoA.
__device__ __forceinline__ void gen_one_fp64(const RayGenParams& P, const IcCamera& C, uint64_t i, uint64_t o) {
    double y[10];
    ray_ic_fp64(P, C, i, y);
    double* out = reinterpret_cast<double*>(P.rays);
    const uint64_t st = P.stride;
#pragma unroll
    for (int k = 0; k < 10; ++k) out[k * st + o] = y[k];
    if (P.per_ray_spin) out[10 * st + o] = double(P.spin);   // the spin of the equations, per ray (batched cameras)
}

__global__ void __launch_bounds__(256) k_generate_rays_fp64(RayGenParams P, const IcCamera* __restrict__ ic) {
    const uint64_t n = uint64_t(P.W) * P.H * P.spp;
    const IcCamera C = *ic;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
        gen_one_fp64(P, C, i, i);
}

// Batched cameras: ray i belongs to camera c with offsets[c] <= i < offsets[c + 1]
// (at most kMaxBatchCameras); each camera's rays are laid out as a single
// camera's, sample-major, with the camera's own spin in component 10.
__global__ void __launch_bounds__(256) k_generate_rays_batch(const RayGenParams* cams, const uint64_t* offsets, uint32_t n_cam,
                                                             uint64_t n, const IcCamera* ic) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x) {
        uint32_t lo = 0, hi = n_cam;   // last c with offsets[c] <= i
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (offsets[mid] <= i) lo = mid; else hi = mid;
        }
        gen_one_fp64(cams[lo], ic[lo], i - offsets[lo], i);
    }
}

// ---------------------------------------------------------------------------
// Shading: one thread per pixel.
// ---------------------------------------------------------------------------
constexpr int kShadeBlock = 256;

__device__ __forceinline__ float star_lookup(const ShadeParams& P, float c, float ph) {
    float sn = sqrtf(fmaxf(0.0f, 1.0f - c * c));
    float dx = sn * cosf(ph), dz = sn * sinf(ph);
    uint32_t ic = min(P.grid_nc - 1, uint32_t(fmaxf(0.0f, (c + 1.0f) * 0.5f * float(P.grid_nc))));
    uint32_t ip = min(P.grid_np - 1, uint32_t(fmaxf(0.0f, (ph + kPi) * (0.5f / kPi) * float(P.grid_np))));
    uint32_t cell = ic * P.grid_np + ip;
    float acc = 0.0f;
    for (uint32_t k = P.cell_start[cell]; k < P.cell_start[cell + 1]; ++k) {
        float4 s = P.stars[P.cell_items[k]];
        float ex = s.x - dx, ey = s.y - c, ez = s.z - dz;   // chord: |s - d|^2 <= (2 sin(rho/2))^2
        if (ex * ex + ey * ey + ez * ez <= P.chord2_star_r) acc += s.w;
    }
    return acc;
}

// Spectral response of the bins to one unit of irradiance observed at lambda,
//   R(lambda) = sum_i W_i (Phi(e_{i+1} - lambda) - Phi(e_i - lambda)),
// with bin edges e_i, the bins' CMF weights W_i and the normal CDF Phi of
// width sigma: the Gaussian redistribution of kernel_redistribute_radiance.cu:55-92
// followed by the splat of kernel_splat_to_renderbuffer.cu:37-57.  The splat is
// linear in the bins, so a path adds E R(lambda) to its pixel and no per-pixel
// bin array is needed.  Tabulated in FP64 at kRespN points over the hero range;
// linear interpolation between them errs by about (step^2 / 8) |R''|, near
// 3e-6 of R for 0.05 nm steps and sigma = 10 nm.
__global__ void __launch_bounds__(256) k_shade_response(ShadeParams P, float4* out) {
    const int j = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (j >= kRespN) return;
    const double lam = double(P.hero_min) + double(P.hero_span) * double(j) / double(kRespN - 1);
    const double k = 1.0 / (double(P.sigma) * 1.4142135623730951);
    double X = 0.0, Y = 0.0, Z = 0.0;
    double prev = 0.5 * (1.0 + erf((double(P.bin_min) - lam) * k));
    for (uint32_t i = 0; i < P.n_bin; ++i) {
        const double e1 = double(P.bin_min) + double(i + 1) * double(P.bin_width);
        const double cur = 0.5 * (1.0 + erf((e1 - lam) * k));
        const double d = cur - prev;
        X += d * double(P.weights[3 * i]);
        Y += d * double(P.weights[3 * i + 1]);
        Z += d * double(P.weights[3 * i + 2]);
        prev = cur;
    }
    out[j] = make_float4(float(X), float(Y), float(Z), 0.0f);
}

// One thread per pixel: the pixel's spp paths, n_hero wavelengths each.
__global__ void __launch_bounds__(kShadeBlock) k_shade(ShadeParams P) {
    const uint32_t npix = P.W * P.H;
    const uint32_t pixel = blockIdx.x * kShadeBlock + threadIdx.x;
    if (pixel >= npix) return;
    const float resp_scale = P.hero_span > 0.0f ? float(kRespN - 1) / P.hero_span : 0.0f;
    float X = 0.0f, Yv = 0.0f, Zv = 0.0f;
    for (uint32_t s = 0; s < P.spp; ++s) {
        mgk_hit_t h = P.hits[uint64_t(s) * npix + pixel];
        if (h.cls != MGK_HIT_DISK && h.cls != MGK_HIT_SPHERE && h.cls != MGK_HIT_ESCAPE && h.cls != MGK_HIT_BOX_EMITTER &&
            h.cls != MGK_HIT_MESH_EMITTER)
            continue;
        float u[4];
        path_uniforms(P.seed, P.camera_id, pixel, P.sample_begin + s, 0, u);
        const float g = h.g, g5 = g * g * g * g * g;
        float emit_scale = 0.0f, T = 0.0f, star = 0.0f;
        if (h.cls == MGK_HIT_DISK) {
            T = P.disk_T_in * powf(h.a0 / P.disk_r_in, -0.75f);
            emit_scale = 1.0f;
        } else if (h.cls == MGK_HIT_SPHERE) {
            float th = acosf(fmaxf(-1.0f, fminf(1.0f, h.a0)));
            float ph = h.a1 + kPi;
            uint32_t n = P.sphere_checker;
            uint32_t ci = min(n - 1, uint32_t(th / kPi * float(n)));
            uint32_t cj = min(n - 1, uint32_t(ph / (2.0f * kPi) * float(n)));
            T = P.sphere_T;
            emit_scale = ((ci + cj) & 1u) ? 0.0f : 1.0f;
        } else {
            if (P.flags & MGK_SCENE_STARS) star = star_lookup(P, h.a0, h.a1);
        }
        for (uint32_t k = 0; k < P.n_hero; ++k) {
            float f = u[2] + float(k) / float(P.n_hero);
            f -= floorf(f);
            float lam, thr = 0.0f;
            if (P.lambda_cmf) {
                // the primary sample through the inverse CDF; throughput
                // 1 / (n_hero pdf), pdf = 1 / (255 (c[k + 1] - c[k]))
                // (device_sensor_utilities.cuh:25-33)
                const float x = f * float(kCmfCdfN - 1);
                const uint32_t kk = uint32_t(fminf(x, float(kCmfCdfN - 2)));
                const float fr = x - float(kk), c0 = P.inv_cdf[kk], c1 = P.inv_cdf[kk + 1];
                lam = c0 * (1.0f - fr) + c1 * fr;
                thr = float(kCmfCdfN - 1) * (c1 - c0) / float(P.n_hero);
            } else {
                lam = P.hero_min + P.hero_span * f;
            }
            const float le = lam * g;
            float L;
            if (h.cls == MGK_HIT_ESCAPE) {
                L = 0.0f;
                if (P.flags & MGK_SCENE_BACKGROUND) L += P.background_scale * planck(le, P.background_T);
                if (star > 0.0f) L += star * planck(le, P.star_T);
            } else if (h.cls == MGK_HIT_BOX_EMITTER || h.cls == MGK_HIT_MESH_EMITTER) {
                L = P.led_scale * white_led(le);
            } else {
                L = emit_scale > 0.0f ? planck(le, T) : 0.0f;
            }
            L *= g5;
            if (isnan(L) || isinf(L) || L < 0.0f) L = 0.0f;
            if (!(L > 0.0f)) continue;
            const float E = P.lambda_cmf ? L * (P.irr_base * thr) : L * P.irr_scale;
            const float t = fmaxf(0.0f, (lam - P.hero_min) * resp_scale);
            const int i = min(int(t), kRespN - 2);
            const float w = t - float(i);
            const float4 r0 = P.resp[i], r1 = P.resp[i + 1];
            X += E * (r0.x + w * (r1.x - r0.x));
            Yv += E * (r0.y + w * (r1.y - r0.y));
            Zv += E * (r0.z + w * (r1.z - r0.z));
        }
    }
    const float inv = 1.0f / float(P.spp);
    P.xyz[3 * uint64_t(pixel) + 0] = X * inv;
    P.xyz[3 * uint64_t(pixel) + 1] = Yv * inv;
    P.xyz[3 * uint64_t(pixel) + 2] = Zv * inv;
}

// ---------------------------------------------------------------------------
// Reflection off diffuse box matter (kernel_bxdf.cu:69-95 with
// bxdf_utilities.cuh:56-151, matter_examples.cuh:36-60, brdf.cuh:12-29).
// ---------------------------------------------------------------------------
// Self-intersection offset of one float coordinate along n (Waechter & Binder,
// Ray Tracing Gems ch. 6, as bxdf_utilities.cuh:112-134 uses it): int(1024 n)
// units in the last place, or 2^-14 n within 1/32 of the origin.
__device__ __forceinline__ float offset_coord(float x, float n) {
    const int of = int(1024.0f * n);
    const float p = __int_as_float(__float_as_int(x) + ((x < 0.0f) ? -of : of));
    return fabsf(x) < (1.0f / 32.0f) ? x + (4.0f / 65536.0f) * n : p;
}

// Spatial legs of the Gram-Schmidt tetrad of tetrad_apply (order z, x, y).
// Each leg is normalised by one reciprocal square root and three multiplies:
// FP64 division is a multi-instruction sequence, and the legs only need to be
// unit to rounding.
template <class S>
__device__ __forceinline__ void tetrad_legs(const Sym4<S>& g, V4<S> u, V4<S>& ex, V4<S>& ey, V4<S>& ez) {
#pragma clang fp contract(off)
    const V4<S> X{S(1), S(0), S(0), S(0)}, Y{S(0), S(1), S(0), S(0)}, Z{S(0), S(0), S(1), S(0)};
    ez = axpy(inner(g, Z, u), u, Z);
    ez = scale(ez, S(1) / sqrt_<false>(inner(g, ez, ez)));
    ex = axpy(-inner(g, X, ez), ez, axpy(inner(g, X, u), u, X));
    ex = scale(ex, S(1) / sqrt_<false>(inner(g, ex, ex)));
    ey = axpy(-inner(g, Y, ex), ex, axpy(-inner(g, Y, ez), ez, axpy(inner(g, Y, u), u, Y)));
    ey = scale(ey, S(1) / sqrt_<false>(inner(g, ey, ey)));
}

template <class S> __device__ __forceinline__ void norm3(S& x, S& y, S& z) {
#pragma clang fp contract(off)
    const S il = S(1) / sqrt_<false>(x * x + y * y + z * z);
    x = x * il; y = y * il; z = z * il;
}

// One thread per traced ray; rays that ended on diffuse matter get new
// initial conditions (in place, components 0-7 and 9; the step size of
// component 8 is kept) and are appended to the out-queue, one atomic per wave.
// The matter is at rest: its four-velocity is the reference observer's.  The
// incoming direction and the face normal are taken into that observer's
// frame; faces with an x normal reflect specularly, the others send
// (p + n) / |p + n| for p uniform on the sphere, cosine-weighted about the
// facing normal n.  The new u.p carries the redshift of the path so far.
// Threads per block of k_reflect: the reflected rays of a block take their
// out-queue slots with one atomic per block (a single counter taken once per
// wave had serialised ~0.5 M atomics per launch on one address).
constexpr int kReflectBlock = 512;

template <int METRIC, class Sst, class S> __global__ void __launch_bounds__(kReflectBlock) k_reflect(ReflectParams P) {
#pragma clang fp contract(off)
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const uint32_t lane = __lane_id();
    uint64_t ray = 0;
    bool go = false;
    if (i < P.n_in) {
        ray = P.queue_in ? uint64_t(P.queue_in[i]) : i;
        const uint32_t c = P.hits[ray].cls;
        go = c == MGK_HIT_BOX_DIFFUSE || c == MGK_HIT_MESH_DIFFUSE;
    }
    if (go) {
        const uint64_t sd = P.stride;
        const double qxd = P.state[0 * sd + ray], qyd = P.state[1 * sd + ray], qzd = P.state[2 * sd + ray];
        const S qx = S(qxd), qy = S(qyd), qz = S(qzd);
        const V4<S> pin{S(P.state[4 * sd + ray]), S(P.state[5 * sd + ray]), S(P.state[6 * sd + ray]), S(P.state[7 * sd + ray])};
        Sst* rays = reinterpret_cast<Sst*>(P.rays);
        const S contraction = S(rays[9 * sd + ray]);
        const S a = S(P.spin), ms = S(P.mass_sign);
        const Sym4<S> g = metric_g<METRIC, S>(qx, qy, qz, a, ms, false);
        const Sym4<S> gi = metric_g<METRIC, S>(qx, qy, qz, a, ms, true);
        const V4<S> u = reference_observer<METRIC, S>(qx, qy, qz, a, ms);
        V4<S> ex, ey, ez;
        tetrad_legs(g, u, ex, ey, ez);
        // incoming direction in the matter's frame (bxdf_utilities.cuh:92-96)
        const V4<S> pup = lower(gi, pin);
        S dx = inner(g, ex, pup), dy = inner(g, ey, pup), dz = inner(g, ez, pup);
        norm3(dx, dy, dz);
        S z = contraction / dot4(u, pin);
        if (isnan(z) || isinf(z)) z = S(1e-10f);
        // geometric normal into the same frame, turned against the incoming direction (:101-104).
        // Boxes: the face's axis from the hit's code, against the chord already.  Meshes: the
        // triangle's normal through the instance's normal map, turned against the incoming
        // coordinate velocity; mesh matter is Lambertian on every face.
        const mgk_hit_t hr = P.hits[ray];
        int axis;
        V4<S> ng;
        if (hr.cls == MGK_HIT_MESH_DIFFUSE) {
            const float4* I = P.mesh_inst + uint64_t(kInstF4) * uint32_t(hr.a1);
            const float4 n3 = I[3], n4 = I[4], n5 = I[5];
            const float4 no = P.mesh_normals[__float_as_uint(n4.w) + uint32_t(hr.a0)];
            S wx = S(n3.x) * S(no.x) + S(n3.y) * S(no.y) + S(n3.z) * S(no.z);
            S wy = S(n4.x) * S(no.x) + S(n4.y) * S(no.y) + S(n4.z) * S(no.z);
            S wz = S(n5.x) * S(no.x) + S(n5.y) * S(no.y) + S(n5.z) * S(no.z);
            norm3(wx, wy, wz);
            if (wx * pup.x + wy * pup.y + wz * pup.z > S(0)) { wx = -wx; wy = -wy; wz = -wz; }
            ng = V4<S>{wx, wy, wz, S(0)};
            axis = -1;
        } else {
            const int code = int(hr.a0);
            axis = code % 3;
            const S ns = code >= 3 ? S(-1) : S(1);
            ng = V4<S>{axis == 0 ? ns : S(0), axis == 1 ? ns : S(0), axis == 2 ? ns : S(0), S(0)};
        }
        const V4<S> nup = lower(gi, ng);
        S nx = inner(g, ex, nup), ny = inner(g, ey, nup), nz = inner(g, ez, nup);
        norm3(nx, ny, nz);
        if (nx * dx + ny * dy + nz * dz > S(0)) { nx = -nx; ny = -ny; nz = -nz; }
        S ox, oy, oz;
        if (axis == 0) {
            const S k = S(2) * (dx * nx + dy * ny + dz * nz);
            ox = dx - nx * k; oy = dy - ny * k; oz = dz - nz * k;
        } else {
            float uu[4];
            const uint32_t pixel = uint32_t(ray % P.npix), sample = P.sample_begin + uint32_t(ray / P.npix);
            path_uniforms(P.seed, P.camera_id, pixel, sample, P.block, uu);
            // p uniform on the sphere: cos(theta) = 1 - 2 u0, sin(theta) = 2 sqrt(u0 (1 - u0))
            const S u0 = S(uu[0]), ct = S(1) - S(2) * u0, st = S(2) * sqrt_<false>(u0 * (S(1) - u0));
            S sp, cp;
            if constexpr (sizeof(S) == 4) sincosf(2.0f * kPi * uu[1], &sp, &cp);
            else sincos(S(2) * S(kPi) * S(uu[1]), &sp, &cp);
            ox = cp * st + nx; oy = sp * st + ny; oz = ct + nz;
        }
        norm3(ox, oy, oz);
        // outgoing momentum from the frame (relativity_utilities.cuh:186-195)
        const S lt = sqrt_<false>((ox * ox + oy * oy + oz * oz) - S(0));
        V4<S> vup = scale(u, lt);
        vup = axpy(ox, ex, vup);
        vup = axpy(oy, ey, vup);
        vup = axpy(oz, ez, vup);
        const V4<S> pout = lower(g, vup);
        // offset along the global facing normal (the box normal already opposes the chord)
        // the position keeps the storage precision: the float offset is added to the stored hit point
        const float fx = float(qxd), fy = float(qyd), fz = float(qzd);
        rays[0 * sd + ray] = Sst(qxd) + Sst(offset_coord(fx, float(ng.x)) - fx);
        rays[1 * sd + ray] = Sst(qyd) + Sst(offset_coord(fy, float(ng.y)) - fy);
        rays[2 * sd + ray] = Sst(qzd) + Sst(offset_coord(fz, float(ng.z)) - fz);
        rays[3 * sd + ray] = Sst(P.state[3 * sd + ray]);
        rays[4 * sd + ray] = Sst(pout.x); rays[5 * sd + ray] = Sst(pout.y); rays[6 * sd + ray] = Sst(pout.z);
        rays[7 * sd + ray] = Sst(pout.t);
        rays[9 * sd + ray] = Sst(dot4(u, pout) * z);
    }
    // append the reflected rays: per-wave counts in LDS, one atomic per block
    __shared__ uint32_t wave_count[kReflectBlock / 64];
    __shared__ unsigned long long block_base;
    const uint64_t m = __ballot(go);
    const uint32_t wave = threadIdx.x / 64u;
    if (lane == 0) wave_count[wave] = uint32_t(__popcll(m));
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t total = 0;
        for (uint32_t w = 0; w < kReflectBlock / 64; ++w) total += wave_count[w];
        block_base = total ? atomicAdd(P.count, (unsigned long long)total) : 0ull;
    }
    __syncthreads();
    if (go) {
        uint64_t off = block_base;
        for (uint32_t w = 0; w < wave; ++w) off += wave_count[w];
        const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
        P.queue_out[off + uint64_t(__popcll(m & lt_mask))] = uint32_t(ray);
    }
}

// precision: MGK_PRECISION_* of the trace.  The frame of the reflection is
// computed in FP32 on the FP32-reference and mixed paths (the reference's own
// bxdf arithmetic is FP32, and the walls lie far from the FP64 zone), in FP64
// on the FP64 path; initial conditions keep the ray buffer's precision.
hipError_t launch_reflect(const ReflectParams& p, int precision, hipStream_t s) {
    if (p.n_in == 0) return hipSuccess;
    const uint64_t blocks = (p.n_in + kReflectBlock - 1) / kReflectBlock;
    if (blocks > 0x7FFFFFFFull) return hipErrorInvalidValue;
    const dim3 g{unsigned(blocks)}, b{unsigned(kReflectBlock)};
    auto go = [&](auto m) -> hipError_t {
        constexpr int M = decltype(m)::value;
        switch (precision) {
            case MGK_PRECISION_FP32_REFERENCE: hipLaunchKernelGGL((k_reflect<M, float, float>), g, b, 0, s, p); break;
            case MGK_PRECISION_MIXED: hipLaunchKernelGGL((k_reflect<M, double, float>), g, b, 0, s, p); break;
            default: hipLaunchKernelGGL((k_reflect<M, double, double>), g, b, 0, s, p); break;
        }
        return hipGetLastError();
    };
    switch (p.metric) {
        case KERR: return go(std::integral_constant<int, KERR>{});
        case SCHWARZSCHILD: return go(std::integral_constant<int, SCHWARZSCHILD>{});
        case MINKOWSKI: return go(std::integral_constant<int, MINKOWSKI>{});
        default: return hipErrorInvalidValue;
    }
}

// ---------------------------------------------------------------------------
// magik_trace_geodesic: one ray, FP32 reference arithmetic, every step
// recorded (relativity/host_trace_geodesic.cu:13-130).
// ---------------------------------------------------------------------------
template <int METRIC> __global__ void k_geodesic_steps(GeodesicStepParams P) {
#pragma clang fp contract(off)
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    using S = float;
    const S L = P.length_scale, a = P.spin, ms = P.mass_sign;
    const float* m = P.rot;
    S dx = m[0] * S(0) + m[1] * S(0) + m[2] * S(-1);
    S dy = m[3] * S(0) + m[4] * S(0) + m[5] * S(-1);
    S dz = m[6] * S(0) + m[7] * S(0) + m[8] * S(-1);
    S sq = S(0);
    if (P.massive) {
        S f = P.speed / kC;
        dx = dx * f; dy = dy * f; dz = dz * f;
        sq = S(-1);
    }
    S qx = P.position[0] / L, qy = P.position[1] / L, qz = P.position[2] / L, qt = S(0) / (L / kC);
    Sym4<S> g = metric_g<METRIC, S>(qx, qy, qz, a, ms, false);
    V4<S> uref = reference_observer<METRIC, S>(qx, qy, qz, a, ms);
    S wx = P.lpv[0] / kC, wy = P.lpv[1] / kC, wz = P.lpv[2] / kC;
    if (P.time_reversed) { wx = -wx; wy = -wy; wz = -wz; }
    V4<S> RX{m[0], m[3], m[6], S(0)}, RY{m[1], m[4], m[7], S(0)}, RZ{m[2], m[5], m[8], S(0)};
    V4<S> u = tetrad_apply(g, RX, RY, RZ, uref, V4<S>{wx, wy, wz, sqrt_<false>((wx * wx + wy * wy + wz * wz) - S(-1))});
    V4<S> X{S(1), S(0), S(0), S(0)}, Y{S(0), S(1), S(0), S(0)}, Z{S(0), S(0), S(1), S(0)};
    S lx = dx / L, ly = dy / L, lz = dz / L;
    V4<S> pup = tetrad_apply(g, X, Y, Z, u, V4<S>{lx, ly, lz, sqrt_<false>((lx * lx + ly * ly + lz * lz) - sq)});
    V4<S> pm = lower(g, pup);
    St<S> y{qx, qy, qz, qt, pm.x, pm.y, pm.z, pm.t};
    St<S> d = rhs_ref<METRIC, S>(y, a, ms);
    auto put = [&](uint32_t k, const St<S>& s, const St<S>& dd, S con, S aff, S dh) {
        magik_trace_geodesic_step_t& o = P.steps[k];
        o.four_position[0] = s.qx; o.four_position[1] = s.qy; o.four_position[2] = s.qz; o.four_position[3] = s.qt;
        o.four_momentum[0] = s.px; o.four_momentum[1] = s.py; o.four_momentum[2] = s.pz; o.four_momentum[3] = s.pt;
        o.moment_derivative[0] = dd.px; o.moment_derivative[1] = dd.py; o.moment_derivative[2] = dd.pz; o.moment_derivative[3] = dd.pt;
        o.four_vector_contraction = con;
        o.affine_parameter = aff;
        o.delta_affine_parameter = dh;
    };
    const S contraction = dot4(u, pm);
    S h = METRIC == MINKOWSKI ? S(1000) : S(1);
    uint32_t n = 0;
    if (P.n_alloc > 0) {
        put(0, y, d, contraction, S(0), h);
        n = 1;
    }
    S affine = S(0);
    while (n < P.n_alloc) {
        St<S> y1;
        S ht;
        if constexpr (METRIC == MINKOWSKI) {
            ht = h;
            y1 = add(y, mul(h, rhs_ref<METRIC, S>(y, a, ms)));
        } else {
            for (uint32_t i = 0; i < P.n_max; ++i) {
                S hn;
                bool acc;
                y1 = trial_reference<METRIC, S>(y, h, a, ms, S(P.abs_tol), S(P.rel_tol), S(P.h_max), hn, acc);
                ht = h;
                h = hn;
                if (acc) break;
            }
        }
        St<S> d1 = rhs_ref<METRIC, S>(y1, a, ms);
        affine = affine + ht;
        put(n, y1, d1, contraction, affine, h);
        ++n;
        y = y1;
        bool inside = false;
        if constexpr (METRIC == KERR) {
            if (P.spin_in_range) inside = !(P.horizon_r < kerr_schild_r(y1.qx, y1.qy, y1.qz, a));
        } else if constexpr (METRIC == SCHWARZSCHILD) {
            S r = sqrt_<false>(y1.qx * y1.qx + y1.qy * y1.qy + y1.qz * y1.qz);
            inside = ms < S(0) ? (r < S(1e-3f)) : (r <= P.horizon_r);
        }
        if (inside) break;
    }
    for (uint32_t k = 0; k < n; ++k)
        for (int c = 0; c < 4; ++c) P.steps[k].four_position[c] *= L;
    *P.n_written = n;
}

// ---------------------------------------------------------------------------
// Launchers (compile-time dispatch over the metric plugin, as the reference's
// dispatch_list over supported_metrics, relativity_utilities.cuh:197-202).
// ---------------------------------------------------------------------------
template <template <int> class K, class Fn> static hipError_t dispatch_metric(int metric, Fn&& fn) {
    switch (metric) {
        case MINKOWSKI: return fn(std::integral_constant<int, MINKOWSKI>{});
        case SCHWARZSCHILD: return fn(std::integral_constant<int, SCHWARZSCHILD>{});
        case KERR: return fn(std::integral_constant<int, KERR>{});
        default: return hipErrorInvalidValue;
    }
}
template <int> struct tag_t {};

hipError_t launch_ic_camera(const RayGenParams& p, IcCamera* d_ic, hipStream_t s) {
    hipLaunchKernelGGL(k_ic_camera, dim3(1), dim3(64), 0, s, p, d_ic);
    return hipGetLastError();
}

hipError_t launch_generate_rays(const RayGenParams& p, IcCamera* d_ic, hipStream_t s) {
    const uint64_t n = uint64_t(p.W) * p.H * p.spp;
    if (n == 0) return hipSuccess;
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 65536ull * 16) blocks = 65536ull * 16;
    if (p.precision != MGK_PRECISION_FP32_REFERENCE) {
        hipLaunchKernelGGL(k_ic_camera, dim3(1), dim3(64), 0, s, p, d_ic);
        hipLaunchKernelGGL(k_generate_rays_fp64, dim3(unsigned(blocks)), dim3(256), 0, s, p, d_ic);
        return hipGetLastError();
    }
    return dispatch_metric<tag_t>(p.metric, [&](auto m) -> hipError_t {
        constexpr int M = decltype(m)::value;
        hipLaunchKernelGGL((k_generate_rays_ref<M>), dim3(unsigned(blocks)), dim3(256), 0, s, p);
        return hipGetLastError();
    });
}

template <bool MESH> hipError_t launch_trace_t(const TraceParams& p, int grid, hipStream_t s) {
    return dispatch_metric<tag_t>(p.metric, [&](auto m) -> hipError_t {
        constexpr int M = decltype(m)::value;
        switch (p.precision) {
            case MGK_PRECISION_FP32_REFERENCE:
                hipLaunchKernelGGL((k_trace<M, MGK_PRECISION_FP32_REFERENCE, MESH>), dim3(grid), dim3(256), 0, s, p);
                break;
            case MGK_PRECISION_MIXED:
                hipLaunchKernelGGL((k_trace<M, MGK_PRECISION_MIXED, MESH>), dim3(grid), dim3(256), 0, s, p);
                break;
            default:
                hipLaunchKernelGGL((k_trace<M, MGK_PRECISION_FP64, MESH>), dim3(grid), dim3(256), 0, s, p);
                break;
        }
        return hipGetLastError();
    });
}

hipError_t launch_trace(const TraceParams& p, int grid, hipStream_t s) {
    return (p.flags & MGK_SCENE_MESHES) ? launch_trace_t<true>(p, grid, s) : launch_trace_t<false>(p, grid, s);
}

template <int FEAT>
hipError_t launch_trace_zone_t(const TraceParams& p, const ZoneParams& z, bool near, int grid, hipStream_t s) {
    return dispatch_metric<tag_t>(p.metric, [&](auto m) -> hipError_t {
        constexpr int M = decltype(m)::value;
        if constexpr (M == MINKOWSKI) {
            return hipErrorInvalidValue;   // flat space has no zones: one persistent launch (launch_trace)
        } else {
            if (near) hipLaunchKernelGGL((k_trace_zone<M, true, FEAT>), dim3(grid), dim3(256), 0, s, p, z);
            else hipLaunchKernelGGL((k_trace_zone<M, false, FEAT>), dim3(grid), dim3(256), 0, s, p, z);
            return hipGetLastError();
        }
    });
}

hipError_t launch_trace_zone(const TraceParams& p, const ZoneParams& z, bool near, int grid, hipStream_t s) {
    const bool mesh = (p.flags & MGK_SCENE_MESHES) != 0;
    const bool geo = (p.flags & MGK_SCENE_DISK) && (p.flags & MGK_SCENE_DISK_ON_GEODESIC);
    if (mesh) return geo ? launch_trace_zone_t<kZoneMesh | kZoneDiskGeo>(p, z, near, grid, s) : launch_trace_zone_t<kZoneMesh>(p, z, near, grid, s);
    return geo ? launch_trace_zone_t<kZoneDiskGeo>(p, z, near, grid, s) : launch_trace_zone_t<0>(p, z, near, grid, s);
}

hipError_t launch_generate_rays_batch(const RayGenParams* d_cams, const uint64_t* d_offsets, uint32_t n_cam, uint64_t n, IcCamera* d_ic,
                                      hipStream_t s) {
    if (n == 0 || n_cam == 0) return hipSuccess;
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 65536ull * 16) blocks = 65536ull * 16;
    hipLaunchKernelGGL(k_ic_cameras, dim3((n_cam + 63) / 64), dim3(64), 0, s, d_cams, n_cam, d_ic);
    hipLaunchKernelGGL(k_generate_rays_batch, dim3(unsigned(blocks)), dim3(256), 0, s, d_cams, d_offsets, n_cam, n,
                       static_cast<const IcCamera*>(d_ic));
    return hipGetLastError();
}

hipError_t launch_shade(const ShadeParams& p, hipStream_t s) {
    const uint32_t npix = p.W * p.H;
    if (npix == 0) return hipSuccess;
    hipLaunchKernelGGL(k_shade_response, dim3((kRespN + 255) / 256), dim3(256), 0, s, p, const_cast<float4*>(p.resp));
    hipLaunchKernelGGL(k_shade, dim3((npix + kShadeBlock - 1) / kShadeBlock), dim3(kShadeBlock), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_geodesic_steps(const GeodesicStepParams& p, hipStream_t s) {
    return dispatch_metric<tag_t>(p.metric, [&](auto m) -> hipError_t {
        constexpr int M = decltype(m)::value;
        hipLaunchKernelGGL((k_geodesic_steps<M>), dim3(1), dim3(64), 0, s, p);
        return hipGetLastError();
    });
}

// Step histogram: path id (i, or queue[i]) lands in bin clamp(steps[id], 1,
// n_bin - 1) of limit_out when its hit class is MGK_HIT_STEP_LIMIT, else of
// end_out (hits NULL: always end_out).  The first kHistLds bins are summed in
// LDS per block, one global atomic per non-empty bin and block.
constexpr uint32_t kHistLds = 2048;
constexpr int kHistBlock = 256;

__global__ __launch_bounds__(kHistBlock) void k_step_hist(const uint32_t* __restrict__ steps, const mgk_hit_t* __restrict__ hits,
                                                           const uint32_t* __restrict__ queue, uint64_t n, uint32_t n_bin,
                                                           uint32_t* end_out, uint32_t* limit_out) {
    __shared__ uint32_t h_end[kHistLds], h_lim[kHistLds];
    const uint32_t nl = n_bin < kHistLds ? n_bin : kHistLds;
    for (uint32_t b = threadIdx.x; b < nl; b += kHistBlock) {
        h_end[b] = 0u;
        h_lim[b] = 0u;
    }
    __syncthreads();
    const uint64_t stride = uint64_t(gridDim.x) * kHistBlock;
    for (uint64_t i = uint64_t(blockIdx.x) * kHistBlock + threadIdx.x; i < n; i += stride) {
        const uint32_t id = queue ? queue[i] : uint32_t(i);
        const uint32_t s = steps[id];
        const uint32_t b = s < 1u ? 1u : (s > n_bin - 1u ? n_bin - 1u : s);
        const bool lim = hits && hits[id].cls == uint32_t(MGK_HIT_STEP_LIMIT);
        if (b < nl)
            atomicAdd(lim ? &h_lim[b] : &h_end[b], 1u);
        else
            atomicAdd(lim ? &limit_out[b] : &end_out[b], 1u);
    }
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < nl; b += kHistBlock) {
        if (h_end[b]) atomicAdd(&end_out[b], h_end[b]);
        if (h_lim[b]) atomicAdd(&limit_out[b], h_lim[b]);
    }
}

hipError_t launch_step_hist(const uint32_t* steps, const mgk_hit_t* hits, const uint32_t* queue, uint64_t n, uint32_t n_bin,
                            uint32_t* end_out, uint32_t* limit_out, int n_cu, hipStream_t s) {
    if (n == 0) return hipSuccess;
    if (n_bin < 2) return hipErrorInvalidValue;
    const uint64_t want = (n + kHistBlock - 1) / kHistBlock, cap = uint64_t(n_cu > 0 ? n_cu : 1) * 4u;
    const unsigned grid = unsigned(want < cap ? want : cap);
    hipLaunchKernelGGL(k_step_hist, dim3(grid), dim3(kHistBlock), 0, s, steps, hits, queue, n, n_bin, end_out, limit_out);
    return hipGetLastError();
}

}  // namespace mgk
