// Device-side arithmetic of the geodesic tracer for gfx950 (CDNA4, wave64).
//
// Everything here is __device__ code written for one ray per lane: SoA loads,
// registers only, no local memory.  The spacetime is a compile-time template
// parameter (the reference's metric<T> plugin seam,
// relativity_utilities.cuh:97-120); the real type S selects FP32 or FP64.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "magik_amd.h"

namespace mgk {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kC = 299792458.0f;
constexpr float kPlanck = 6.62607015E-34f;
constexpr float kBoltzmann = 1.380649E-23f;
constexpr float kG = 6.67430E-11f;

// ---------------------------------------------------------------------------
// Reciprocal / square root.  FAST selects the single-instruction gfx950
// v_rcp_f32 / v_sqrt_f32 (1 ulp) for the bulk FP32 right-hand side; otherwise
// IEEE-rounded division and square root (the reference-parity and FP64 paths).
// ---------------------------------------------------------------------------
template <bool FAST> __device__ __forceinline__ float rcp(float x) {
    if constexpr (FAST) return __builtin_amdgcn_rcpf(x);
    else return 1.0f / x;
}
// FP64 FAST: v_rcp_f64 / v_rsq_f64 seeds refined by one Newton step (the
// seed's relative error e becomes e^2, below 1e-13 for any seed good to
// 2^-22), far cheaper than the IEEE division / square-root sequences; used
// where the result only has to be far more accurate than the integration
// tolerance (near-zone right-hand side, null renormalisation).
template <bool FAST> __device__ __forceinline__ double rcp(double x) {
    if constexpr (FAST) {
        const double y = __builtin_amdgcn_rcp(x);
        return fma(y, fma(-x, y, 1.0), y);
    } else {
        return 1.0 / x;
    }
}
// Reciprocal square root: v_rsq_f32 (1 ulp), or for FP64 the v_rsq_f64 seed
// with one Newton step (relative error e -> 1.5 e^2).
template <bool FAST> __device__ __forceinline__ float rsq(float x) {
    if constexpr (FAST) return __builtin_amdgcn_rsqf(x);
    else return 1.0f / sqrtf(x);
}
template <bool FAST> __device__ __forceinline__ double rsq(double x) {
    if constexpr (FAST) {
        const double y = __builtin_amdgcn_rsq(x);
        return y * fma(-0.5 * x * y, y, 1.5);
    } else {
        return 1.0 / __dsqrt_rn(x);
    }
}
// The IEEE square root is sqrtf: on this toolchain it compiles to v_sqrt_f32
// plus the one-ulp correction that rounds correctly, while __fsqrt_rn compiles
// to the bare v_sqrt_f32 (1 ulp), which made FP32 reference trajectories part
// from the CPU's within a step.
template <bool FAST> __device__ __forceinline__ float sqrt_(float x) {
    if constexpr (FAST) return __builtin_amdgcn_sqrtf(x);
    else return sqrtf(x);
}
template <bool FAST> __device__ __forceinline__ double sqrt_(double x) {
    if constexpr (FAST) {
        // y ~ 1/sqrt(x): one Newton step, then s = x y with one correction
        double y = __builtin_amdgcn_rsq(x);
        double h = 0.5 * x;
        y = y * fma(-h * y, y, 1.5);
        double s = x * y;
        double r = fma(-s, s, x);
        return x > 0.0 ? fma(r, 0.5 * y, s) : (x == 0.0 ? 0.0 : __dsqrt_rn(x));
    } else {
        return __dsqrt_rn(x);
    }
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11).  Counter (pixel, sample, block, 0),
// key derived from (seed, camera): the per-path random stream.
// ---------------------------------------------------------------------------
__host__ __device__ __forceinline__ uint32_t mulhi32(uint32_t a, uint32_t b) {
    return uint32_t((uint64_t(a) * uint64_t(b)) >> 32);
}
__host__ __device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint32_t lo0 = 0xD2511F53u * c[0], hi0 = mulhi32(0xD2511F53u, c[0]);
        uint32_t lo1 = 0xCD9E8D57u * c[2], hi1 = mulhi32(0xCD9E8D57u, c[2]);
        uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}
__host__ __device__ __forceinline__ float u01(uint32_t v) { return float(v >> 8) * (1.0f / 16777216.0f); }
__device__ __forceinline__ void path_uniforms(uint64_t seed, uint32_t cam, uint32_t pixel, uint32_t sample, uint32_t block,
                                              float u[4]) {
    uint32_t c[4] = {pixel, sample, block, 0u};
    philox4x32_10(c, uint32_t(seed) ^ (cam * 0x9E3779B9u), uint32_t(seed >> 32) + cam);
    u[0] = u01(c[0]); u[1] = u01(c[1]); u[2] = u01(c[2]); u[3] = u01(c[3]);
}

// ---------------------------------------------------------------------------
// Spacetime parameters shared by every kernel of one launch.
// ---------------------------------------------------------------------------
struct MetricArgs {
    float spin;          // signed spin of the equations (negated when time reversed)
    float mass_sign;     // Schwarzschild
    float horizon_scalar;
    float length_scale;  // G M / c^2 in metres (1 for Minkowski)
    int time_reversed;
};

// Geodesic state x^mu (q) and covariant p_mu, t last.
template <class S> struct St {
    S qx, qy, qz, qt, px, py, pz, pt;
};

// Right-hand side split as dq = P(p) - c, dp = (dpx, dpy, dpz, 0), with
// P(p) = (p_x, p_y, p_z, -p_t) and c = f (k.p - p_t) (k_x, k_y, k_z, -1): the
// Kerr-Schild Hamiltonian equations of metric_kerr.cuh:38-105 (Schwarzschild:
// metric_schwarzschild.cuh:8-72; Minkowski: c = 0, dp = 0).
template <class S> struct Rhs {
    S cx, cy, cz, ct, dpx, dpy, dpz;
};

template <int METRIC, class S, bool FAST>
__device__ __forceinline__ Rhs<S> rhs(S x, S y, S z, S px, S py, S pz, S pt, S a, S mass_sign) {
    Rhs<S> o;
    if constexpr (METRIC == MINKOWSKI) {
        o.cx = o.cy = o.cz = o.ct = o.dpx = o.dpy = o.dpz = S(0);
        return o;
    } else {
        // r from r^4 - (rho^2 - a^2) r^2 - a^2 y^2 = 0: 2 r^2 = w + sigma with
        // w = rho^2 - a^2, sigma = sqrt(w^2 + 4 a^2 y^2).  The same quartic
        // gives r^4 + a^2 y^2 = r^2 sigma, hence
        //   f = 2 r^3 / (r^4 + a^2 y^2) = 2 r / sigma,
        //   dr/dx = x r / sigma, dr/dz = z r / sigma, dr/dy = y (r^2 + a^2) / (r sigma),
        // and with A = x p_x + z p_z, B = x p_z - z p_x, d = r^2 + a^2, the
        // spatial k.p = (r A + a B) / d + p_y y / r, whose gradient at fixed p is
        //   p.d_i k = dr_i [(A - 2 r u) / d - p_y y / r^2] + (r dA_i + a dB_i) / d + delta_iy p_y / r
        // with u = (r A + a B) / d.  Three hardware seeds (two rsq, one rcp).
        const S y_sq = y * y;
        S a_sq, sigma, inv_sigma, r_sq;
        if constexpr (METRIC == KERR) {
            a_sq = a * a;
            const S w = (x * x + y_sq + z * z) - a_sq;
            const S sig_sq = S(4) * a_sq * y_sq + w * w;
            inv_sigma = rsq<FAST>(sig_sq);
            sigma = sig_sq * inv_sigma;
            r_sq = S(0.5) * (w + sigma);
        } else {
            a_sq = S(0);
            r_sq = x * x + y_sq + z * z;
        }
        const S inv_r = rsq<FAST>(r_sq);
        const S r = r_sq * inv_r;
        const S inv_rsq = inv_r * inv_r;
        if constexpr (METRIC != KERR) { sigma = r_sq; inv_sigma = inv_rsq; }
        const S d = r_sq + a_sq;
        const S inv_d = METRIC == KERR ? rcp<FAST>(d) : inv_rsq;
        const S s_xz = r * inv_sigma, s_y = d * inv_r * inv_sigma;
        const S drx = x * s_xz, dry = y * s_y, drz = z * s_xz;
        const S inv_sig_sq = inv_sigma * inv_sigma;
        // half_f = f / 2 and the halved gradient of f: the factor 1/2 of the
        // Hamiltonian's (k.p - p_t)^2 term is folded in here
        S half_f, dfs, dfy_extra;
        if constexpr (METRIC == KERR) {
            half_f = s_xz;   // r / sigma
            dfs = (S(3) * a_sq * y_sq - r_sq * r_sq) * (inv_rsq * inv_sig_sq);
            dfy_extra = S(2) * a_sq * y * (inv_r * inv_sig_sq);
        } else {
            half_f = mass_sign * inv_r;
            dfs = -mass_sign * inv_rsq;
            dfy_extra = S(0);
        }
        const S dfx = dfs * drx, dfy = dfs * dry - dfy_extra, dfz = dfs * drz;
        const S A = x * px + z * pz;
        const S ky = y * inv_r;
        S u, kx, kz, ex, ez;
        if constexpr (METRIC == KERR) {
            const S B = x * pz - z * px;
            u = (r * A + a * B) * inv_d;
            kx = (r * x - a * z) * inv_d;
            kz = (r * z + a * x) * inv_d;
            ex = (r * px + a * pz) * inv_d;
            ez = (r * pz - a * px) * inv_d;
        } else {
            u = A * inv_r;
            kx = x * inv_r;
            kz = z * inv_r;
            ex = px * inv_r;
            ez = pz * inv_r;
        }
        const S kTp = (u + py * ky) - pt;
        const S Hc = (A - S(2) * r * u) * inv_d - py * ky * inv_r;
        const S bx = drx * Hc + ex;
        const S by = dry * Hc + py * inv_r;
        const S bz = drz * Hc + ez;
        const S f_kTp = S(2) * (half_f * kTp);
        const S kTp_sq = kTp * kTp;
        o.cx = f_kTp * kx;
        o.cy = f_kTp * ky;
        o.cz = f_kTp * kz;
        o.ct = -f_kTp;
        o.dpx = kTp_sq * dfx + f_kTp * bx;
        o.dpy = kTp_sq * dfy + f_kTp * by;
        o.dpz = kTp_sq * dfz + f_kTp * bz;
        return o;
    }
}

// The same equations in the exact operation order of the reference
// (metric_kerr.cuh:38-105), used by the FP32 reference-parity path so its
// rounding follows the reference integrator's.
template <int METRIC, class S>
__device__ __forceinline__ St<S> rhs_ref(const St<S>& s, S a, S mass_sign) {
#pragma clang fp contract(off)
    St<S> o;
    if constexpr (METRIC == MINKOWSKI) {
        o.qx = s.px; o.qy = s.py; o.qz = s.pz; o.qt = -s.pt;
        o.px = o.py = o.pz = o.pt = S(0);
        return o;
    } else {
        S x = s.qx, y = s.qy, z = s.qz, pt = s.pt, px = s.px, py = s.py, pz = s.pz;
        S a_sq, y_sq = y * y, w_sq, sigma;
        if constexpr (METRIC == KERR) {
            a_sq = a * a;
            w_sq = (x * x + y * y + z * z) - a_sq;
            sigma = sqrt_<false>(S(4) * a_sq * y_sq + w_sq * w_sq);
        } else {
            a_sq = S(0);
            a = S(0);
            w_sq = (x * x + y * y + z * z);
            sigma = w_sq;
        }
        S r = sqrt_<false>(S(0.5) * (w_sq + sigma));
        S r_sq = r * r, r_cu = r_sq * r, r_quad = r_sq * r_sq, r_six = r_cu * r_cu;
        S inv_r = S(1) / r;
        S inv_rsq = inv_r * inv_r;
        S inv_sigma = S(1) / sigma;
        S scale_xz = S(0.5) * inv_r * (S(1) + w_sq * inv_sigma);
        S scale_y = METRIC == KERR ? S(0.5) * inv_r * (S(1) + (w_sq + S(2) * a_sq) * inv_sigma)
                                   : S(0.5) * inv_r * (S(1) + w_sq * inv_sigma);
        S drx = x * scale_xz, dry = y * scale_y, drz = z * scale_xz;
        S d = r_sq + a_sq;
        S inv_d = S(1) / d;
        S inv_dsq = inv_d * inv_d;
        S kx_num = METRIC == KERR ? r * x - a * z : r * x;
        S kz_num = METRIC == KERR ? r * z + a * x : r * z;
        S kx = kx_num * inv_d, kz = kz_num * inv_d, ky = y * inv_r;
        S f_den = METRIC == KERR ? r_quad + a_sq * y_sq : r_quad;
        S inv_f_den_sq = S(1) / (f_den * f_den);
        S f, dfs, dfx, dfy, dfz;
        if constexpr (METRIC == KERR) {
            f = S(2) * r_cu / f_den;
            dfs = S(2) * (S(3) * a_sq * r_sq * y_sq - r_six) * inv_f_den_sq;
            dfx = dfs * drx;
            dfy = dfs * dry - S(4) * a_sq * r_cu * y * inv_f_den_sq;
            dfz = dfs * drz;
        } else {
            f = mass_sign * S(2) * r_cu / f_den;
            dfs = mass_sign * S(2) * -r_six * inv_f_den_sq;
            dfx = dfs * drx;
            dfy = dfs * dry;
            dfz = dfs * drz;
        }
        S dkx_dx = ((x * drx + r) * d - S(2) * r * drx * kx_num) * inv_dsq;
        S dkx_dy = ((x * dry) * d - S(2) * r * dry * kx_num) * inv_dsq;
        S dkx_dz = (METRIC == KERR ? (x * drz - a) * d : (x * drz) * d) - S(2) * r * drz * kx_num;
        dkx_dz = dkx_dz * inv_dsq;
        S dky_dx = -y * drx * inv_rsq;
        S dky_dy = (r - y * dry) * inv_rsq;
        S dky_dz = -y * drz * inv_rsq;
        S dkz_dx = (METRIC == KERR ? (a + z * drx) * d : (z * drx) * d) - S(2) * r * drx * kz_num;
        dkz_dx = dkz_dx * inv_dsq;
        S dkz_dy = ((z * dry) * d - S(2) * r * dry * kz_num) * inv_dsq;
        S dkz_dz = ((r + z * drz) * d - S(2) * r * drz * kz_num) * inv_dsq;
        S kTp = px * kx + py * ky + pz * kz - pt;
        S f_kTp = f * kTp;
        S half_kTp_sq = S(0.5) * kTp * kTp;
        S bx = px * dkx_dx + py * dky_dx + pz * dkz_dx;
        S by = px * dkx_dy + py * dky_dy + pz * dkz_dy;
        S bz = px * dkx_dz + py * dky_dz + pz * dkz_dz;
        o.qx = px - f_kTp * kx; o.qy = py - f_kTp * ky; o.qz = pz - f_kTp * kz; o.qt = -pt + f_kTp;
        o.px = half_kTp_sq * dfx + f_kTp * bx;
        o.py = half_kTp_sq * dfy + f_kTp * by;
        o.pz = half_kTp_sq * dfz + f_kTp * bz;
        o.pt = S(0);
        return o;
    }
}

// ---------------------------------------------------------------------------
// Metric tensor pieces: f and k of g = eta + f l l (metric_kerr.cuh:8-33,
// :177-205; metric_schwarzschild.cuh:143-165).
// ---------------------------------------------------------------------------
template <class S> struct Sym4 { S tt, tx, ty, tz, xx, xy, xz, yy, yz, zz; };

template <class S> __device__ __forceinline__ S kerr_schild_r(S x, S y, S z, S a) {
#pragma clang fp contract(off)
    S R = (x * x + y * y + z * z - a * a);
    S r2 = S(0.5) * R + S(0.5) * sqrt_<false>(R * R + S(4) * a * a * y * y);
    return sqrt_<false>(r2);
}

template <int METRIC, class S> __device__ __forceinline__ Sym4<S> metric_g(S x, S y, S z, S a, S mass_sign, bool inverse) {
#pragma clang fp contract(off)
    if constexpr (METRIC == MINKOWSKI) {
        return {S(-1), S(0), S(0), S(0), S(1), S(0), S(0), S(1), S(0), S(1)};
    } else {
        S f, kx, ky, kz;
        if constexpr (METRIC == KERR) {
            S r = kerr_schild_r(x, y, z, a);
            f = S(2) * r * r * r / (r * r * r * r + a * a * y * y);
            kx = (r * x - a * z) / (r * r + a * a);
            ky = y / r;
            kz = (r * z + a * x) / (r * r + a * a);
        } else {
            S r = sqrt_<false>(x * x + y * y + z * z);
            f = S(2) * mass_sign / r;
            kx = x / r; ky = y / r; kz = z / r;
        }
        if (!inverse)
            return {S(-1) + f, f * kx, f * ky, f * kz, S(1) + f * kx * kx, f * kx * ky, f * kx * kz, S(1) + f * ky * ky,
                    f * ky * kz, S(1) + f * kz * kz};
        return {S(-1) - f, f * kx, f * ky, f * kz, S(1) - f * kx * kx, -f * kx * ky, -f * kx * kz, S(1) - f * ky * ky,
                -f * ky * kz, S(1) - f * kz * kz};
    }
}

template <class S> struct V4 { S x, y, z, t; };

template <class S> __device__ __forceinline__ S inner(const Sym4<S>& g, V4<S> u, V4<S> v) {
#pragma clang fp contract(off)
    return g.tt * u.t * v.t + g.tx * (u.t * v.x + u.x * v.t) + g.ty * (u.t * v.y + u.y * v.t) + g.tz * (u.t * v.z + u.z * v.t) +
           g.xx * u.x * v.x + g.xy * (u.x * v.y + u.y * v.x) + g.xz * (u.x * v.z + u.z * v.x) + g.yy * u.y * v.y +
           g.yz * (u.y * v.z + u.z * v.y) + g.zz * u.z * v.z;
}
template <class S> __device__ __forceinline__ V4<S> lower(const Sym4<S>& g, V4<S> u) {
#pragma clang fp contract(off)
    return {g.tx * u.t + g.xx * u.x + g.xy * u.y + g.xz * u.z, g.ty * u.t + g.xy * u.x + g.yy * u.y + g.yz * u.z,
            g.tz * u.t + g.xz * u.x + g.yz * u.y + g.zz * u.z, g.tt * u.t + g.tx * u.x + g.ty * u.y + g.tz * u.z};
}
template <class S> __device__ __forceinline__ V4<S> axpy(S s, V4<S> a, V4<S> b) {   // b + s a
#pragma clang fp contract(off)
    return {b.x + s * a.x, b.y + s * a.y, b.z + s * a.z, b.t + s * a.t};
}

template <class S> __device__ __forceinline__ V4<S> scale(V4<S> a, S s) { return {a.x * s, a.y * s, a.z * s, a.t * s}; }
template <class S> __device__ __forceinline__ V4<S> divs(V4<S> a, S s) { return {a.x / s, a.y / s, a.z / s, a.t / s}; }
template <class S> __device__ __forceinline__ S dot4(V4<S> a, V4<S> b) {
#pragma clang fp contract(off)
    return ((a.x * b.x) + (a.y * b.y) + (a.z * b.z)) + (a.t * b.t);
}

// Reference observer (metric_kerr.cuh:207-226; metric_schwarzschild.cuh:167-171).
template <int METRIC, class S> __device__ __forceinline__ V4<S> reference_observer(S x, S y, S z, S a, S mass_sign) {
#pragma clang fp contract(off)
    if constexpr (METRIC == KERR) {
        S r = kerr_schild_r(x, y, z, a);
        S r_sq = r * r, a_sq = a * a;
        S ct = y / r;
        S s2 = S(1) - ct * ct;
        S sigma = r_sq + a_sq * ct * ct;
        S delta = r_sq - S(2) * r + a_sq;
        S Lam = (r_sq + a_sq) * (r_sq + a_sq) - a_sq * delta * s2;
        S ut = sqrt_<false>(Lam / (delta * sigma));
        S uph = S(2) * a * r / sqrt_<false>(Lam * delta * sigma);
        return {z * uph, S(0), -x * uph, ut};
    } else if constexpr (METRIC == SCHWARZSCHILD) {
        Sym4<S> g = metric_g<METRIC, S>(x, y, z, a, mass_sign, false);
        return {S(0), S(0), S(0), S(1) / sqrt_<false>(-g.tt)};
    } else {
        return {S(0), S(0), S(0), S(1)};
    }
}

// Gram-Schmidt tetrad with e_t = u in the order z, x, y (relativity_utilities.cuh:141-156);
// returns the global vector of local components (l.x, l.y, l.z, l.t) in that frame.
template <class S>
__device__ __forceinline__ V4<S> tetrad_apply(const Sym4<S>& g, V4<S> X, V4<S> Y, V4<S> Z, V4<S> u, V4<S> l) {
#pragma clang fp contract(off)
    V4<S> ez = axpy(inner(g, Z, u), u, Z);
    ez = divs(ez, sqrt_<false>(inner(g, ez, ez)));
    V4<S> ex = axpy(-inner(g, X, ez), ez, axpy(inner(g, X, u), u, X));
    ex = divs(ex, sqrt_<false>(inner(g, ex, ex)));
    V4<S> ey = axpy(-inner(g, Y, ex), ex, axpy(-inner(g, Y, ez), ez, axpy(inner(g, Y, u), u, Y)));
    ey = divs(ey, sqrt_<false>(inner(g, ey, ey)));
    V4<S> r = scale(u, l.t);
    r = axpy(l.x, ex, r);
    r = axpy(l.y, ey, r);
    r = axpy(l.z, ez, r);
    return r;
}

// Keplerian emitter co-rotating with the equations' spin (Bardeen, Press &
// Teukolsky 1972); see the oracle's kepler_velocity for the derivation.
template <int METRIC, class S> __device__ __forceinline__ void kepler_ut_w(S r, S a, S& ut, S& w) {
#pragma clang fp contract(off)
    if constexpr (METRIC == MINKOWSKI) {
        ut = S(1);
        w = S(0);
    } else {
        S aa = METRIC == KERR ? (a < S(0) ? -a : a) : S(0);
        S s = (METRIC == KERR && a < S(0)) ? S(-1) : S(1);
        S sr = sqrt_<false>(r);
        S r32 = r * sr;
        ut = (r32 + aa) / (sqrt_<false>(r32) * sqrt_<false>(r32 - S(3) * sr + S(2) * aa));
        w = ut * s / (r32 + aa);
    }
}

template <int METRIC, class S> __device__ __forceinline__ V4<S> kepler_velocity(S x, S z, S r, S a) {
    if constexpr (METRIC == MINKOWSKI) {
        return {S(0), S(0), S(0), S(1)};
    } else {
        S aa = METRIC == KERR ? (a < S(0) ? -a : a) : S(0);
        S s = (METRIC == KERR && a < S(0)) ? S(-1) : S(1);
        S sr = sqrt_<false>(r);
        S r32 = r * sr;
        S ut = (r32 + aa) / (sqrt_<false>(r32) * sqrt_<false>(r32 - S(3) * sr + S(2) * aa));
        S w = ut * s / (r32 + aa);
        return {z * w, S(0), -x * w, ut};
    }
}

__device__ __forceinline__ float planck(float lambda_nm, float T) {
    if (T <= 0.0f) return 0.0f;
    float lm = lambda_nm * 1e-9f;
    const float t0 = kPlanck * kC;
    const float t1 = 2.0f * t0 * kC;
    float t2 = lm * lm * lm * lm * lm;
    // hardware reciprocals (1 ulp) for the two quotients: IEEE FP32 division is
    // a multi-instruction sequence, and shading evaluates this per hero wavelength
    float t3 = expf(t0 * __builtin_amdgcn_rcpf(lm * kBoltzmann * T)) - 1.0f;
    return t1 * __builtin_amdgcn_rcpf(t2 * t3);
}

// White LED spectrum: five Gaussians exp(-(x - mu)^2 / (2 w^2)) k
// (internal_math.h:2441-2445, :2501-2509).
__device__ __forceinline__ float white_led(float x) {
    auto gd = [](float x, float mu, float w, float k) { return expf(-0.5f * (((x - mu) * (x - mu)) / (w * w))) * k; };
    return gd(x, 577.0f, 50.7937003968f, 157.0f) + gd(x, 451.0f, 7.07106781187f, 200.0f) + gd(x, 455.0f, 14.1421356237f, 50.0f) +
           gd(x, 528.0f, 13.7840487521f, 20.0f) + gd(x, 690.0f, 44.72135955f, 10.0f);
}

}  // namespace mgk

namespace mgk {

// Jittered sensor point of pixel (px, py) in metres on the z = 0 sensor plane
// (device_sensor_utilities.cuh:35-80: u = -(x/W - 0.5), v = -(y/H - 0.5)
// scaled by the aspect ratio, the sensor width as the full width, jitter of
// one pixel side from the diagonal geometry).  Shared by the ray generator
// and the surrogate's screen coordinates.
template <class S>
__device__ __forceinline__ void sensor_point(uint32_t W, uint32_t H, float sensor_width_mm, uint32_t px, uint32_t py, float u0,
                                             float u1, S& sx, S& sy) {
#pragma clang fp contract(off)
    const S sw = S(sensor_width_mm);
    S inv_s = sw / sqrt_<false>(S(float(W * W + H * H)));
    S width_mm = S(float(W)) * inv_s, height_mm = S(float(H)) * inv_s;
    S area = (width_mm * height_mm) / S(1E6f);
    S side = sqrt_<false>(area / S(float(W * H)));
    S ju = (S(u0) - S(0.5)) * side, jv = (S(u1) - S(0.5)) * side;
    S uu = -((S(float(px)) / S(float(W))) - S(0.5));
    S vv = -(S(float(py)) / S(float(H)) - S(0.5));
    S ratio = S(float(H)) / S(float(W));
    if (ratio < S(1)) vv = vv * ratio; else uu = uu / ratio;
    sx = ((-uu * sw) / S(1000)) + ju;
    sy = ((-vv * sw) / S(1000)) + jv;
}

// ---------------------------------------------------------------------------
// Surrogate feature space.  A camera at Boyer-Lindquist (r, theta) of a Kerr
// hole with spin a sits at the Kerr-Schild point (0, r cos theta,
// sqrt(r^2 + a^2) sin theta), looks at the origin with +y up, and a ray's
// screen coordinates are alpha = r d_x / (-d_z), beta = r d_y / (-d_z) for its
// camera-local unit direction d.  ic_from_features builds the ray's initial
// conditions exactly as the render path does (reference observer, camera
// tetrad, t0 = c/L, |p| ~ 1/L).
// ---------------------------------------------------------------------------
struct Camframe {
    double P[3];           // dimensionless position
    double X[3], Y[3], Z[3];   // camera axes (columns of the orientation)
};

__host__ __device__ inline void bl_camera_frame(double a, double r, double theta, Camframe& f) {
    f.P[0] = 0.0;
    f.P[1] = r * cos(theta);
    f.P[2] = sqrt(r * r + a * a) * sin(theta);
    double n = sqrt(f.P[1] * f.P[1] + f.P[2] * f.P[2]);
    f.Z[0] = 0.0; f.Z[1] = f.P[1] / n; f.Z[2] = f.P[2] / n;
    double d = f.Z[1];
    double y0 = -d * f.Z[0], y1 = 1.0 - d * f.Z[1], y2 = -d * f.Z[2];
    double yl = sqrt(y0 * y0 + y1 * y1 + y2 * y2);
    f.Y[0] = y0 / yl; f.Y[1] = y1 / yl; f.Y[2] = y2 / yl;
    f.X[0] = f.Y[1] * f.Z[2] - f.Y[2] * f.Z[1];
    f.X[1] = f.Y[2] * f.Z[0] - f.Y[0] * f.Z[2];
    f.X[2] = f.Y[0] * f.Z[1] - f.Y[1] * f.Z[0];
}

// Initial conditions of a Kerr ray from a camera frame and a camera-local unit
// direction d (the metric's spin `a_eq` is the time-reversed one).
template <class S>
__device__ __forceinline__ void ic_from_frame(const Camframe& f, S a_eq, S L, const S d[3], St<S>& y, S& contraction, V4<S>& u_out) {
    S qx = S(f.P[0]), qy = S(f.P[1]), qz = S(f.P[2]), qt = S(1) / (L / S(kC));
    Sym4<S> g = metric_g<KERR, S>(qx, qy, qz, a_eq, S(1), false);
    V4<S> uref = reference_observer<KERR, S>(qx, qy, qz, a_eq, S(1));
    V4<S> X{S(1), S(0), S(0), S(0)}, Y{S(0), S(1), S(0), S(0)}, Z{S(0), S(0), S(1), S(0)};
    V4<S> u = tetrad_apply(g, X, Y, Z, uref, V4<S>{S(0), S(0), S(0), sqrt_<false>(S(0) - S(-1))});
    V4<S> CX{S(f.X[0]), S(f.X[1]), S(f.X[2]), S(0)}, CY{S(f.Y[0]), S(f.Y[1]), S(f.Y[2]), S(0)}, CZ{S(f.Z[0]), S(f.Z[1]), S(f.Z[2]), S(0)};
    S lx = d[0] / L, ly = d[1] / L, lz = d[2] / L;
    V4<S> pup = tetrad_apply(g, CX, CY, CZ, u, V4<S>{lx, ly, lz, sqrt_<false>(lx * lx + ly * ly + lz * lz)});
    V4<S> pm = lower(g, pup);
    y = St<S>{qx, qy, qz, qt, pm.x, pm.y, pm.z, pm.t};
    contraction = dot4(u, pm);
    u_out = u;
}

}  // namespace mgk
