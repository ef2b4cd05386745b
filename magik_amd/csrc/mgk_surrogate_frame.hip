// The surrogate render path (mgk_render_frame_surrogate): every path of the
// frame is answered by the MLP unless a fallback rule sends it to the RKF45
// integrator; a held-out judge sample is traced as well and compared.
//
//   k_mlp_infer_res: (mgk_mlp_infer.hpp) pixel + sample -> screen point
//                    (alpha, beta) of the lens point's camera frame ->
//                    normalised features, computed in the kernel -> network
//                    outputs (n x 8; outputs 6 and 7 carry alpha and beta)
//   k_sur_decode   : outputs -> class / disk radius / escape direction, the
//                    fallback rules, the hit record of answered paths, and
//                    the compacted queue of paths to trace (one atomic per
//                    wave, ballot ranks within it)
//   k_sur_mino     : (fallback.analytic) the Mino-time check of the class of
//                    every confident path and the refinement of a disk
//                    path's radius (mgk_mino.hpp), and the same integrals for
//                    the unsure paths the decode deferred, in a kernel of its
//                    own so that its registers allow four waves per SIMD
//   mgk_trace_rays : the queue, written back into the same hit slots
//   k_sur_judge    : surrogate vs RKF45 on the judge sample
//   mgk_shade      : the unchanged spectral shading of every hit
//
// Redshift of an answered path comes from the ray's conserved quantities:
// g = (u.p)_camera / p_t for an escaping ray and
// g = (u.p)_camera / (u^t p_t + Omega u^t p_phi) for a circular Keplerian
// emitter at the predicted disk radius, with p_t and p_phi = z p_x - x p_z
// read from the initial conditions (both are conserved along a Kerr geodesic
// in Kerr-Schild coordinates with the spin along +y).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>

#include "magik_amd.h"
#include "mgk_device.hpp"
#include "mgk_internal.hpp"
#include "mgk_mino.hpp"
#include "mgk_raygen.hpp"
#include "mgk_mlp.hpp"
#include "mgk_mlp_infer.hpp"
#include "mgk_surrogate.hpp"

using namespace mgk;

#define HIP_TRY(expr)                                                                            \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess) {                                                                  \
            std::fprintf(stderr, "magik_amd: %s failed: %s\n", #expr, hipGetErrorString(e_));   \
            return MGK_ERROR_HIP;                                                                \
        }                                                                                        \
    } while (0)

namespace {

constexpr uint32_t kNotAnswered = 0xFFFFFFFFu;

// device counters of one frame (u64 slots; the double sums live in the
// same buffer, bit-cast)
enum {
    C_QUEUE = 0, C_CONF, C_CRIT, C_DISK, C_STAR, C_DOMAIN, C_ANALYTIC, C_JUDGE, C_AGREE, C_JDISK, C_JESC,
    C_SUM_R, C_SUM_G, C_SUM_DIR, C_MAX_DIR, C_N
};
constexpr int C_REASON_LAST = C_ANALYTIC;   // the fallback reasons are C_CONF .. C_REASON_LAST

struct FeatureParams {
    SurrogateRanges R;
    uint32_t W, H, camera_id, spp;
    uint64_t seed, n;
    float sensor_width_mm, focal_mm;
    double a, r_lens, th_lens_deg;
    // sensor_point's per-camera parts in float, and the screen scale: the
    // normalised alpha = -sx k_screen, beta = -sy k_screen
    float inv_w, inv_h, aspect, side, sw_m, k_screen;
    uint32_t wide;
};

// The input rows of the inference kernel, computed in it: lane l of a wave
// computes the screen point of path ray0 + l from its pixel and Philox
// samples (the path's own ray-generation samples), and the lanes of each
// 32-ray tile gather theirs by cross-lane reads.  Spin, r and theta are the
// camera's (the same normalised values for every path).  The normalised
// alpha and beta also leave the kernel as outputs 6 and 7 of the row.
struct InferFromCamera {
    FeatureParams P;
    float x0, x1, x2;
    static constexpr bool kExtra = true;
    __device__ __forceinline__ void load(int64_t ray0, int64_t n, int lane, bf16x8 (&xb)[kInferTiles],
                                         float (&extra)[kInferTiles][2]) const {
        const uint64_t i = uint64_t(ray0) + uint64_t(lane);
        float al_n = 0.0f, be_n = 0.0f;
        if (int64_t(i) < n) {
            // sensor_point in float with the camera's divisions done on the host
            // (the features enter the network in bf16); 32-bit indices (a frame
            // holds fewer than 2^32 paths)
            const uint32_t npix = P.W * P.H, i32 = uint32_t(i);
            const uint32_t pixel = i32 % npix, sample = i32 / npix;
            float u[4];
            path_uniforms(P.seed, P.camera_id, pixel, sample, 0, u);
            float a = float(pixel % P.W) * P.inv_w - 0.5f, b = float(pixel / P.W) * P.inv_h - 0.5f;
            if (P.wide) b *= P.aspect;
            else a *= P.aspect;
            const float sx = a * P.sw_m + (u[0] - 0.5f) * P.side, sy = b * P.sw_m + (u[1] - 0.5f) * P.side;
            // camera-local direction (-sx, -sy, -f): alpha = r d_x / (-d_z)
            al_n = -sx * P.k_screen;
            be_n = -sy * P.k_screen;
        }
        const int g = lane >> 5, r = lane & 31;
#pragma unroll
        for (int c = 0; c < kInferTiles; ++c) {
            const float a3 = __shfl(al_n, 32 * c + r), a4 = __shfl(be_n, 32 * c + r);
            const bool lo = g == 0;
            xb[c][0] = __bf16(lo ? x0 : 0.0f); xb[c][1] = __bf16(lo ? x1 : 0.0f); xb[c][2] = __bf16(lo ? x2 : 0.0f);
            xb[c][3] = __bf16(lo ? a3 : 0.0f); xb[c][4] = __bf16(lo ? a4 : 0.0f);
            xb[c][5] = xb[c][6] = xb[c][7] = __bf16(0.0f);
            extra[c][0] = a3;
            extra[c][1] = a4;
        }
    }
};

// Per-camera constants of the Mino-time check's own initial conditions (FP64,
// from the host): the sensor mapping of sensor_point and the direction
// normalisation of ray_ic_fp64 with the camera's divisions done once.
struct MinoIc {
    double inv_w, inv_h;   // 1 / W, 1 / H
    double aspect;         // H / W on the vertical offset (wide) or W / H on the horizontal one
    double side;           // jitter square side (mm)
    double sw_m;           // sensor width / 1000
    double f_m;            // focal length (m)
    double inv_l;          // 1 / L
    uint32_t wide, npix, W, sample_begin;
};

struct DecodeParams {
    SurrogateRanges R;
    uint64_t n;
    const float* Y;              // n x 8 network outputs (6 and 7: the normalised alpha, beta)
    RayGenParams G;              // the camera's ray generation, and its FP64 constants:
    const IcCamera* ic;          // a path's initial conditions are formed in registers
    MinoIc mi;
    double* rays;                // 10 x n doubles (q, p, h0, u.p), written for the paths to trace
    mgk_hit_t* hits;
    uint32_t* queue;
    uint8_t* reason;             // the fallback reasons of the deferred paths
    float* cand;                 // (fallback.analytic) per path, the class for k_sur_mino: the network's
                                 // disk radius (> 0), -1 captured, -2 escape, -10 - the network's radius
                                 // for a deferred path, 0 none
    mgk_hit_t* pred;             // judge predictions (one per judge slot)
    unsigned long long* cnt;
    const float* crit;           // kCritN screen radii of the critical curve (< 0: none)
    StarView stars;
    float star_chord2_in, star_chord2_out;
    uint32_t flags;
    double a_eq, disk_r_in, disk_r_out, r_horizon;
    MinoDisk dk;                 // the same radii with their reciprocals
    float min_conf, crit_margin, edge_margin;
    int star_test;
    uint32_t judge_stride, trace_disk, analytic;
};

// Any star whose edge lies within the band [chord_in, chord_out] of direction d.
__device__ bool near_star_edge(const DecodeParams& P, float dx, float dy, float dz, float ph) {
    const StarView& S = P.stars;
    const uint32_t ic = min(S.nc - 1, uint32_t(fmaxf(0.0f, (dy + 1.0f) * 0.5f * float(S.nc))));
    const uint32_t ip = min(S.np - 1, uint32_t(fmaxf(0.0f, (ph + kPi) * (0.5f / kPi) * float(S.np))));
    const uint32_t cell = ic * S.np + ip;
    for (uint32_t k = S.cell_start[cell]; k < S.cell_start[cell + 1]; ++k) {
        const float4 s = S.stars[S.cell_items[k]];
        const float ex = s.x - dx, ey = s.y - dy, ez = s.z - dz;
        const float c2 = ex * ex + ey * ey + ez * ez;
        if (c2 > P.star_chord2_in && c2 <= P.star_chord2_out) return true;
    }
    return false;
}

// Paths decoded per workgroup; the workgroup compacts its fallback paths in
// LDS and publishes them with one global atomic.
constexpr int kDecodeBlock = 256;
constexpr int kDecodeChunk = 16 * kDecodeBlock;

// A disk hit at radius r: the redshift of the emitter on its circular
// (Keplerian) orbit seen along the path.
// From the path's conserved p_t and p_phi = z p_x - x p_z and its u.p at the camera.
__device__ __forceinline__ void disk_hit(const DecodeParams& P, double pt, double pphi, double con, double r, mgk_hit_t& h) {
    double ut, w;
    kepler_ut_w<KERR, double>(r, P.a_eq, ut, w);
    double g = con / (ut * pt + w * pphi);
    if (isnan(g) || isinf(g)) g = 1e-10;
    h = mgk_hit_t{MGK_HIT_DISK, float(g), float(r), 0.0f};
}

enum DecodeStatus : int { kAnswered = 0, kTrace = 1, kDefer = 2, kMino = 3 };
constexpr float kCandCaptured = -1.0f, kCandEscape = -2.0f, kCandDeferred = -10.0f;

// The answer of path i from its network outputs: kAnswered (in h), kTrace or
// kDefer (the fallback reason in reason), or kMino: confident, the class to be
// confirmed by k_sur_mino, in cand; h then holds a captured or escaping path's
// answer.  C: the camera's FP64 constants, from which the path's initial
// conditions y = (x, y, z, t, p_x, p_y, p_z, p_t, h0, u.p) are formed where an
// answer needs them.
__device__ __forceinline__ int decode_one(const DecodeParams& P, const IcCamera& C, uint64_t i, mgk_hit_t& h, int& reason,
                                          float& cand) {
    bool fallback = false;
    reason = -1;
    h = mgk_hit_t{MGK_HIT_NONE, 0.0f, 0.0f, 0.0f};
    const float4 y0 = reinterpret_cast<const float4*>(P.Y + i * kMlpOut)[0];
    const float4 y1 = reinterpret_cast<const float4*>(P.Y + i * kMlpOut)[1];
    // logits (captured, disk, escape = 0)
    const float mx = fmaxf(fmaxf(y0.x, y0.y), 0.0f);
    const float e0 = __expf(y0.x - mx), e1 = __expf(y0.y - mx), e2 = __expf(-mx);
    const float inv = 1.0f / (e0 + e1 + e2);
    const float p0 = e0 * inv, p1 = e1 * inv, p2 = e2 * inv;
    const int cls = (p0 >= p1 && p0 >= p2) ? 0 : (p1 >= p2 ? 1 : 2);
    const float conf = fmaxf(p0, fmaxf(p1, p2));
    const float al = y1.z * float(P.R.s_max), be = y1.w * float(P.R.s_max);   // outputs 6, 7: alpha, beta normalised
    if (fabsf(al) > float(P.R.s_max) || fabsf(be) > float(P.R.s_max)) {
        fallback = true; reason = C_DOMAIN;
    } else if (conf < P.min_conf) {
        fallback = true; reason = C_CONF;
    } else {
        const float rho = sqrtf(al * al + be * be), psi = atan2f(be, al);
        float fk = (psi + kPi) * (float(kCritN) / (2.0f * kPi));
        int k0 = int(floorf(fk));
        const float t = fk - float(k0);
        k0 = ((k0 % kCritN) + kCritN) % kCritN;
        const int k1 = (k0 + 1) % kCritN;
        const float c0 = P.crit[k0], c1 = P.crit[k1];
        if (c0 >= 0.0f && c1 >= 0.0f && fabsf(rho - (c0 + t * (c1 - c0))) < P.crit_margin) {
            fallback = true; reason = C_CRIT;
        }
    }
    const double r = exp(P.R.lnr_mid + double(y0.z) * P.R.lnr_half);
    if (fallback) {
        // the network is not trusted here (outside its domain, unsure, or near
        // the critical curve): traced, or with the Mino-time integrals on,
        // deferred to k_sur_mino, the network's radius as the seed
        if (!P.analytic || P.trace_disk) return kTrace;
        cand = kCandDeferred - ((r > 0.0 && r < 1e30) ? float(r) : 1.0f);
        return kDefer;
    }
    if (cls == 1 && P.trace_disk) {
        reason = C_DISK;
        return kTrace;
    }
    // with the Mino-time integrals on, k_sur_mino must confirm the network's
    // class, and a disk path's radius becomes the one they give, seeded by
    // the network's
    if (P.analytic && cls == 1) {
        cand = float(r);
        return kMino;
    }
    if (!P.analytic && cls == 1 && (r < P.disk_r_in * (1.0 + P.edge_margin) || r > P.disk_r_out * (1.0 - P.edge_margin))) {
        reason = C_DISK;
        return kTrace;
    }
    if (cls == 0) {
        h.cls = MGK_HIT_HORIZON;
    } else if (cls == 1) {
        double y[10];
        ray_ic_fp64(P.G, C, i, y);
        disk_hit(P, y[7], y[2] * y[4] - y[0] * y[6], y[9], r, h);
    } else {
        const float c = fminf(1.0f, fmaxf(-1.0f, y0.w));
        const float ph = atan2f(y1.x, y1.y);
        const float sn = sqrtf(fmaxf(0.0f, 1.0f - c * c));
        float sph, cph;
        __sincosf(ph, &sph, &cph);   // ph in [-pi, pi]: the hardware sine and cosine (~1e-6)
        if (P.star_test && near_star_edge(P, sn * cph, c, sn * sph, ph)) {
            reason = C_STAR;
            return kTrace;
        }
        double y[10];
        ray_ic_fp64(P.G, C, i, y);
        double g = y[9] / y[7];
        if (isnan(g) || isinf(g)) g = 1e-10;
        h = mgk_hit_t{MGK_HIT_ESCAPE, float(g), c, ph};
    }
    if (P.analytic) {
        cand = cls == 0 ? kCandCaptured : kCandEscape;
        return kMino;
    }
    return kAnswered;
}

// Appends the lanes' flagged path indices to a block-local list (one LDS
// atomic per wave, ballot ranks within it).
__device__ __forceinline__ void wave_append(bool flag, uint32_t i, uint32_t lane, uint64_t lt_mask, uint32_t* n_local, uint32_t* buf) {
    const uint64_t m = __ballot(flag);
    if (m == 0) return;
    uint32_t off = 0;
    if (lane == 0) off = atomicAdd(n_local, uint32_t(__popcll(m)));
    off = __shfl(off, 0);
    if (flag) buf[off + __popcll(m & lt_mask)] = i;
}

// Held to six waves per SIMD (80 registers, a few spilled; 114 and four waves
// unbounded): 0.685 -> 0.612 ms per 1080p frame; its LDS (17 KB per
// workgroup) allows it.
__global__ void __launch_bounds__(kDecodeBlock, 6) k_sur_decode(DecodeParams P) {
    __shared__ uint32_t qbuf[kDecodeChunk];
    __shared__ uint32_t n_local[1];
    __shared__ uint32_t reasons[C_REASON_LAST + 1];
    __shared__ unsigned long long base_s[1];
    const uint32_t lane = __lane_id();
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    // the camera's FP64 constants: uniform and read before any store, so
    // scalar loads into SGPRs (a copy in LDS took 40 more vector registers)
    const IcCamera C = *P.ic;
    if (threadIdx.x < 1) n_local[threadIdx.x] = 0;
    if (threadIdx.x <= C_REASON_LAST) reasons[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t begin = uint64_t(blockIdx.x) * kDecodeChunk;
#pragma unroll 1
    for (int it = 0; it < kDecodeChunk / kDecodeBlock; ++it) {
        const uint64_t i = begin + uint64_t(it) * kDecodeBlock + threadIdx.x;
        const bool live = i < P.n;
        int st = kAnswered;
        int reason = -1;
        mgk_hit_t h;
        if (live) {
            float cand = 0.0f;
            st = decode_one(P, C, i, h, reason, cand);
            const bool judge = P.judge_stride && (i % P.judge_stride) == 0;
            if (judge && st == kMino) {
                // judge sample left to k_sur_mino, which records a disk path's
                // answer and queues the path; a captured or escaping path's
                // answer is recorded here
                if (cand < 0.0f) P.pred[i / P.judge_stride] = h;
            } else if (judge) {
                // judge sample: the answer recorded, the path traced as well
                // (a deferred one's answer is recorded by k_sur_mino)
                mgk_hit_t pr = h;
                if (st != kAnswered) pr.cls = kNotAnswered;
                P.pred[i / P.judge_stride] = pr;
                if (st == kAnswered) st = kTrace;
            } else if (st == kAnswered || (st == kMino && cand < 0.0f)) {
                P.hits[i] = h;   // a disk path's is written by k_sur_mino
            }
            if (P.cand) P.cand[i] = cand;
            if (st == kDefer) P.reason[i] = uint8_t(reason);
            if (st == kTrace) {
                // traced: its initial conditions go to HBM
                double y[10];
                ray_ic_fp64(P.G, C, i, y);
#pragma unroll
                for (int k = 0; k < 10; ++k) P.rays[k * P.n + i] = y[k];
            }
        }
        // a deferred path's reason counts only if k_sur_mino traces it
        for (int c = C_CONF; c <= C_REASON_LAST; ++c) {
            const uint64_t m = __ballot(st == kTrace && reason == c);
            if (m != 0 && lane == 0) atomicAdd(&reasons[c], uint32_t(__popcll(m)));
        }
        wave_append(live && st == kTrace, uint32_t(i), lane, lt_mask, &n_local[0], qbuf);
    }
    __syncthreads();
    const uint32_t n_q = n_local[0];
    if (threadIdx.x == 0) base_s[0] = n_q ? atomicAdd(&P.cnt[C_QUEUE], (unsigned long long)n_q) : 0ull;
    if (threadIdx.x >= C_CONF && threadIdx.x <= C_REASON_LAST && reasons[threadIdx.x])
        atomicAdd(&P.cnt[threadIdx.x], (unsigned long long)reasons[threadIdx.x]);
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < n_q; k += kDecodeBlock) P.queue[base_s[0] + k] = qbuf[k];
}

// A path's covariant momentum p and u_obs . p for the Mino-time check: the
// steps of ray_ic_fp64 (sensor point, normalised direction, l = d / (|d| L),
// p = M l) with the camera's divisions hoisted into mi, reciprocals for the
// path's own, contraction on and 32-bit index arithmetic (a frame holds fewer
// than 2^32 paths).  Within a few ulp of ray_ic_fp64's, which the traced
// paths keep: the integrals need nothing closer.
__device__ __forceinline__ void mino_ic(const DecodeParams& P, const IcCamera& C, uint32_t i, double (&p)[4], double& con) {
    const MinoIc& F = P.mi;
    const uint32_t pixel = i % F.npix, sample = F.sample_begin + i / F.npix;
    const uint32_t px = pixel % F.W, py = pixel / F.W;
    float u[4];
    path_uniforms(P.G.seed, P.G.camera_id, pixel, sample, 0, u);
    double a = double(px) * F.inv_w - 0.5, b = double(py) * F.inv_h - 0.5;
    if (F.wide) b *= F.aspect;
    else a *= F.aspect;
    const double sx = a * F.sw_m + (double(u[0]) - 0.5) * F.side;
    const double sy = b * F.sw_m + (double(u[1]) - 0.5) * F.side;
    const double d2 = sx * sx + sy * sy + F.f_m * F.f_m;
    double y = __builtin_amdgcn_rsq(d2);
    y = y * fma(-0.5 * d2 * y, y, 1.5);   // 1 / |d| to ~1e-15
    const double k = -y * F.inv_l;
    const double l[4] = {sx * k, sy * k, F.f_m * k, F.inv_l};
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) p[mu] = C.M[mu][0] * l[0] + C.M[mu][1] * l[1] + C.M[mu][2] * l[2] + C.M[mu][3] * l[3];
    con = C.w[0] * l[0] + C.w[1] * l[1] + C.w[2] * l[2] + C.w[3] * l[3];
}

// The Mino-time check of the confident paths (fallback.analytic): a path
// whose class the integrals confirm keeps the decode's answer, a disk path's
// with the radius they give; the others are traced (reason C_ANALYTIC).  The
// paths the decode deferred (the network not trusted) take the same integrals
// seeded by the network's radius: one they find captured or on the disk is
// answered; an escape (its direction would be the network's) or an unresolved
// path is traced under the reason the decode gave it.  A judge path is traced
// as well, its answer recorded in pred or marked unanswered.  Inside the
// decode the integrals' registers held it to two waves per SIMD, with spills,
// and a loop over several paths per thread kept loop-invariant constants in
// vector registers (246).  One path per thread, it takes 91 registers (five
// waves per SIMD); held to six (80 registers, a few bytes spilled) it runs
// 0.3 ms faster per 1080p frame, and eight (64, more spills) gain nothing
// more.
__global__ void __launch_bounds__(kDecodeBlock, 6) k_sur_mino(DecodeParams P) {
    const IcCamera C = *P.ic;   // uniform, read before any store: scalar loads into SGPRs
    const uint32_t lane = __lane_id();
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint32_t i = blockIdx.x * kDecodeBlock + threadIdx.x;
    const float c = i < P.n ? P.cand[i] : 0.0f;
    bool fail = false, trace = false;
    int reason = -1;
    if (c != 0.0f) {
        double p[4], con;
        mino_ic(P, C, i, p, con);
        const double q[3] = {C.q[0], C.q[1], C.q[2]};
        const double pt = p[3], pphi = q[2] * p[0] - q[0] * p[2];
        const bool deferred = c <= kCandDeferred;
        const double seed = deferred ? double(kCandDeferred - c) : (c > 0.0f ? double(c) : -1.0);
        const MinoResult mr = mino_disk(mino_ray(q, p, P.a_eq), P.dk, seed);
        const bool judge = P.judge_stride && (i % P.judge_stride) == 0;
        bool ok;
        if (deferred) {
            ok = mr.kind == kMinoCaptured || mr.kind == kMinoDisk;
            if (!ok) reason = P.reason[i];
        } else {
            ok = mr.kind == (c > 0.0f ? kMinoDisk : (c == kCandCaptured ? kMinoCaptured : kMinoEscapes));
            fail = !ok;
        }
        if (ok && (c > 0.0f || deferred)) {
            mgk_hit_t h{MGK_HIT_HORIZON, 0.0f, 0.0f, 0.0f};
            if (mr.kind == kMinoDisk) disk_hit(P, pt, pphi, con, mr.r, h);
            if (judge) P.pred[i / P.judge_stride] = h;
            else P.hits[i] = h;
        }
        if (!ok && judge) P.pred[i / P.judge_stride] = mgk_hit_t{kNotAnswered, 0.0f, 0.0f, 0.0f};
        trace = !ok || judge;
        if (trace) {
            // its initial conditions go to HBM (formed again, so that they
            // need not stay live through the integrals)
            double z[10];
            ray_ic_fp64(P.G, C, i, z);
#pragma unroll
            for (int k = 0; k < 10; ++k) P.rays[k * P.n + i] = z[k];
        }
    }
    const uint64_t mf = __ballot(fail);
    if (mf != 0 && lane == 0) atomicAdd(&P.cnt[C_ANALYTIC], (unsigned long long)__popcll(mf));
    if (__ballot(reason >= 0) != 0) {
        for (int r = C_CONF; r <= C_REASON_LAST; ++r) {
            const uint64_t mr = __ballot(reason == r);
            if (mr != 0 && lane == 0) atomicAdd(&P.cnt[r], (unsigned long long)__popcll(mr));
        }
    }
    const uint64_t m = __ballot(trace);
    if (m != 0) {
        unsigned long long off = 0;
        if (lane == 0) off = atomicAdd(&P.cnt[C_QUEUE], (unsigned long long)__popcll(m));
        off = __shfl(off, 0);
        if (trace) P.queue[off + __popcll(m & lt_mask)] = i;
    }
}

__device__ inline void atomic_max_double(unsigned long long* p, double v) {
    // non-negative doubles order like their bit patterns
    atomicMax(p, (unsigned long long)__double_as_longlong(v));
}

__global__ void __launch_bounds__(256) k_sur_judge(const mgk_hit_t* pred, const mgk_hit_t* hits, uint64_t n_slots, uint32_t stride,
                                                   unsigned long long* cnt) {
    const uint64_t j = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= n_slots) return;
    const mgk_hit_t p = pred[j];
    if (p.cls == kNotAnswered) return;
    const mgk_hit_t h = hits[j * stride];
    atomicAdd(&cnt[C_JUDGE], 1ull);
    if (p.cls != h.cls) return;
    atomicAdd(&cnt[C_AGREE], 1ull);
    double* sums = reinterpret_cast<double*>(cnt);
    if (h.cls == MGK_HIT_DISK) {
        atomicAdd(&cnt[C_JDISK], 1ull);
        atomicAdd(&sums[C_SUM_R], fabs(double(p.a0) - double(h.a0)) / fabs(double(h.a0)));
        atomicAdd(&sums[C_SUM_G], fabs(double(p.g) - double(h.g)) / fabs(double(h.g)));
    } else if (h.cls == MGK_HIT_ESCAPE) {
        atomicAdd(&cnt[C_JESC], 1ull);
        const double sp = sqrt(fmax(0.0, 1.0 - double(p.a0) * p.a0)), sh = sqrt(fmax(0.0, 1.0 - double(h.a0) * h.a0));
        const double dx = sp * cos(double(p.a1)) - sh * cos(double(h.a1)), dy = double(p.a0) - double(h.a0);
        const double dz = sp * sin(double(p.a1)) - sh * sin(double(h.a1));
        const double chord = sqrt(dx * dx + dy * dy + dz * dz);
        const double ang = 2.0 * asin(fmin(1.0, 0.5 * chord));
        atomicAdd(&sums[C_SUM_DIR], ang);
        atomic_max_double(&cnt[C_MAX_DIR], ang);
    }
}

// Critical curve by bisection along each screen angle: captured at rho = 0,
// not captured at rho = s_max, traced in FP64 without disk or stars.
struct CritParams {
    SurrogateRanges R;
    double L, a, r, th_deg;
    int time_reversed;
    int mode;                    // 0: rho = 0, 1: rho = hi, 2: midpoint
    double* lo;
    double* hi;
    int* ok;
    double* rays;
};

__global__ void k_crit_rays(CritParams P) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= kCritN) return;
    const double psi = -3.14159265358979323846 + 2.0 * 3.14159265358979323846 * double(k) / double(kCritN);
    const double rho = P.mode == 0 ? 0.0 : (P.mode == 1 ? P.hi[k] : 0.5 * (P.lo[k] + P.hi[k]));
    write_feature_ray(P.R, P.L, P.time_reversed, P.a, P.r, P.th_deg, rho * cos(psi), rho * sin(psi), P.rays, kCritN, uint64_t(k),
                      nullptr);
}

__global__ void k_crit_update(CritParams P, const mgk_hit_t* hits, float* table) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= kCritN) return;
    // a ray that orbits until the step limit counts as captured
    const bool captured = hits[k].cls == MGK_HIT_HORIZON || hits[k].cls == MGK_HIT_STEP_LIMIT;
    if (P.mode == 0) {
        P.ok[k] = captured ? 1 : 0;
    } else if (P.mode == 1) {
        P.ok[k] = P.ok[k] && !captured;
    } else if (P.mode == 2) {
        const double mid = 0.5 * (P.lo[k] + P.hi[k]);
        if (captured) P.lo[k] = mid; else P.hi[k] = mid;
    }
    if (table) table[k] = P.ok[k] ? float(0.5 * (P.lo[k] + P.hi[k])) : -1.0f;
}

__global__ void k_crit_init(double* lo, double* hi, double s_max) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= kCritN) return;
    lo[k] = 0.0;
    hi[k] = s_max;
}

__global__ void k_fill_u64(unsigned long long* p, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = 0ull;
}

double kerr_r(double x, double y, double z, double a) {   // Kerr-Schild r with the spin along +y
    const double rho2 = x * x + y * y + z * z - a * a;
    return std::sqrt(0.5 * (rho2 + std::sqrt(rho2 * rho2 + 4.0 * a * a * y * y)));
}

double isco(double a) {   // prograde ISCO (Bardeen, Press & Teukolsky 1972)
    a = std::fabs(a);
    const double z1 = 1.0 + std::cbrt(1.0 - a * a) * (std::cbrt(1.0 + a) + std::cbrt(1.0 - a));
    const double z2 = std::sqrt(3.0 * a * a + z1 * z1);
    return 3.0 + z2 - std::sqrt((3.0 - z1) * (3.0 + z1 + 2.0 * z2));
}

e_magik_result_types ensure_critical_curve(mgk_surrogate* s, mgk_context_handle ctx, const mgk_integrator_t* integ,
                                           const mgk_scene_t* scene, double L, double r, double th_deg, hipStream_t st) {
    const double key[6] = {double(integ->spin), r, th_deg, s->ranges.s_max, double(scene->celestial_radius_m),
                           double(integ->time_reversed)};
    if (std::memcmp(key, s->crit_key, sizeof(key)) == 0) return MAGIK_SUCCESS;
    HIP_TRY(s->crit.reserve(kCritN * sizeof(float)));
    HIP_TRY(s->crit_lo.reserve(kCritN * sizeof(double)));
    HIP_TRY(s->crit_hi.reserve(kCritN * sizeof(double)));
    HIP_TRY(s->crit_ok.reserve(kCritN * sizeof(int)));
    HIP_TRY(s->crit_rays.reserve(size_t(kCritN) * 11 * sizeof(double)));
    HIP_TRY(s->crit_hits.reserve(kCritN * sizeof(mgk_hit_t)));
    mgk_integrator_t it = *integ;
    it.precision = MGK_PRECISION_FP64;
    mgk_scene_t sc{};
    sc.celestial_radius_m = scene->celestial_radius_m;
    sc.flags = 0;
    CritParams P;
    P.R = s->ranges;
    P.L = L;
    P.a = integ->spin;
    P.r = r;
    P.th_deg = th_deg;
    P.time_reversed = integ->time_reversed ? 1 : 0;
    P.lo = s->crit_lo.as<double>();
    P.hi = s->crit_hi.as<double>();
    P.ok = s->crit_ok.as<int>();
    P.rays = s->crit_rays.as<double>();
    const dim3 grid(kCritN / 256), block(256);
    hipLaunchKernelGGL(k_crit_init, grid, block, 0, st, P.lo, P.hi, s->ranges.s_max);
    HIP_TRY(hipGetLastError());
    const float origin[3] = {0.0f, 0.0f, 0.0f};
    const int iters = 22;   // s_max / 2^22 ~ 4e-6 M
    for (int step = 0; step < iters + 2; ++step) {
        P.mode = step < 2 ? step : 2;
        hipLaunchKernelGGL(k_crit_rays, grid, block, 0, st, P);
        HIP_TRY(hipGetLastError());
        e_magik_result_types rr = mgk_trace_rays_per_ray(ctx, &it, &sc, origin, P.rays, kCritN, kCritN, s->crit_hits.as<mgk_hit_t>(), st);
        if (rr != MAGIK_SUCCESS) return rr;
        hipLaunchKernelGGL(k_crit_update, grid, block, 0, st, P, s->crit_hits.as<const mgk_hit_t>(),
                           step == iters + 1 ? s->crit.as<float>() : nullptr);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipStreamSynchronize(st));
    std::memcpy(s->crit_key, key, sizeof(key));
    return MAGIK_SUCCESS;
}

float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0.0f;
    return hipEventElapsedTime(&ms, a, b) == hipSuccess ? ms : 0.0f;
}

}  // namespace

extern "C" {

MGK_API e_magik_result_types mgk_render_frame_surrogate(mgk_context_handle ctx, mgk_surrogate_handle s, const mgk_integrator_t* integ,
                                                        const mgk_scene_t* scene, const mgk_spectral_t* spec, const mgk_camera_t* cam,
                                                        uint64_t seed, uint32_t spp, const mgk_fallback_t* fb, float* d_xyz,
                                                        mgk_hit_t* d_hits, mgk_surrogate_stats_t* stats, void* stream) {
    if (!ctx || !s || !integ || !scene || !spec || !cam || !fb || !d_xyz) return MAGIK_ERROR_INVALID_POINTER;
    if (integ->metric_type != KERR || integ->precision == MGK_PRECISION_FP32_REFERENCE || spp == 0) return MGK_ERROR_INVALID_ARGUMENT;
    if (cam->observer_velocity[0] != 0.0f || cam->observer_velocity[1] != 0.0f || cam->observer_velocity[2] != 0.0f)
        return MGK_ERROR_INVALID_ARGUMENT;
    if (!(fb->min_confidence >= 0.0f) || !(fb->critical_margin >= 0.0f) || !(fb->direction_margin >= 0.0f) ||
        !(fb->disk_edge_margin >= 0.0f))
        return MGK_ERROR_INVALID_ARGUMENT;
    const SurrogateRanges& R = s->ranges;
    const double a = integ->spin;
    if (a < R.a_lo || a > R.a_hi) return MGK_ERROR_INVALID_ARGUMENT;
    // the disk the network was trained on: r_ISCO(a) .. disk_r_max
    if ((scene->flags & MGK_SCENE_DISK) &&
        (std::fabs(double(scene->disk_r_in) - isco(a)) > 1e-4 * isco(a) || std::fabs(double(scene->disk_r_out) - R.disk_r_max) > 1e-4 * R.disk_r_max))
        return MGK_ERROR_INVALID_ARGUMENT;
    if (scene->flags & (MGK_SCENE_SPHERE | MGK_SCENE_BACKGROUND)) return MGK_ERROR_INVALID_ARGUMENT;
    // lens point of the camera in units of M (the rays' origin) and its Boyer-Lindquist (r, theta)
    const double L = double(kG * integ->mass_kg / (kC * kC));
    const double f_m = double(cam->focal_length_mm) / 1000.0;
    const float* m = cam->orientation;
    const double o[3] = {(double(m[2]) * -f_m + double(cam->position[0])) / L, (double(m[5]) * -f_m + double(cam->position[1])) / L,
                         (double(m[8]) * -f_m + double(cam->position[2])) / L};
    const double r_lens = kerr_r(o[0], o[1], o[2], a);
    const double th_lens = std::acos(std::max(-1.0, std::min(1.0, o[1] / r_lens)));
    const double th_deg = th_lens * 180.0 / 3.14159265358979323846;
    if (r_lens < R.r_lo || r_lens > R.r_hi || th_deg < R.th_lo || th_deg > R.th_hi) return MGK_ERROR_INVALID_ARGUMENT;
    // the camera must be the Boyer-Lindquist camera of that point (looking at the hole, +y up)
    Camframe f;
    bl_camera_frame(a, r_lens, th_lens, f);
    for (int row = 0; row < 3; ++row) {
        if (std::fabs(m[3 * row + 0] - f.X[row]) > 1e-4 || std::fabs(m[3 * row + 1] - f.Y[row]) > 1e-4 ||
            std::fabs(m[3 * row + 2] - f.Z[row]) > 1e-4)
            return MGK_ERROR_INVALID_ARGUMENT;
    }
    if (std::fabs(o[0]) > 1e-5 * r_lens) return MGK_ERROR_INVALID_ARGUMENT;
    HIP_TRY(hipSetDevice(s->device));
    hipStream_t st = hipStream_t(stream);
    for (hipEvent_t& e : s->ev)
        if (!e) HIP_TRY(hipEventCreate(&e));

    const uint64_t n = uint64_t(cam->x_resolution) * cam->y_resolution * spp;
    const uint64_t n_slots = fb->judge_stride ? (n + fb->judge_stride - 1) / fb->judge_stride : 0;
    HIP_TRY(s->fr_rays.reserve(n * 10 * sizeof(double)));
    HIP_TRY(s->fr_y.reserve(n * kMlpOut * sizeof(float)));
    if (!d_hits) HIP_TRY(s->fr_hits.reserve(n * sizeof(mgk_hit_t)));
    mgk_hit_t* hits = d_hits ? d_hits : s->fr_hits.as<mgk_hit_t>();
    HIP_TRY(s->fr_queue.reserve(n * sizeof(uint32_t)));
    if (fb->analytic && !fb->trace_disk) HIP_TRY(s->fr_reason.reserve(n));
    if (fb->analytic) HIP_TRY(s->fr_cand.reserve(n * sizeof(float)));
    HIP_TRY(s->fr_pred.reserve(std::max<uint64_t>(1, n_slots) * sizeof(mgk_hit_t)));
    HIP_TRY(s->fr_counters.reserve(C_N * sizeof(unsigned long long)));
    if (n > 0xFFFFFFFFull) return MGK_ERROR_CAPACITY_EXCEEDED;

    e_magik_result_types r = ensure_critical_curve(s, ctx, integ, scene, L, r_lens, th_deg, st);
    if (r != MAGIK_SUCCESS) return r;
    StarView sv;
    const bool star_test = (scene->flags & MGK_SCENE_STARS) && scene->n_star > 0 && fb->direction_margin > 0.0f;
    std::memset(&sv, 0, sizeof(sv));
    if (star_test && (r = mgk_star_band_view(ctx, scene, fb->direction_margin, &sv)) != MAGIK_SUCCESS) return r;
    unsigned long long* cnt = s->fr_counters.as<unsigned long long>();

    HIP_TRY(hipEventRecord(s->ev[0], st));
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, st, cnt, int(C_N));
    HIP_TRY(hipGetLastError());
    RayGenParams G;
    const IcCamera* d_ic = nullptr;
    if ((r = mgk_camera_constants(ctx, integ, cam, seed, spp, &G, &d_ic, st)) != MAGIK_SUCCESS) return r;
    G.rays = s->fr_rays.ptr;
    HIP_TRY(hipEventRecord(s->ev[1], st));

    InferFromCamera src;
    FeatureParams& F = src.P;
    F.R = R;
    F.W = cam->x_resolution; F.H = cam->y_resolution; F.camera_id = cam->camera_id; F.spp = spp;
    F.seed = seed; F.n = n;
    F.sensor_width_mm = cam->sensor_width_mm; F.focal_mm = cam->focal_length_mm;
    F.a = a; F.r_lens = r_lens; F.th_lens_deg = th_deg;
    {
        const uint32_t W = F.W, H = F.H;
        const double sw = double(cam->sensor_width_mm);
        const double inv_s = sw / std::sqrt(double(float(W * W + H * H)));
        const double area = (double(float(W)) * inv_s) * (double(float(H)) * inv_s) / 1e6;
        const double ratio = double(float(H)) / double(float(W));
        F.side = float(std::sqrt(area / double(float(W * H))));
        F.inv_w = float(1.0 / double(W));
        F.inv_h = float(1.0 / double(H));
        F.wide = ratio < 1.0 ? 1u : 0u;
        F.aspect = float(ratio < 1.0 ? ratio : 1.0 / ratio);
        F.sw_m = float(sw / 1000.0);
        F.k_screen = float(r_lens / (double(cam->focal_length_mm) / 1000.0) / R.s_max);
    }
    {
        float x[kMlpInPad];
        normalise_features(R, a, r_lens, th_deg, 0.0, 0.0, x);
        src.x0 = x[0]; src.x1 = x[1]; src.x2 = x[2];
    }
    InferArgs ia;
    if ((r = surrogate_infer_args(s, ia, st)) != MAGIK_SUCCESS) return r;
    ia.X = nullptr;
    ia.x_stride = 0;
    ia.n = int64_t(n);
    ia.Y = s->fr_y.as<float>();
    hipLaunchKernelGGL(k_mlp_infer_res<InferFromCamera>, dim3(infer_grid(ia.n)), dim3(kInferBlock), 0, st, ia, src);
    HIP_TRY(hipGetLastError());

    DecodeParams D;
    std::memset(&D, 0, sizeof(D));
    D.R = R;
    D.n = n;
    D.Y = s->fr_y.as<const float>();
    D.G = G;
    D.ic = d_ic;
    {
        // the per-camera parts of sensor_point and ray_ic_fp64, in FP64
        MinoIc& F = D.mi;
        const uint32_t W = cam->x_resolution, H = cam->y_resolution;
        const double sw = double(cam->sensor_width_mm);
        const double inv_s = sw / std::sqrt(double(float(W * W + H * H)));
        const double area = (double(float(W)) * inv_s) * (double(float(H)) * inv_s) / 1e6;
        F.side = std::sqrt(area / double(float(W * H)));
        F.inv_w = 1.0 / double(W);
        F.inv_h = 1.0 / double(H);
        const double ratio = double(float(H)) / double(float(W));
        F.wide = ratio < 1.0 ? 1u : 0u;
        F.aspect = ratio < 1.0 ? ratio : 1.0 / ratio;
        F.sw_m = sw / 1000.0;
        F.f_m = double(cam->focal_length_mm) / 1000.0;
        F.inv_l = 1.0 / double(G.length_scale);
        F.npix = W * H;
        F.W = W;
        F.sample_begin = G.sample_begin;
    }
    D.rays = s->fr_rays.as<double>();
    D.hits = hits;
    D.queue = s->fr_queue.as<uint32_t>();
    D.reason = s->fr_reason.as<uint8_t>();
    D.cand = fb->analytic ? s->fr_cand.as<float>() : nullptr;
    D.pred = s->fr_pred.as<mgk_hit_t>();
    D.cnt = cnt;
    D.crit = s->crit.as<const float>();
    D.stars = sv;
    D.star_test = star_test ? 1 : 0;
    if (star_test) {
        const double rin = std::max(0.0, double(scene->star_radius_rad) - fb->direction_margin);
        const double rout = double(scene->star_radius_rad) + fb->direction_margin;
        D.star_chord2_in = float(4.0 * std::sin(0.5 * rin) * std::sin(0.5 * rin));
        D.star_chord2_out = float(4.0 * std::sin(0.5 * rout) * std::sin(0.5 * rout));
    }
    D.flags = scene->flags;
    D.a_eq = integ->time_reversed ? -a : a;
    D.r_horizon = 1.0 + std::sqrt(std::max(0.0, 1.0 - a * a));
    D.disk_r_in = scene->disk_r_in;
    D.disk_r_out = scene->disk_r_out;
    D.dk = mino_disk_radii(D.r_horizon, D.disk_r_in, D.disk_r_out);
    D.min_conf = fb->min_confidence;
    D.crit_margin = fb->critical_margin;
    D.edge_margin = fb->disk_edge_margin;
    D.judge_stride = fb->judge_stride;
    D.trace_disk = fb->trace_disk;
    D.analytic = fb->analytic;
    hipLaunchKernelGGL(k_sur_decode, dim3(unsigned((n + kDecodeChunk - 1) / kDecodeChunk)), dim3(kDecodeBlock), 0, st, D);
    HIP_TRY(hipGetLastError());
    if (D.analytic) {
        hipLaunchKernelGGL(k_sur_mino, dim3(unsigned((n + kDecodeBlock - 1) / kDecodeBlock)), dim3(kDecodeBlock), 0, st, D);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipEventRecord(s->ev[2], st));

    unsigned long long nq = 0;
    HIP_TRY(hipMemcpyAsync(&nq, cnt + C_QUEUE, sizeof(nq), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (nq > 0 && (r = mgk_trace_rays(ctx, integ, scene, cam->position, s->fr_rays.ptr, n, nq, s->fr_queue.as<uint32_t>(),
                                      hits, nullptr, st)) != MAGIK_SUCCESS)
        return r;
    if (n_slots) {
        hipLaunchKernelGGL(k_sur_judge, dim3(unsigned((n_slots + 255) / 256)), dim3(256), 0, st, s->fr_pred.as<const mgk_hit_t>(),
                           hits, n_slots, fb->judge_stride, cnt);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipEventRecord(s->ev[3], st));
    if ((r = mgk_shade(ctx, scene, spec, cam, seed, 0, spp, hits, d_xyz, st)) != MAGIK_SUCCESS) return r;
    HIP_TRY(hipEventRecord(s->ev[4], st));

    if (stats) {
        HIP_TRY(hipStreamSynchronize(st));
        unsigned long long h[C_N];
        HIP_TRY(hipMemcpy(h, cnt, sizeof(h), hipMemcpyDeviceToHost));
        double dsum[C_N];
        std::memcpy(dsum, h, sizeof(h));
        std::memset(stats, 0, sizeof(*stats));
        stats->n_paths = n;
        stats->n_fallback = h[C_QUEUE];
        stats->n_fallback_confidence = h[C_CONF];
        stats->n_fallback_critical = h[C_CRIT];
        stats->n_fallback_disk = h[C_DISK];
        stats->n_fallback_star = h[C_STAR];
        stats->n_fallback_domain = h[C_DOMAIN];
        stats->n_fallback_analytic = h[C_ANALYTIC];
        stats->n_judge = h[C_JUDGE];
        stats->judge_class_agree = h[C_AGREE];
        stats->judge_n_disk = h[C_JDISK];
        stats->judge_n_escape = h[C_JESC];
        stats->judge_mean_rel_r = h[C_JDISK] ? dsum[C_SUM_R] / double(h[C_JDISK]) : 0.0;
        stats->judge_mean_rel_g = h[C_JDISK] ? dsum[C_SUM_G] / double(h[C_JDISK]) : 0.0;
        stats->judge_mean_dir = h[C_JESC] ? dsum[C_SUM_DIR] / double(h[C_JESC]) : 0.0;
        stats->judge_max_dir = dsum[C_MAX_DIR];
        mgk_trace_stats_t ts;
        if (nq > 0 && mgk_last_trace_stats(ctx, &ts) == MAGIK_SUCCESS) stats->n_trials = ts.n_trials;
        stats->ms_raygen = elapsed(s->ev[0], s->ev[1]);
        stats->ms_mlp = elapsed(s->ev[1], s->ev[2]);
        stats->ms_trace = elapsed(s->ev[2], s->ev[3]);
        stats->ms_shade = elapsed(s->ev[3], s->ev[4]);
        stats->ms_total = elapsed(s->ev[0], s->ev[4]);
    }
    return MAGIK_SUCCESS;
}

MGK_API e_magik_result_types mgk_mino_disk(const mgk_integrator_t* integ, const mgk_scene_t* scene, const double* rays,
                                           uint64_t ray_stride, uint64_t n, const double* r_seed, int32_t* kind, double* r_hit) {
    if (!integ || !scene || (n > 0 && (!rays || !kind || !r_hit))) return MAGIK_ERROR_INVALID_POINTER;
    if (integ->metric_type != KERR || n > ray_stride) return MGK_ERROR_INVALID_ARGUMENT;
    if (!(scene->disk_r_out > scene->disk_r_in) || !(scene->disk_r_in > 0.0f)) return MGK_ERROR_INVALID_ARGUMENT;
    const double a = integ->spin;
    if (!(std::fabs(a) < 1.0)) return MGK_ERROR_INVALID_ARGUMENT;
    const double a_eq = integ->time_reversed ? -a : a;   // as the frame's decode takes it
    const double r_h = 1.0 + std::sqrt(1.0 - a * a);
    for (uint64_t i = 0; i < n; ++i) {
        const double q[3] = {rays[0 * ray_stride + i], rays[1 * ray_stride + i], rays[2 * ray_stride + i]};
        const double p[4] = {rays[4 * ray_stride + i], rays[5 * ray_stride + i], rays[6 * ray_stride + i], rays[7 * ray_stride + i]};
        const MinoResult mr = mino_disk(mino_ray(q, p, a_eq), r_h, scene->disk_r_in, scene->disk_r_out, r_seed ? r_seed[i] : -1.0);
        kind[i] = mr.kind;
        r_hit[i] = mr.r;
    }
    return MAGIK_SUCCESS;
}

}  // extern "C"
