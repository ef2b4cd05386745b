// Feature space, device state and launch arguments of the surrogate
// (mgk_surrogate.hip: training; mgk_surrogate_frame.hip: the render path).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>

#include "magik_amd.h"
#include "mgk_device.hpp"
#include "mgk_mlp.hpp"

namespace mgk {

// Feature ranges (normalised to [-1, 1]) and the normalisation of the disk
// radius target (ln r mapped to [-1, 1] over [1, disk_r_max]).
struct SurrogateRanges {
    double a_lo, a_hi, r_lo, r_hi, th_lo, th_hi, s_max;
    double lnr_mid, lnr_half;
    double disk_r_max;
};

inline SurrogateRanges make_ranges(const mgk_surrogate_ranges_t& in) {
    SurrogateRanges R;
    R.a_lo = in.spin_min; R.a_hi = in.spin_max;
    R.r_lo = in.r_min; R.r_hi = in.r_max;
    R.th_lo = in.theta_min_deg; R.th_hi = in.theta_max_deg;
    R.s_max = in.screen_max;
    const double lo = std::log(1.0), hi = std::log(double(in.disk_r_max));
    R.lnr_mid = 0.5 * (lo + hi);
    R.lnr_half = 0.5 * (hi - lo);
    R.disk_r_max = in.disk_r_max;
    return R;
}

__host__ __device__ inline void normalise_features(const SurrogateRanges& R, double a, double r, double th_deg, double al, double be,
                                                   float* x) {
    auto nrm = [](double v, double lo, double hi) { return hi > lo ? float(2.0 * (v - lo) / (hi - lo) - 1.0) : 0.0f; };
    x[0] = nrm(a, R.a_lo, R.a_hi);
    x[1] = nrm(r, R.r_lo, R.r_hi);
    x[2] = nrm(th_deg, R.th_lo, R.th_hi);
    x[3] = float(al / R.s_max);
    x[4] = float(be / R.s_max);
    for (int k = 5; k < kMlpInPad; ++k) x[k] = 0.0f;
}

// Normalised features of one ray and its FP64 initial conditions: a camera
// at Boyer-Lindquist (r, theta_deg) of a hole with physical spin a, screen
// point (alpha, beta).  rays (11 x stride doubles: q, p, h0, u.p, metric spin)
// and X (16 floats per row) are each optional.
__device__ inline void write_feature_ray(const SurrogateRanges& R, double L, int time_reversed, double a, double r, double th_deg,
                                         double al, double be, double* rays, uint64_t stride, uint64_t i, float* X) {
    const double th = th_deg * 3.14159265358979323846 / 180.0;
    Camframe f;
    bl_camera_frame(a, r, th, f);
    double d[3] = {al / r, be / r, -1.0};
    const double dl = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    d[0] /= dl; d[1] /= dl; d[2] /= dl;
    const double a_eq = time_reversed ? -a : a;
    St<double> y;
    double con;
    V4<double> u;
    ic_from_frame<double>(f, a_eq, L, d, y, con, u);
    if (rays) {
        rays[0 * stride + i] = y.qx; rays[1 * stride + i] = y.qy; rays[2 * stride + i] = y.qz; rays[3 * stride + i] = y.qt;
        rays[4 * stride + i] = y.px; rays[5 * stride + i] = y.py; rays[6 * stride + i] = y.pz; rays[7 * stride + i] = y.pt;
        rays[8 * stride + i] = 1.0;
        rays[9 * stride + i] = con;
        rays[10 * stride + i] = a_eq;
    }
    if (X) normalise_features(R, a, r, th_deg, al, be, X + i * kMlpInPad);
}

struct TrainGen {
    SurrogateRanges ranges;
    double L;
    int time_reversed;
    uint64_t seed, n;
    double* rays;
    uint64_t stride;
    float* X;
};

// Device allocation that grows on demand and is released with its owner.
struct SBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    SBuf() = default;
    SBuf(const SBuf&) = delete;
    SBuf& operator=(const SBuf&) = delete;
    ~SBuf() { release(); }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    hipError_t reserve(size_t n) {
        if (n <= bytes) return hipSuccess;
        release();
        hipError_t e = hipMalloc(&ptr, n);
        if (e == hipSuccess) bytes = n;
        return e;
    }
    template <class T> T* as() const { return reinterpret_cast<T*>(ptr); }
};

constexpr int kCritN = 1024;   // angles of the critical-curve table

}  // namespace mgk

struct mgk_surrogate {
    int device = 0;
    float* W[mgk::kMlpLayers] = {};
    float* b[mgk::kMlpLayers] = {};
    float* mW[mgk::kMlpLayers] = {};
    float* vW[mgk::kMlpLayers] = {};
    float* mb[mgk::kMlpLayers] = {};
    float* vb[mgk::kMlpLayers] = {};
    float* gW[mgk::kMlpLayers] = {};
    float* gb[mgk::kMlpLayers] = {};
    __bf16* Wf[mgk::kMlpLayers] = {};   // forward A fragments (inference and training)
    __bf16* Wg[mgk::kMlpLayers] = {};   // backward A fragments (training; layers 1..4)
    bool frag_dirty = true;
    int t = 0;   // Adam step
    // training scratch: the layer inputs h_l and output gradients g_l of a
    // batch, feature-major with rows of cap_pad rays
    int cap_pad = 0;
    __bf16* HT[mgk::kMlpLayers] = {};
    __bf16* GT[mgk::kMlpLayers] = {};
    float* part = nullptr;   // per 32-ray tile: partial bias-gradient and loss sums
    float* gWp[mgk::kMlpLayers] = {};   // partial weight gradients of the batch splits
    float* sums = nullptr;
    mgk::SurrogateRanges ranges{};
    // render-path buffers
    mgk::SBuf fr_rays, fr_y, fr_hits, fr_queue, fr_pred, fr_counters, fr_reason, fr_cand;
    // critical-curve table (screen radius per angle) and its key
    mgk::SBuf crit, crit_lo, crit_hi, crit_ok, crit_rays, crit_hits;
    double crit_key[6] = {-1, -1, -1, -1, -1, -1};
    hipEvent_t ev[5] = {};
};

namespace mgk {
// The fused inference over n feature rows of x_stride floats (kMlpInPad, the
// public layout, or kFrameXStride, a frame's own).
e_magik_result_types surrogate_forward(mgk_surrogate* s, const float* d_x, int x_stride, uint64_t n, float* d_y, hipStream_t st);
// Refreshes the inference fragments after a weight change (a no-op otherwise)
// and fills the weight and bias pointers of an inference launch.
e_magik_result_types surrogate_infer_args(mgk_surrogate* s, InferArgs& a, hipStream_t st);
}  // namespace mgk

// Internal (not exported): trace rays whose spin is stored per ray (component
// 10), with the escape sphere about each ray's origin and the disk starting at
// each ray's ISCO.
e_magik_result_types mgk_trace_rays_per_ray(mgk_context_handle ctx, const mgk_integrator_t* integ, const mgk_scene_t* scene,
                                            const float origin[3], const double* d_rays, uint64_t stride, uint64_t n,
                                            mgk_hit_t* d_hits, hipStream_t stream);
