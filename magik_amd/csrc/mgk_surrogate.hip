// The learned surrogate of the Kerr tracer: training-ray generation (FP64
// RKF45 targets), training on the bf16 MFMA kernels of mgk_mlp.hip, and
// inference.  The render path that uses it (mgk_render_frame_surrogate) lives
// in mgk_surrogate_frame.hip.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "magik_amd.h"
#include "mgk_device.hpp"
#include "mgk_internal.hpp"
#include "mgk_mlp.hpp"
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

const int* const kK = kLayerK;
const int* const kN = kLayerN;

template <class T> hipError_t alloc(T** p, size_t n) { return hipMalloc(reinterpret_cast<void**>(p), n * sizeof(T)); }

// ---------------------------------------------------------------------------
// Training rays: seeded uniform (a, r_obs, theta_obs, alpha, beta); FP64 ICs
// with the spin in component 10; normalised features.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_training_rays(TrainGen G) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= G.n) return;
    uint32_t c[4] = {uint32_t(i), uint32_t(i >> 32), 0x54524e31u, 0u};
    philox4x32_10(c, uint32_t(G.seed), uint32_t(G.seed >> 32) ^ 0x53555252u);
    uint32_t c2[4] = {uint32_t(i), uint32_t(i >> 32), 0x54524e32u, 0u};
    philox4x32_10(c2, uint32_t(G.seed), uint32_t(G.seed >> 32) ^ 0x53555252u);
    const SurrogateRanges& R = G.ranges;
    double a = R.a_lo + (R.a_hi - R.a_lo) * double(u01(c[0]));
    double r = R.r_lo + (R.r_hi - R.r_lo) * double(u01(c[1]));
    double th = R.th_lo + (R.th_hi - R.th_lo) * double(u01(c[2]));
    double al = R.s_max * (2.0 * double(u01(c[3])) - 1.0);
    double be = R.s_max * (2.0 * double(u01(c2[0])) - 1.0);
    write_feature_ray(G.ranges, G.L, G.time_reversed, a, r, th, al, be, G.rays, G.stride, i, G.X);
}

// Targets from the FP64 hits: class (0 captured / unresolved, 1 disk,
// 2 escape), normalised log r of the disk hit, escape direction.
__global__ void __launch_bounds__(256) k_targets(const mgk_hit_t* hits, uint64_t n, SurrogateRanges R, float* T) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    mgk_hit_t h = hits[i];
    float* t = T + i * kTargetStride;
    for (int k = 0; k < kTargetStride; ++k) t[k] = 0.0f;
    if (h.cls == MGK_HIT_DISK) {
        t[0] = 1.0f;
        t[1] = float((log(double(h.a0)) - R.lnr_mid) / R.lnr_half);
    } else if (h.cls == MGK_HIT_ESCAPE) {
        t[0] = 2.0f;
        t[2] = h.a0;
        t[3] = sinf(h.a1);
        t[4] = cosf(h.a1);
    }
}

}  // namespace

namespace {

void release_scratch(mgk_surrogate* s) {
    for (int l = 0; l < kMlpLayers; ++l) {
        if (s->HT[l]) (void)hipFree(s->HT[l]), s->HT[l] = nullptr;
        if (s->GT[l]) (void)hipFree(s->GT[l]), s->GT[l] = nullptr;
    }
    if (s->part) (void)hipFree(s->part), s->part = nullptr;
    for (int l = 0; l < kMlpLayers; ++l)
        if (s->gWp[l]) (void)hipFree(s->gWp[l]), s->gWp[l] = nullptr;
    s->cap_pad = 0;
}

e_magik_result_types ensure_scratch(mgk_surrogate* s, int n) {
    const int np = train_n_pad(n);
    if (np <= s->cap_pad) return MAGIK_SUCCESS;
    release_scratch(s);
    for (int l = 0; l < kMlpLayers; ++l) {
        HIP_TRY(alloc(&s->HT[l], size_t(kK[l]) * np));
        HIP_TRY(alloc(&s->GT[l], size_t(kN[l]) * np));
    }
    HIP_TRY(alloc(&s->part, size_t(np / 32) * kPartStride));
    for (int l = 0; l < kMlpLayers; ++l) HIP_TRY(alloc(&s->gWp[l], size_t(kDwSplits) * kK[l] * kN[l]));
    s->cap_pad = np;
    return MAGIK_SUCCESS;
}

AdamArgs adam_args(mgk_surrogate* s) {
    AdamArgs a;
    std::memset(&a, 0, sizeof(a));
    for (int l = 0; l < kMlpLayers; ++l) {
        a.W[l] = s->W[l]; a.b[l] = s->b[l];
        a.mW[l] = s->mW[l]; a.vW[l] = s->vW[l]; a.mb[l] = s->mb[l]; a.vb[l] = s->vb[l];
        a.gWp[l] = s->gWp[l]; a.gW[l] = s->gW[l]; a.gb[l] = s->gb[l];
        a.Wf[l] = s->Wf[l]; a.Wg[l] = s->Wg[l];
    }
    return a;
}

// Forward, loss (targets T) or a given dL/dy (dY), backward and the weight
// and bias gradients of one batch, into gW and gb; the loss sums into sums.
e_magik_result_types train_gradients(mgk_surrogate* s, const float* X, const float* T, const float* dY, float* Y, int n, float w_r,
                                     float w_d, hipStream_t st) {
    TrainArgs a;
    std::memset(&a, 0, sizeof(a));
    a.X = X; a.T = T; a.dY = dY; a.Y = Y;
    a.n = n;
    a.n_pad = train_n_pad(n);
    a.w_r = w_r; a.w_d = w_d;
    DwArgs d;
    std::memset(&d, 0, sizeof(d));
    d.n_pad = a.n_pad;
    for (int l = 0; l < kMlpLayers; ++l) {
        a.Wf[l] = s->Wf[l]; a.Wg[l] = s->Wg[l]; a.b[l] = s->b[l];
        a.HT[l] = s->HT[l]; a.GT[l] = s->GT[l];
        d.HT[l] = s->HT[l]; d.GT[l] = s->GT[l]; d.gWp[l] = s->gWp[l]; d.gb[l] = s->gb[l];
    }
    a.part = s->part;
    d.part = s->part;
    d.sums = s->sums;
    HIP_TRY(launch_train_fb(a, st));
    HIP_TRY(launch_train_dw(d, st));
    return MAGIK_SUCCESS;
}

// The bf16 forward and backward fragments of the current master weights.
e_magik_result_types refresh_fragments(mgk_surrogate* s, hipStream_t st) {
    AdamArgs a = adam_args(s);
    a.update = 0;
    HIP_TRY(launch_train_adam(a, 0, st));
    s->frag_dirty = false;
    return MAGIK_SUCCESS;
}

}  // namespace


extern "C" {

MGK_API e_magik_result_types mgk_surrogate_create(int device, uint64_t seed, const mgk_surrogate_ranges_t* ranges,
                                                  mgk_surrogate_handle* out) {
    if (!out || !ranges) return MAGIK_ERROR_INVALID_POINTER;
    *out = nullptr;
    int nd = 0;
    if (hipGetDeviceCount(&nd) != hipSuccess || device < 0 || device >= nd) return MAGIK_INVALID_CUDA_DEVICE;
    if (!(ranges->spin_max >= ranges->spin_min) || !(ranges->r_max > ranges->r_min) || !(ranges->theta_max_deg > ranges->theta_min_deg) ||
        !(ranges->screen_max > 0.0f) || !(ranges->disk_r_max > 1.0f))
        return MGK_ERROR_INVALID_ARGUMENT;
    HIP_TRY(hipSetDevice(device));
    mgk_surrogate* s = new (std::nothrow) mgk_surrogate();
    if (!s) return MAGIK_ERROR_HOST_OUT_OF_MEMORY;
    s->device = device;
    s->ranges = make_ranges(*ranges);
    auto fail = [&](e_magik_result_types r) {
        mgk_surrogate_destroy(s);
        return r;
    };
    for (int l = 0; l < kMlpLayers; ++l) {
        size_t nw = size_t(kK[l]) * kN[l];
        if (alloc(&s->W[l], nw) || alloc(&s->mW[l], nw) || alloc(&s->vW[l], nw) || alloc(&s->gW[l], nw) || alloc(&s->Wg[l], nw) ||
            alloc(&s->Wf[l], nw) || alloc(&s->b[l], size_t(kN[l])) || alloc(&s->mb[l], size_t(kN[l])) ||
            alloc(&s->vb[l], size_t(kN[l])) || alloc(&s->gb[l], size_t(kN[l])))
            return fail(MAGIK_ERROR_DEVICE_ALLOCATION_FAILED);
    }
    if (alloc(&s->sums, 4)) return fail(MAGIK_ERROR_DEVICE_ALLOCATION_FAILED);
    // He-uniform initialisation from Philox; padded input rows and output
    // columns start (and stay) zero.
    std::vector<float> w, z;
    for (int l = 0; l < kMlpLayers; ++l) {
        const int K = kK[l], N = kN[l];
        const int fan_in = l == 0 ? 5 : K;
        const float lim = std::sqrt(6.0f / float(fan_in));
        w.assign(size_t(K) * N, 0.0f);
        for (int k = 0; k < K; ++k)
            for (int n = 0; n < N; ++n) {
                if ((l == 0 && k >= 5) || (l == kMlpLayers - 1 && n >= kMlpOut)) continue;
                uint32_t c[4] = {uint32_t(k), uint32_t(n), uint32_t(l), 0x494e4954u};
                philox4x32_10(c, uint32_t(seed), uint32_t(seed >> 32));
                w[size_t(k) * N + n] = lim * (2.0f * u01(c[0]) - 1.0f) * (l == kMlpLayers - 1 ? 0.1f : 1.0f);
            }
        z.assign(size_t(K) * N, 0.0f);
        if (hipMemcpy(s->W[l], w.data(), w.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(s->mW[l], z.data(), z.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(s->vW[l], z.data(), z.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(s->b[l], z.data(), size_t(N) * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(s->mb[l], z.data(), size_t(N) * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(s->vb[l], z.data(), size_t(N) * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(s->gb[l], z.data(), size_t(N) * 4, hipMemcpyHostToDevice) != hipSuccess)
            return fail(MGK_ERROR_HIP);
    }
    if (refresh_fragments(s, nullptr) != MAGIK_SUCCESS || hipDeviceSynchronize() != hipSuccess) return fail(MGK_ERROR_HIP);
    *out = s;
    return MAGIK_SUCCESS;
}

MGK_API e_magik_result_types mgk_surrogate_destroy(mgk_surrogate_handle s) {
    if (!s) return MAGIK_ERROR_INVALID_POINTER;
    (void)hipSetDevice(s->device);
    (void)hipDeviceSynchronize();
    release_scratch(s);
    for (int l = 0; l < kMlpLayers; ++l)
        for (void* p : {(void*)s->W[l], (void*)s->b[l], (void*)s->mW[l], (void*)s->vW[l], (void*)s->mb[l], (void*)s->vb[l],
                        (void*)s->gW[l], (void*)s->gb[l], (void*)s->Wg[l], (void*)s->Wf[l]})
            if (p) (void)hipFree(p);
    if (s->sums) (void)hipFree(s->sums);
    for (hipEvent_t e : s->ev)
        if (e) (void)hipEventDestroy(e);
    delete s;
    return MAGIK_SUCCESS;
}

MGK_API e_magik_result_types mgk_surrogate_n_params(uint64_t* n_weights, uint64_t* n_biases) {
    uint64_t w = 0, b = 0;
    for (int l = 0; l < kMlpLayers; ++l) {
        w += uint64_t(kK[l]) * kN[l];
        b += uint64_t(kN[l]);
    }
    if (n_weights) *n_weights = w;
    if (n_biases) *n_biases = b;
    return MAGIK_SUCCESS;
}

MGK_API e_magik_result_types mgk_surrogate_get_weights(mgk_surrogate_handle s, float* W, float* b) {
    if (!s || !W || !b) return MAGIK_ERROR_INVALID_POINTER;
    HIP_TRY(hipSetDevice(s->device));
    HIP_TRY(hipDeviceSynchronize());
    size_t ow = 0, ob = 0;
    for (int l = 0; l < kMlpLayers; ++l) {
        HIP_TRY(hipMemcpy(W + ow, s->W[l], size_t(kK[l]) * kN[l] * 4, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(b + ob, s->b[l], size_t(kN[l]) * 4, hipMemcpyDeviceToHost));
        ow += size_t(kK[l]) * kN[l];
        ob += kN[l];
    }
    return MAGIK_SUCCESS;
}

MGK_API e_magik_result_types mgk_surrogate_set_weights(mgk_surrogate_handle s, const float* W, const float* b) {
    if (!s || !W || !b) return MAGIK_ERROR_INVALID_POINTER;
    HIP_TRY(hipSetDevice(s->device));
    size_t ow = 0, ob = 0;
    for (int l = 0; l < kMlpLayers; ++l) {
        HIP_TRY(hipMemcpy(s->W[l], W + ow, size_t(kK[l]) * kN[l] * 4, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(s->b[l], b + ob, size_t(kN[l]) * 4, hipMemcpyHostToDevice));
        ow += size_t(kK[l]) * kN[l];
        ob += kN[l];
    }
    e_magik_result_types r = refresh_fragments(s, nullptr);
    if (r != MAGIK_SUCCESS) return r;
    HIP_TRY(hipDeviceSynchronize());
    return MAGIK_SUCCESS;
}

}  // extern "C"

namespace mgk {
e_magik_result_types surrogate_infer_args(mgk_surrogate* s, InferArgs& a, hipStream_t st) {
    HIP_TRY(hipSetDevice(s->device));
    if (s->frag_dirty) {
        e_magik_result_types r = refresh_fragments(s, st);
        if (r != MAGIK_SUCCESS) return r;
    }
    for (int l = 0; l < kMlpLayers; ++l) {
        a.Wf[l] = s->Wf[l];
        a.b[l] = s->b[l];
    }
    return MAGIK_SUCCESS;
}

e_magik_result_types surrogate_forward(mgk_surrogate* s, const float* d_x, int x_stride, uint64_t n, float* d_y, hipStream_t st) {
    InferArgs a;
    e_magik_result_types r = surrogate_infer_args(s, a, st);
    if (r != MAGIK_SUCCESS) return r;
    a.X = d_x;
    a.x_stride = x_stride;
    a.n = int64_t(n);
    a.Y = d_y;
    HIP_TRY(launch_infer(a, st));
    return MAGIK_SUCCESS;
}
}  // namespace mgk

extern "C" {

MGK_API e_magik_result_types mgk_surrogate_forward(mgk_surrogate_handle s, const float* d_x, uint64_t n, float* d_y, void* stream) {
    if (!s || (n && (!d_x || !d_y))) return MAGIK_ERROR_INVALID_POINTER;
    return surrogate_forward(s, d_x, kMlpInPad, n, d_y, hipStream_t(stream));
}


MGK_API e_magik_result_types mgk_surrogate_gradients(mgk_surrogate_handle s, const float* d_x, const float* d_dy, uint32_t n, float* d_y,
                                                     float* W_grad, float* b_grad) {
    if (!s || !d_x || !d_dy || !W_grad || !b_grad) return MAGIK_ERROR_INVALID_POINTER;
    if (n == 0 || n > (1u << 24)) return MGK_ERROR_INVALID_ARGUMENT;
    HIP_TRY(hipSetDevice(s->device));
    e_magik_result_types r = ensure_scratch(s, int(n));
    if (r != MAGIK_SUCCESS) return r;
    if (s->frag_dirty && (r = refresh_fragments(s, nullptr)) != MAGIK_SUCCESS) return r;
    if ((r = train_gradients(s, d_x, nullptr, d_dy, d_y, int(n), 0.0f, 0.0f, nullptr)) != MAGIK_SUCCESS) return r;
    AdamArgs a = adam_args(s);   // sums the weight gradients' partials (the weights stay)
    a.update = 0;
    HIP_TRY(launch_train_adam(a, 0, nullptr));
    HIP_TRY(hipDeviceSynchronize());
    size_t ow = 0, ob = 0;
    for (int l = 0; l < kMlpLayers; ++l) {
        HIP_TRY(hipMemcpy(W_grad + ow, s->gW[l], size_t(kK[l]) * kN[l] * 4, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(b_grad + ob, s->gb[l], size_t(kN[l]) * 4, hipMemcpyDeviceToHost));
        ow += size_t(kK[l]) * kN[l];
        ob += kN[l];
    }
    return MAGIK_SUCCESS;
}

MGK_API e_magik_result_types mgk_surrogate_train_step(mgk_surrogate_handle s, const float* d_x, const float* d_t, uint32_t n, float lr,
                                                      float w_r, float w_d, float* losses) {
    if (!s || !d_x || !d_t) return MAGIK_ERROR_INVALID_POINTER;
    if (n == 0 || n > (1u << 24)) return MGK_ERROR_INVALID_ARGUMENT;
    HIP_TRY(hipSetDevice(s->device));
    e_magik_result_types r = ensure_scratch(s, int(n));
    if (r != MAGIK_SUCCESS) return r;
    hipStream_t st = nullptr;
    if (s->frag_dirty && (r = refresh_fragments(s, st)) != MAGIK_SUCCESS) return r;
    if ((r = train_gradients(s, d_x, d_t, nullptr, nullptr, int(n), w_r, w_d, st)) != MAGIK_SUCCESS) return r;
    s->t += 1;
    AdamArgs a = adam_args(s);
    a.lr = lr;
    a.update = 1;
    HIP_TRY(launch_train_adam(a, s->t, st));   // also writes the fragments of the new weights
    if (losses) {
        float h[4];
        HIP_TRY(hipMemcpy(h, s->sums, sizeof(h), hipMemcpyDeviceToHost));
        for (int k = 0; k < 4; ++k) losses[k] = h[k] / float(n);
    }
    return MAGIK_SUCCESS;
}

MGK_API e_magik_result_types mgk_surrogate_train(mgk_surrogate_handle s, const float* d_x, const float* d_t, uint64_t n,
                                                 const mgk_train_config_t* cfg, float* epoch_losses) {
    if (!s || !d_x || !d_t || !cfg) return MAGIK_ERROR_INVALID_POINTER;
    if (cfg->batch == 0 || cfg->batch > n || cfg->epochs == 0) return MGK_ERROR_INVALID_ARGUMENT;
    const uint64_t steps_per_epoch = n / cfg->batch;
    const uint64_t total = steps_per_epoch * cfg->epochs;
    uint64_t step = 0;
    for (uint32_t e = 0; e < cfg->epochs; ++e) {
        double acc[4] = {0, 0, 0, 0};
        for (uint64_t k = 0; k < steps_per_epoch; ++k, ++step) {
            // cosine decay from lr to lr_final
            double frac = double(step) / double(total > 1 ? total - 1 : 1);
            float lr = float(cfg->lr_final + 0.5 * (cfg->lr - cfg->lr_final) * (1.0 + std::cos(3.14159265358979323846 * frac)));
            const uint64_t off = ((k + uint64_t(e) * 7919u) % steps_per_epoch) * cfg->batch;
            float l4[4];
            bool last = (k + 1 == steps_per_epoch);
            e_magik_result_types r = mgk_surrogate_train_step(s, d_x + off * kMlpInPad, d_t + off * kTargetStride, cfg->batch, lr,
                                                              cfg->w_r, cfg->w_d, (last || (k % 64) == 0) ? l4 : nullptr);
            if (r != MAGIK_SUCCESS) return r;
            if (last) for (int q = 0; q < 4; ++q) acc[q] = l4[q];
        }
        if (epoch_losses)
            for (int q = 0; q < 4; ++q) epoch_losses[4 * e + q] = float(acc[q]);
    }
    HIP_TRY(hipDeviceSynchronize());
    return MAGIK_SUCCESS;
}

MGK_API e_magik_result_types mgk_surrogate_generate_data(mgk_context_handle ctx, mgk_surrogate_handle s, const mgk_integrator_t* integ,
                                                         const mgk_scene_t* scene, uint64_t seed, uint64_t n, double* d_rays,
                                                         float* d_x, float* d_t, mgk_hit_t* d_hits, void* stream) {
    if (!ctx || !s || !integ || !scene || !d_rays || !d_x || !d_t || !d_hits) return MAGIK_ERROR_INVALID_POINTER;
    if (integ->metric_type != KERR || integ->precision != MGK_PRECISION_FP64) return MGK_ERROR_INVALID_ARGUMENT;
    // the disk of the targets runs from each ray's ISCO to disk_r_max
    if ((scene->flags & MGK_SCENE_DISK) && std::fabs(double(scene->disk_r_out) - s->ranges.disk_r_max) > 1e-4 * s->ranges.disk_r_max)
        return MGK_ERROR_INVALID_ARGUMENT;
    if (scene->flags & (MGK_SCENE_SPHERE | MGK_SCENE_BACKGROUND)) return MGK_ERROR_INVALID_ARGUMENT;
    HIP_TRY(hipSetDevice(s->device));
    hipStream_t st = hipStream_t(stream);
    TrainGen G;
    G.ranges = s->ranges;
    G.L = double(mgk::kG * integ->mass_kg / (mgk::kC * mgk::kC));   // the FP32 length scale of the render path
    G.time_reversed = integ->time_reversed ? 1 : 0;
    G.seed = seed;
    G.n = n;
    G.rays = d_rays;
    G.stride = n;
    G.X = d_x;
    if (n == 0) return MAGIK_SUCCESS;
    hipLaunchKernelGGL(k_training_rays, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, G);
    HIP_TRY(hipGetLastError());
    const float origin[3] = {0.0f, 0.0f, 0.0f};
    e_magik_result_types r = mgk_trace_rays_per_ray(ctx, integ, scene, origin, d_rays, n, n, d_hits, st);
    if (r != MAGIK_SUCCESS) return r;
    hipLaunchKernelGGL(k_targets, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, d_hits, n, s->ranges, d_t);
    HIP_TRY(hipGetLastError());
    return MAGIK_SUCCESS;
}

}  // extern "C"
