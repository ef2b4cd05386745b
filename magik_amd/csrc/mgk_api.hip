// C-ABI layer of the tracer: argument validation, per-device context (work
// counters, star index, cached frame buffers), host-side colour pipeline and
// camera placement, and the launches of the kernels in mgk_kernels.hip.
//
// Error convention: every entry point returns e_magik_result_types
// (magik.h:38-110); HIP failures map to MGK_ERROR_HIP and nothing calls exit()
// (the reference's check_cuda exits with 99, cuda_host_utilities.h:12-22).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "magik_amd.h"
#include "mgk_device.hpp"
#include "mgk_display.hpp"
#include "mgk_internal.hpp"
#include "mgk_surrogate.hpp"

namespace {

#include "cie1931_table.inc"   // kCie1931[471][3], 360..830 nm at 1 nm (generated from data/cie1931_2deg_1nm.csv)

using namespace mgk;

#define HIP_TRY(expr)                                  \
    do {                                               \
        hipError_t e_ = (expr);                        \
        if (e_ != hipSuccess) {                        \
            std::fprintf(stderr, "magik_amd: %s failed: %s\n", #expr, hipGetErrorString(e_)); \
            return MGK_ERROR_HIP;                      \
        }                                              \
    } while (0)

__global__ void k_zero_u64(unsigned long long* p, int n) {
    int i = threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// The trace kernels' work counters (kHeads of them, kHeadStride apart).
__global__ void k_zero_heads(unsigned long long* p) {
    const uint32_t i = threadIdx.x;
    if (i < mgk::kHeads) p[uint64_t(i) * mgk::kHeadStride] = 0ull;
}

// Counter block of a context, in u64: the work counters on their own 128-byte
// lines, then the statistics, the two hand-off queue counts and the reflected
// ray count, each line apart.
constexpr uint64_t kCtrStats = uint64_t(mgk::kHeads) * mgk::kHeadStride;   // [0] steps [1] trials [2] fp64 trials
constexpr uint64_t kCtrZone0 = kCtrStats + 16, kCtrZone1 = kCtrStats + 32, kCtrReflect = kCtrStats + 48;
constexpr uint64_t kCtrTotal = kCtrStats + 64;

struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    hipError_t reserve(size_t need) {
        if (need <= bytes) return hipSuccess;
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
        hipError_t e = hipMalloc(&ptr, need);
        if (e == hipSuccess) bytes = need;
        return e;
    }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
};

// Device memory of mgk_trace_rays_noalloc (its work counters, statistics and,
// for the mixed precision, the zone hand-off queues and parked states), so that
// it allocates nothing.  Mixed-precision calls take at most kNoAllocZoneRays
// rays: a pass reserves hand-off slots for at most its rays plus kOutBlock per
// wave of its launch, and a launch of n <= kNoAllocZoneRays rays has at most
// ceil(n / 256) blocks of 4 waves.  About 20 MB per device, held until the
// process ends.
constexpr uint64_t kNoAllocZoneRays = 16384;
constexpr uint64_t kNoAllocZoneCap = kNoAllocZoneRays + (kNoAllocZoneRays + 255) / 256 * 4 * kOutBlock;
constexpr uint64_t kScrCounters = 0;                                                     // kCtrTotal u64
constexpr uint64_t kScrState = (kCtrTotal * 8 + 255) / 256 * 256;                        // 2 x kParked x cap doubles
constexpr uint64_t kScrSteps = kScrState + 2 * kNoAllocZoneCap * kParked * 8;            // 2 x cap u32
constexpr uint64_t kScrQueues = kScrSteps + 2 * kNoAllocZoneCap * 4;                     // 2 x cap u32
constexpr uint64_t kScratchBytes = kScrQueues + 2 * kNoAllocZoneCap * 4;

// The buffers of one trace that the caller does not pass: either the
// context's (grown on demand) or a fixed workspace (zone_cap > 0).
struct Workspace {
    unsigned long long* counters = nullptr;   // kCtr* layout
    double* zone_state[2] = {nullptr, nullptr};
    uint32_t* zone_steps[2] = {nullptr, nullptr};
    uint32_t* zone_q[2] = {nullptr, nullptr};
    uint64_t zone_cap = 0;
};

}  // namespace

// In host code `scratch` names a host-side stand-in; its device address comes
// from hipGetSymbolAddress (per device, with that device current) and only that
// address is handed to kernels and copies.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxScratchDevices] = {};

// Device address of this device's scratch array, resolved once per device; the
// device must be current.
hipError_t scratch_address(int device, void** out) {
    if (device < 0 || device >= kMaxScratchDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = ptr;
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

Workspace scratch_workspace(void* base) {
    unsigned char* b = static_cast<unsigned char*>(base);
    Workspace w;
    w.counters = reinterpret_cast<unsigned long long*>(b + kScrCounters);
    for (int k = 0; k < 2; ++k) {
        w.zone_state[k] = reinterpret_cast<double*>(b + kScrState + uint64_t(k) * kNoAllocZoneCap * kParked * 8);
        w.zone_steps[k] = reinterpret_cast<uint32_t*>(b + kScrSteps + uint64_t(k) * kNoAllocZoneCap * 4);
        w.zone_q[k] = reinterpret_cast<uint32_t*>(b + kScrQueues + uint64_t(k) * kNoAllocZoneCap * 4);
    }
    w.zone_cap = kNoAllocZoneCap;
    return w;
}
}  // namespace

struct mgk_context {
    int device = 0;
    int n_cu = 0;
    void* scratch = nullptr;                     // this device's scratch array (device address)
    unsigned long long* d_counters = nullptr;   // work counters, statistics, queue counts (kCtr*)
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool timed = false;
    uint64_t last_n = 0;
    // star index
    DeviceBuffer stars, cell_start, cell_items;
    uint32_t grid_nc = 0, grid_np = 0;
    uint64_t star_key = ~0ull;
    uint32_t star_n = 0;
    float star_rho = -1.0f, star_scale = 0.0f;
    // star grid widened by a margin (surrogate star-edge test)
    DeviceBuffer band_start, band_items;
    uint64_t band_key = ~0ull;
    uint32_t band_n = 0, band_nc = 0, band_np = 0;
    float band_rho = -1.0f;
    // frame buffers
    DeviceBuffer rays, hits;
    DeviceBuffer batch_params;   // per-camera ray-generation parameters, ray offsets and IC constants
    DeviceBuffer ic_camera;      // IC constants of the last single-camera generation
    DeviceBuffer shade_resp;     // tabulated spectral response of the last shading call
    // zone-split integration: parked states and the two hand-off queues
    DeviceBuffer zone_state[2], zone_steps[2], zone_q0, zone_q1;
    int last_passes = 0;
    // paths continued after reflections: terminal states, step counts, reflected-ray queues
    DeviceBuffer path_state, path_steps, refl_q[2];
    uint64_t last_reflections = 0;
    DeviceBuffer step_hist;      // end / limit / reflect step histograms of a render that asks for them
    // display parameters derived from the last display settings (the blackbody
    // white point alone is ~1e3 Planck evaluations): reused while the settings repeat
    mgk_display_t disp_key{};
    disp::Params disp_params{};
    bool disp_valid = false;
    // triangle meshes (mgk_set_meshes): the two-level hierarchy
    DeviceBuffer mesh_nodes, mesh_tris, mesh_normals, mesh_inst;
    uint32_t mesh_root = 0;
    float mesh_lo[3] = {0, 0, 0}, mesh_hi[3] = {0, 0, 0};
    bool mesh_valid = false, mesh_diffuse = false;
};

namespace {

e_magik_result_types set_device(mgk_context* c) {
    HIP_TRY(hipSetDevice(c->device));
    return MAGIK_SUCCESS;
}

// Geometric length scale G M / c^2 in float (metric_kerr.cuh:154-157).
float length_scale_of(const mgk_integrator_t* in) {
    if (in->metric_type == MINKOWSKI) return 1.0f;
    float m = in->metric_type == SCHWARZSCHILD ? std::fabs(in->mass_kg) : in->mass_kg;
    return kG * m / (kC * kC);
}

bool valid_metric(int m) { return m == MINKOWSKI || m == SCHWARZSCHILD || m == KERR; }

e_magik_result_types check_integrator(const mgk_integrator_t* in) {
    if (!in) return MAGIK_ERROR_INVALID_POINTER;
    if (!valid_metric(in->metric_type)) return MAGIK_ERROR_UNKNOWN_METRIC;
    if (in->precision < MGK_PRECISION_FP32_REFERENCE || in->precision > MGK_PRECISION_FP64) return MGK_ERROR_INVALID_ARGUMENT;
    if (in->metric_type != MINKOWSKI && (in->n_max_substep == 0 || !(in->abs_tol > 0.0f) || !(in->rel_tol >= 0.0f) ||
                                         !(in->h_max > 0.0f)))
        return MGK_ERROR_INVALID_ARGUMENT;
    if (in->metric_type != MINKOWSKI && !(in->mass_kg != 0.0f)) return MGK_ERROR_INVALID_ARGUMENT;
    return MAGIK_SUCCESS;
}

float signed_spin(const mgk_integrator_t* in) {
    if (in->metric_type != KERR) return 0.0f;
    return in->time_reversed ? -in->spin : in->spin;
}
float mass_sign_of(const mgk_integrator_t* in) {
    return (in->metric_type == SCHWARZSCHILD && in->mass_kg < 0.0f) ? -1.0f : 1.0f;
}

// Outer prograde/retrograde photon-orbit radius (Bardeen, Press & Teukolsky
// 1972): r_ph = 2 (1 + cos(2/3 acos(+|a|))) for the retrograde (outer) orbit.
double outer_photon_radius(double a) {
    a = std::fabs(a);
    return 2.0 * (1.0 + std::cos((2.0 / 3.0) * std::acos(a)));
}

void fill_trace_params(const mgk_integrator_t* in, const mgk_scene_t* sc, const float cam[3], TraceParams& P) {
    std::memset(&P, 0, sizeof(P));
    P.metric = in->metric_type;
    P.precision = in->precision;
    P.spin = signed_spin(in);
    P.mass_sign = mass_sign_of(in);
    P.length_scale = length_scale_of(in);
    P.h_max = in->h_max;
    P.abs_tol = in->abs_tol;
    P.rel_tol = in->rel_tol;
    P.n_max = in->n_max_substep ? in->n_max_substep : 1;
    P.step_limit = in->step_limit;
    P.horizon_scalar = in->event_horizon_scalar;
    if (in->metric_type == KERR) {
        float a = P.spin;
        P.spin_in_range = std::fabs(a) > 1.0f ? 0 : 1;
        P.horizon_r_f = (1.0f + std::sqrt(1.0f - a * a)) * in->event_horizon_scalar;
        double ad = double(a);
        P.horizon_r_d = (1.0 + std::sqrt(1.0 - ad * ad)) * double(in->event_horizon_scalar);
        double rp = 1.0 + std::sqrt(std::max(0.0, 1.0 - ad * ad));
        double zr = std::max(1.5 * outer_photon_radius(ad), rp + 0.1);
        P.zone_rho2 = zr * zr + ad * ad;
    } else if (in->metric_type == SCHWARZSCHILD) {
        P.spin_in_range = 1;
        P.horizon_r_f = 2.0f * in->event_horizon_scalar;
        P.horizon_r_d = 2.0 * double(in->event_horizon_scalar);
        double zr = P.mass_sign > 0 ? std::max(1.5 * 3.0, 2.0 + 0.1) : 0.0;
        P.zone_rho2 = zr * zr;
    }
    for (int k = 0; k < 3; ++k) P.cam[k] = cam[k];
    P.cel_r2_f = sc->celestial_radius_m * sc->celestial_radius_m;
    P.cel_r2_d = double(sc->celestial_radius_m) * double(sc->celestial_radius_m);
    P.flags = sc->flags;
    P.disk_r_in = sc->disk_r_in;
    P.disk_r_out = sc->disk_r_out;
    for (int k = 0; k < 3; ++k) P.sphere_c[k] = sc->sphere_center[k];
    P.sphere_r = sc->sphere_radius;
    if (sc->flags & MGK_SCENE_BOXES) {
        P.n_box = sc->n_box < uint32_t(MGK_MAX_BOX) ? sc->n_box : uint32_t(MGK_MAX_BOX);
        for (uint32_t b = 0; b < P.n_box; ++b) {
            for (int k = 0; k < 3; ++k) {
                P.box_min[b][k] = sc->box[b].min[k];
                P.box_max[b][k] = sc->box[b].max[k];
            }
            P.box_matter[b] = sc->box[b].matter;
        }
    }
    P.chord = sc->chord;
}

e_magik_result_types check_scene(const mgk_scene_t* sc) {
    if (!sc) return MAGIK_ERROR_INVALID_POINTER;
    if (sc->flags & MGK_SCENE_BOXES) {
        if (sc->n_box > uint32_t(MGK_MAX_BOX)) return MGK_ERROR_INVALID_ARGUMENT;
        for (uint32_t b = 0; b < sc->n_box; ++b)
            if (sc->box[b].matter != MGK_MATTER_DIFFUSE && sc->box[b].matter != MGK_MATTER_ILLUMINANT) return MGK_ERROR_INVALID_ARGUMENT;
    }
    if (sc->chord != MGK_CHORD_AFFINE && sc->chord != MGK_CHORD_HERMITE) return MGK_ERROR_INVALID_ARGUMENT;
    return MAGIK_SUCCESS;
}

bool has_diffuse_matter(const mgk_context* c, const TraceParams& P) {
    if ((P.flags & MGK_SCENE_MESHES) && c->mesh_diffuse) return true;
    if (!(P.flags & MGK_SCENE_BOXES)) return false;
    for (uint32_t b = 0; b < P.n_box; ++b)
        if (P.box_matter[b] == MGK_MATTER_DIFFUSE) return true;
    return false;
}

int trace_grid(mgk_context* c, const TraceParams& P) {
    // persistent launch: as many 256-thread blocks as can be resident, capped
    // by the work available
    int per_cu = 8;
    if (P.precision != MGK_PRECISION_FP32_REFERENCE) per_cu = 4;
    uint64_t want = uint64_t(c->n_cu) * uint64_t(per_cu);
    uint64_t need = (P.n + 255) / 256;
    if (need < want) want = need;
    if (want < 1) want = 1;
    return int(want);
}

// Mixed precision: alternate the far-zone (FP32) and near-zone (FP64) passes
// until no ray is handed over.  Reads one queue count per pass (a 8-byte copy).
e_magik_result_types do_trace_zones(mgk_context* c, TraceParams& P, hipStream_t s, const Workspace& ws) {
    if (P.n >= uint64_t(kHole)) return MGK_ERROR_CAPACITY_EXCEEDED;   // ray ids travel as uint32 below kHole
    // A pass hands off each of its rays at most once; slots are reserved kOutBlock
    // per wave, so the unused tails of the waves' last blocks add at most
    // (waves of its launch) * kOutBlock.  A later pass reads slots, holes
    // included, so its slot count may exceed the rays and its launch the first
    // pass's: with the context's buffers every launch may take the whole GPU
    // (each wave then starts on one 64-slot chunk, as in the first pass); in a
    // fixed workspace the later launches are capped at the first pass's, which
    // bounds every pass by n rays plus (waves of the first launch) * kOutBlock.
    TraceParams Q0 = P;
    const bool fixed = ws.zone_cap > 0;
    const int grid0 = fixed ? trace_grid(c, Q0) : c->n_cu * 4;
    const uint64_t max_waves = uint64_t(grid0) * 4u;
    uint64_t cap = P.n + max_waves * kOutBlock;
    if (cap >= (1ull << 32)) return MGK_ERROR_CAPACITY_EXCEEDED;   // the kernels hold slot indices in 32 bits
    double* state[2];
    uint32_t* stp[2];
    uint32_t* q[2];
    if (ws.zone_cap > 0) {
        if (cap > ws.zone_cap) return MGK_ERROR_CAPACITY_EXCEEDED;
        cap = ws.zone_cap;
        for (int k = 0; k < 2; ++k) { state[k] = ws.zone_state[k]; stp[k] = ws.zone_steps[k]; q[k] = ws.zone_q[k]; }
    } else {
        for (int k = 0; k < 2; ++k) {
            HIP_TRY(c->zone_state[k].reserve(cap * kParked * sizeof(double)));
            HIP_TRY(c->zone_steps[k].reserve(cap * sizeof(uint32_t)));
            state[k] = reinterpret_cast<double*>(c->zone_state[k].ptr);
            stp[k] = reinterpret_cast<uint32_t*>(c->zone_steps[k].ptr);
        }
        HIP_TRY(c->zone_q0.reserve(cap * sizeof(uint32_t)));
        HIP_TRY(c->zone_q1.reserve(cap * sizeof(uint32_t)));
        q[0] = reinterpret_cast<uint32_t*>(c->zone_q0.ptr);
        q[1] = reinterpret_cast<uint32_t*>(c->zone_q1.ptr);
    }
    unsigned long long* cnt[2] = {ws.counters + kCtrZone0, ws.counters + kCtrZone1};
    const uint32_t* caller_queue = P.queue;
    uint64_t n_in = P.n;
    bool near = false;
    int pass = 0;
    const int max_passes = 64;
    for (;;) {
        ZoneParams Z;
        std::memset(&Z, 0, sizeof(Z));
        Z.from_ic = pass == 0;
        Z.no_handoff = pass + 1 >= max_passes;
        Z.in_queue = pass == 0 ? caller_queue : q[(pass + 1) & 1];
        Z.n_in = n_in;
        Z.out_queue = q[pass & 1];
        Z.out_count = cnt[pass & 1];
        // pass p parks into buffer p % 2 and reads what pass p - 1 parked
        Z.cap = cap;
        Z.state_out = state[pass & 1];
        Z.steps_out = stp[pass & 1];
        Z.state_in = state[(pass + 1) & 1];
        Z.steps_in = stp[(pass + 1) & 1];
        hipLaunchKernelGGL(k_zero_heads, dim3(1), dim3(64), 0, s, ws.counters);
        hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, cnt[pass & 1], 1);
        HIP_TRY(hipGetLastError());
        TraceParams Q = P;
        Q.n = n_in;
        const int grid = std::min(trace_grid(c, Q), grid0);   // the slot bound above
        HIP_TRY(launch_trace_zone(Q, Z, near || Z.no_handoff, grid, s));
        ++pass;
        unsigned long long next = 0;
        HIP_TRY(hipMemcpyAsync(&next, cnt[(pass - 1) & 1], sizeof(next), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        if (next == 0 || Z.no_handoff) break;
        n_in = next;
        near = !near;
    }
    c->last_passes = pass;
    return MAGIK_SUCCESS;
}

// One trace of P.n rays (of P.queue when given): the zone passes, or one
// persistent launch.
e_magik_result_types trace_once(mgk_context* c, TraceParams& P, hipStream_t s, const Workspace& ws) {
    if (P.n == 0) return MAGIK_SUCCESS;
    if (P.precision == MGK_PRECISION_MIXED && P.metric != MINKOWSKI) return do_trace_zones(c, P, s, ws);
    hipLaunchKernelGGL(k_zero_heads, dim3(1), dim3(64), 0, s, ws.counters);
    HIP_TRY(hipGetLastError());
    HIP_TRY(launch_trace(P, trace_grid(c, P), s));
    return MAGIK_SUCCESS;
}

// The Philox stream of a camera's paths, for the reflections' uniforms.
struct PathInfo {
    uint64_t seed;
    uint32_t camera_id, npix, sample_begin;
};

// Bins of a render's step histograms: one per step up to the step limit
// (4 MiB per histogram at most; later steps share the last bin).
constexpr uint64_t kMaxHistBins = uint64_t(1) << 20;

// Trace, then (with `paths` and a scene with diffuse boxes) reflect the paths
// that ended on diffuse matter and trace them again, up to max_reflections
// times (kernel_bxdf.cu:69-95, kernel_enforce_ray_limits.cu:48-50).  Each
// round reads one count back (an 8-byte copy).  Reflected rays get new initial
// conditions in place in P.rays.  With a fixed workspace (fixed->zone_cap > 0)
// nothing is allocated: no reflections, no histograms, and the context's
// statistics and timing events are left as they were.  With hist, the step
// histograms of the paths (after the timed region; one synchronous copy).
e_magik_result_types do_trace(mgk_context* c, TraceParams& P, hipStream_t s, const PathInfo* paths = nullptr,
                              uint32_t max_reflections = 0, const Workspace* fixed = nullptr,
                              StepHistogram* hist = nullptr) {
    Workspace ws;
    if (fixed) {
        ws = *fixed;
    } else {
        ws.counters = c->d_counters;
    }
    if (P.flags & MGK_SCENE_MESHES) {
        if (!c->mesh_valid) return MGK_ERROR_INVALID_ARGUMENT;   // no meshes set on this context
        P.mesh_nodes = reinterpret_cast<const float4*>(c->mesh_nodes.ptr);
        P.mesh_tris = reinterpret_cast<const float4*>(c->mesh_tris.ptr);
        P.mesh_inst = reinterpret_cast<const float4*>(c->mesh_inst.ptr);
        P.mesh_root = c->mesh_root;
        for (int k = 0; k < 3; ++k) { P.mesh_lo[k] = c->mesh_lo[k]; P.mesh_hi[k] = c->mesh_hi[k]; }
    }
    P.head = ws.counters;
    P.stats = ws.counters + kCtrStats;
    hipLaunchKernelGGL(k_zero_heads, dim3(1), dim3(64), 0, s, ws.counters);
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, P.stats, 3);
    HIP_TRY(hipGetLastError());
    if (fixed) return trace_once(c, P, s, ws);
    const bool want_hist = hist && P.n > 0;
    const uint32_t n_bin = uint32_t(std::min<uint64_t>(uint64_t(P.step_limit) + 2u, kMaxHistBins));
    uint32_t* d_hist = nullptr;   // end, limit, reflect: n_bin bins each
    if (want_hist) {
        if (!P.steps_out) {
            HIP_TRY(c->path_steps.reserve(P.stride * sizeof(uint32_t)));
            P.steps_out = reinterpret_cast<uint32_t*>(c->path_steps.ptr);
        }
        HIP_TRY(c->step_hist.reserve(size_t(3) * n_bin * sizeof(uint32_t)));
        d_hist = reinterpret_cast<uint32_t*>(c->step_hist.ptr);
        HIP_TRY(hipMemsetAsync(d_hist, 0, size_t(3) * n_bin * sizeof(uint32_t), s));
    }
    HIP_TRY(hipEventRecord(c->ev0, s));
    c->last_reflections = 0;
    const bool reflect = paths && max_reflections > 0 && P.n > 0 && has_diffuse_matter(c, P);
    if (reflect) {
        if (P.n >= uint64_t(kHole)) return MGK_ERROR_CAPACITY_EXCEEDED;
        HIP_TRY(c->refl_q[0].reserve(P.n * sizeof(uint32_t)));
        HIP_TRY(c->refl_q[1].reserve(P.n * sizeof(uint32_t)));
        if (!P.final_state) {
            HIP_TRY(c->path_state.reserve(P.stride * 8 * sizeof(double)));
            P.final_state = reinterpret_cast<double*>(c->path_state.ptr);
        }
        if (!P.steps_out) {
            HIP_TRY(c->path_steps.reserve(P.stride * sizeof(uint32_t)));
            P.steps_out = reinterpret_cast<uint32_t*>(c->path_steps.ptr);
        }
    }
    e_magik_result_types r = trace_once(c, P, s, ws);
    if (r != MAGIK_SUCCESS) return r;
    if (reflect) {
        const uint32_t* in = P.queue;
        uint64_t n_in = P.n;
        for (uint32_t k = 0; k < max_reflections; ++k) {
            ReflectParams R;
            std::memset(&R, 0, sizeof(R));
            R.metric = P.metric;
            R.hits = P.hits;
            R.state = P.final_state;
            R.rays = const_cast<void*>(P.rays);
            R.stride = P.stride;
            R.queue_in = in;
            R.n_in = n_in;
            R.queue_out = reinterpret_cast<uint32_t*>(c->refl_q[k & 1].ptr);
            R.count = c->d_counters + kCtrReflect;
            R.seed = paths->seed;
            R.camera_id = paths->camera_id;
            R.npix = paths->npix;
            R.sample_begin = paths->sample_begin;
            R.block = 1 + k;
            R.spin = P.spin;
            R.mass_sign = P.mass_sign;
            R.mesh_inst = P.mesh_inst;
            R.mesh_normals = reinterpret_cast<const float4*>(c->mesh_normals.ptr);
            hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, R.count, 1);
            HIP_TRY(hipGetLastError());
            HIP_TRY(launch_reflect(R, P.precision, s));
            unsigned long long cnt = 0;
            HIP_TRY(hipMemcpyAsync(&cnt, R.count, sizeof(cnt), hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            if (cnt == 0) break;
            c->last_reflections += cnt;
            if (want_hist)   // steps_out still holds each reflected path's step at the reflection
                HIP_TRY(launch_step_hist(P.steps_out, nullptr, R.queue_out, cnt, n_bin, d_hist + 2 * n_bin, d_hist + 2 * n_bin,
                                         c->n_cu, s));
            TraceParams Q = P;
            Q.queue = R.queue_out;
            Q.n = cnt;
            Q.steps_in = P.steps_out;
            if ((r = trace_once(c, Q, s, ws)) != MAGIK_SUCCESS) return r;
            in = R.queue_out;
            n_in = cnt;
        }
    }
    HIP_TRY(hipEventRecord(c->ev1, s));
    c->timed = true;
    c->last_n = P.n;
    if (want_hist) {
        HIP_TRY(launch_step_hist(P.steps_out, P.hits, P.queue, P.n, n_bin, d_hist, d_hist + n_bin, c->n_cu, s));
        std::vector<uint32_t> h(size_t(3) * n_bin);
        HIP_TRY(hipMemcpyAsync(h.data(), d_hist, h.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        hist->end.assign(h.begin(), h.begin() + n_bin);
        hist->limit.assign(h.begin() + n_bin, h.begin() + 2 * size_t(n_bin));
        hist->reflect.assign(h.begin() + 2 * size_t(n_bin), h.end());
    } else if (hist) {
        hist->end.clear();
        hist->limit.clear();
        hist->reflect.clear();
    }
    return MAGIK_SUCCESS;
}

// ---------------------------------------------------------------------------
// Host colour pipeline (host_color_utilities.cuh:14-178, color_pipeline.cuh:21-48).
// ---------------------------------------------------------------------------
struct f3 { float x, y, z; };
struct m3 { float m[9]; };
f3 mul(const m3& a, f3 v) {
    return {a.m[0] * v.x + a.m[1] * v.y + a.m[2] * v.z, a.m[3] * v.x + a.m[4] * v.y + a.m[5] * v.z,
            a.m[6] * v.x + a.m[7] * v.y + a.m[8] * v.z};
}
m3 mul(const m3& a, const m3& b) {
    m3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            r.m[3 * i + j] = a.m[3 * i] * b.m[j] + a.m[3 * i + 1] * b.m[3 + j] + a.m[3 * i + 2] * b.m[6 + j];
    return r;
}

f3 cmf_at(float w) {   // linear interpolation in the 1 nm table (color_matching_functions.cuh:35-44)
    const uint32_t n = 471;
    float spacing = (830.0f - 360.0f) / float(n - 1);
    float fi = std::fmin(std::fmax((w - 360.0f) / spacing, 0.0f), float(n - 1));
    uint32_t lo = uint32_t(std::fmin(float(uint32_t(fi)), float(n - 2)));
    float t = fi - float(lo);
    return {(kCie1931[lo][0] * (1.0f - t)) + (kCie1931[lo + 1][0] * t), (kCie1931[lo][1] * (1.0f - t)) + (kCie1931[lo + 1][1] * t),
            (kCie1931[lo][2] * (1.0f - t)) + (kCie1931[lo + 1][2] * t)};
}

float planck_host(float lambda_nm, float T) {
    if (T <= 0.0f) return 0.0f;
    float lm = lambda_nm * 1e-9f;
    const float t0 = kPlanck * kC, t1 = 2.0f * t0 * kC;
    float t2 = lm * lm * lm * lm * lm;
    float t3 = std::exp(t0 / (lm * kBoltzmann * T)) - 1.0f;
    return (t1 / t2) * (1.0f / t3);
}

void bin_weights(const mgk_spectral_t* sp, float* W) {
    float sum_y = 0.0f;
    float bin_range = (sp->bin_max_nm - sp->bin_min_nm) / float(sp->n_bin);
    float tab_range = (830.0f - 360.0f) / float(471 - 1);
    uint32_t sub = uint32_t(std::ceil(bin_range / tab_range));
    float sub_w = bin_range / float(sub);
    for (uint32_t i = 0; i < sp->n_bin; ++i) {
        f3 w{0, 0, 0};
        float x0 = sp->bin_min_nm + float(i) * bin_range;
        for (uint32_t j = 0; j < sub; ++j) {
            f3 c = cmf_at(x0 + (float(j) + 0.5f) * sub_w);
            w = {w.x + c.x * sub_w, w.y + c.y * sub_w, w.z + c.z * sub_w};
        }
        W[3 * i] = w.x; W[3 * i + 1] = w.y; W[3 * i + 2] = w.z;
        sum_y += w.y;
    }
    float k = 1.0f / sum_y, inv_bw = 1.0f / bin_range;
    for (uint32_t i = 0; i < 3 * sp->n_bin; ++i) W[i] = W[i] * (k * inv_bw);
}

// The inverse CDF of x+y+z plus a floor of 5% of its mean over [lo, hi], as
// kCmfCdfN wavelengths at u = k / (kCmfCdfN - 1): the CMF resampled at 471
// points across [lo, hi], a trapezoid CDF accumulated in float, normalised,
// and inverted by one linear walk (host_color_utilities.cuh:188-233; the
// floor of color_pipeline.cuh:36).
void cmf_inverse_cdf(float lo, float hi, float* inv) {
    const uint32_t n_e = 471, n = n_e - 1;
    std::vector<float> s(n_e), cdf(n + 1, 0.0f);
    const float step = (hi - lo) / float(n);
    for (uint32_t i = 0; i < n_e; ++i) {
        const f3 c = cmf_at(lo + float(i) * step);
        s[i] = c.x + c.y + c.z;
    }
    float mean = 0.0f;
    for (uint32_t i = 0; i < n_e; ++i) mean += s[i];
    mean /= float(n_e);
    const float floor_ = 0.05f * mean;
    for (uint32_t i = 0; i < n; ++i) cdf[i + 1] = cdf[i] + (0.5f * (s[i] + s[i + 1]) + floor_) * step;
    const float total = cdf[n];
    for (uint32_t j = 0; j <= n; ++j) cdf[j] /= total;
    uint32_t b = 0;
    for (int k = 0; k < kCmfCdfN; ++k) {
        const float u = float(k) / float(kCmfCdfN - 1);
        while (b < n - 1 && cdf[b + 1] < u) ++b;
        const float t = (u - cdf[b]) / std::fmax(cdf[b + 1] - cdf[b], 1e-12f);
        inv[k] = lo + (float(b) + t) * step;
    }
}

f3 xy_to_XYZ(float x, float y) { return {x / y, 1.0f, (1.0f - x - y) / y}; }

// Blackbody white point shifted by tint along the locus normal (host_color_utilities.cuh:29-57).
// In double: the locus direction is the difference of the uv points at T and
// T + 1 K, about 1e-5 of uv, which FP32 resolves to only ~1e-2.
f3 white_point_xyz(float T, float tint) {
    double a[3] = {0, 0, 0}, b[3] = {0, 0, 0};
    const double h = 6.62607015e-34, c = 299792458.0, k = 1.380649e-23;
    for (uint32_t i = 0; i < 471; ++i) {
        const double lam = (360.0 + double(i)) * 1e-9;
        const double p0 = 1.0 / (std::pow(lam, 5) * std::expm1(h * c / (lam * k * double(T))));
        const double p1 = 1.0 / (std::pow(lam, 5) * std::expm1(h * c / (lam * k * (double(T) + 1.0))));
        for (int j = 0; j < 3; ++j) {
            a[j] += double(kCie1931[i][j]) * p0;
            b[j] += double(kCie1931[i][j]) * p1;
        }
    }
    auto uv = [](const double* v) {
        const double d = v[0] + 15.0 * v[1] + 3.0 * v[2];
        return std::pair<double, double>{4.0 * v[0] / d, 6.0 * v[1] / d};
    };
    const auto u0 = uv(a), u1 = uv(b);
    double nx = u1.first - u0.first, ny = u1.second - u0.second;
    const double nl = std::sqrt(nx * nx + ny * ny);
    nx /= nl; ny /= nl;
    double tx = -ny, ty = nx;
    if (ty < 0.0) { tx = -tx; ty = -ty; }
    const double up = u0.first + tx * double(tint), vp = u0.second + ty * double(tint);
    const double den = 2.0 * up - 8.0 * vp + 4.0;
    return xy_to_XYZ(float(3.0 * up / den), float(2.0 * vp / den));
}

// Bradford chromatic adaptation (host_color_utilities.cuh:59-68).
m3 bradford(f3 src, f3 dst) {
    const m3 M{{0.8951f, 0.2664f, -0.1614f, -0.7502f, 1.7135f, 0.0367f, 0.0389f, -0.0685f, 1.0296f}};
    const m3 Mi{{0.9869929f, -0.1470543f, 0.1599627f, 0.4323053f, 0.5183603f, 0.0492912f, -0.0085287f, 0.0400428f, 0.9684867f}};
    f3 ls = mul(M, src), ld = mul(M, dst);
    m3 D{{ld.x / ls.x, 0, 0, 0, ld.y / ls.y, 0, 0, 0, ld.z / ls.z}};
    return mul(mul(Mi, D), M);
}

// XYZ -> white-balanced linear Rec.709 (IEC 61966-2-1 / ITU-R BT.709 primaries, D65).
m3 xyz_to_scene(float white_K) {
    const m3 to709{{3.24100326f, -1.53739899f, -0.49861587f, -0.96922426f, 1.87592999f, 0.04155422f, 0.05563942f, -0.2040112f,
                    1.05714897f}};
    f3 src = white_point_xyz(white_K, 0.0f);
    f3 d65 = xy_to_XYZ(0.31272f, 0.32903f);
    return mul(to709, bradford(src, d65));
}

// Display pipeline parameters (color_pipeline.cuh:21-48 and the view
// transforms' make_params, view_transform.cuh): white balance into the scene
// space, the scene -> display matrix, the working-space conversions of the view
// transforms (convert_scene_colorspace, host_color_utilities.cuh:70-73).
disp::M3 to_disp(const m3& a) {
    disp::M3 r;
    std::memcpy(r.m, a.m, sizeof(r.m));
    return r;
}
m3 from_disp(const disp::M3& a) {
    m3 r;
    std::memcpy(r.m, a.m, sizeof(r.m));
    return r;
}
m3 space_convert(const disp::Space& from, const disp::Space& to) {
    return mul(mul(from_disp(to.from_xyz), bradford(xy_to_XYZ(from.wx, from.wy), xy_to_XYZ(to.wx, to.wy))), from_disp(from.to_xyz));
}

e_magik_result_types display_params(const mgk_display_t* d, disp::Params& P) {
    std::memset(&P, 0, sizeof(P));
    disp::Space sc, dc, work;
    if (!disp::scene_space(d->scene_colorspace, sc) || !disp::display_space(d->display_colorspace, dc)) return MAGIK_UNKNOWN_ENUM_TYPE;
    if (d->view_transform > MAGIK_VIEW_TRANSFORM_KHRONOS_PBR_NEUTRAL || d->gt7_mode > MAGIK_GRAN_TURISMO_7_MODE_JZAZBZ ||
        d->agx_look > MAGIK_AGX_LOOK_PUNCHY || d->output_format > MAGIK_OUTPUT_FORMAT_FLOAT32)
        return MAGIK_UNKNOWN_ENUM_TYPE;
    if (!(d->white_point_K > 0.0f) || !std::isfinite(d->tint)) return MGK_ERROR_INVALID_ARGUMENT;
    P.view = d->view_transform;
    P.display = d->display_colorspace;
    P.xyz_to_scene = to_disp(mul(from_disp(sc.from_xyz), bradford(white_point_xyz(d->white_point_K, d->tint), xy_to_XYZ(sc.wx, sc.wy))));
    P.scene_to_display = to_disp(space_convert(sc, dc));
    P.lum = {sc.to_xyz.m[3], sc.to_xyz.m[4], sc.to_xyz.m[5]};
    switch (d->view_transform) {
        case MAGIK_VIEW_TRANSFORM_REINHARD_EXTENDED:
            if (!(d->reinhard_white_luminance > 0.0f)) return MGK_ERROR_INVALID_ARGUMENT;
            P.reinhard_white = d->reinhard_white_luminance;
            break;
        case MAGIK_VIEW_TRANSFORM_KHRONOS_PBR_NEUTRAL:
            disp::scene_space(MAGIK_SCENE_COLORSPACE_LINEAR_REC709, work);
            P.to_work = to_disp(space_convert(sc, work));
            P.from_work = to_disp(space_convert(work, sc));
            P.khr_start = d->khronos_start_compression;
            P.khr_desat = d->khronos_desaturation;
            break;
        case MAGIK_VIEW_TRANSFORM_GRAN_TURISMO_7:
            if (!(d->gt7_sdr_paper_white > 0.0f) || !(d->gt7_reference_luminance > 0.0f) || !(d->gt7_jzazbz_exponent_scale > 0.0f))
                return MGK_ERROR_INVALID_ARGUMENT;
            disp::scene_space(MAGIK_SCENE_COLORSPACE_LINEAR_REC2020, work);
            P.to_work = to_disp(space_convert(sc, work));
            P.from_work = to_disp(space_convert(work, sc));
            P.gt7 = disp::gt7_setup(d->gt7_mode, d->gt7_sdr_paper_white, d->gt7_reference_luminance, d->gt7_jzazbz_exponent_scale);
            break;
        case MAGIK_VIEW_TRANSFORM_AGX:
            disp::scene_space(MAGIK_SCENE_COLORSPACE_LINEAR_REC2020, work);
            P.to_work = to_disp(space_convert(sc, work));
            P.from_work = to_disp(space_convert(work, sc));
            P.agx_look = d->agx_look;
            break;
        default:
            break;
    }
    return MAGIK_SUCCESS;
}

// ---------------------------------------------------------------------------
// Star catalogue + equal-area sky grid (cos(theta) about +y x azimuth).
// ---------------------------------------------------------------------------
void star_catalogue(uint64_t seed, uint32_t n, float scale, std::vector<float4>& out) {
    out.resize(n);
    const uint32_t k0 = uint32_t(seed), k1 = uint32_t(seed >> 32) ^ 0x53544152u;
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t c[4] = {i, 0u, 0u, 0x73746172u};
        philox4x32_10(c, k0, k1);
        float cz = 2.0f * u01(c[0]) - 1.0f;
        float ph = 2.0f * kPi * u01(c[1]) - kPi;
        float m = 8.0f * u01(c[2]);
        float sn = std::sqrt(std::fmax(0.0f, 1.0f - cz * cz));
        out[i] = make_float4(sn * std::cos(ph), cz, sn * std::sin(ph), scale * std::pow(10.0f, -0.4f * m));
    }
}

void star_grid(const std::vector<float4>& st, float rho, uint32_t nc, uint32_t np, std::vector<uint32_t>& start,
               std::vector<uint32_t>& items) {
    std::vector<std::vector<uint32_t>> cells(size_t(nc) * np);
    const double pi = 3.14159265358979323846;
    const double half_c = 1.0 / nc, half_p = pi / np;   // half a cell of margin
    for (uint32_t i = 0; i < st.size(); ++i) {
        double c = st[i].y, th = std::acos(std::max(-1.0, std::min(1.0, c)));
        double ph = std::atan2(double(st[i].z), double(st[i].x));
        double t0 = th - rho, t1 = th + rho;
        double c_hi = t0 <= 0 ? 1.0 : std::cos(t0), c_lo = t1 >= pi ? -1.0 : std::cos(t1);
        c_hi = std::min(1.0, c_hi + half_c);
        c_lo = std::max(-1.0, c_lo - half_c);
        bool all_phi = t0 <= 0 || t1 >= pi;
        double dph = pi;
        if (!all_phi) {
            double smin = std::min(std::sin(t0), std::sin(t1));
            if (std::sin(rho) >= smin) all_phi = true;
            else dph = std::asin(std::sin(rho) / smin) + half_p;
        }
        int ic0 = std::max(0, int(std::floor((c_lo + 1.0) * 0.5 * nc)));
        int ic1 = std::min(int(nc) - 1, int(std::floor((c_hi + 1.0) * 0.5 * nc)));
        int ip0, ip1;
        if (all_phi || dph >= pi) { ip0 = 0; ip1 = int(np) - 1; }
        else {
            ip0 = int(std::floor((ph - dph + pi) / (2 * pi) * np));
            ip1 = int(std::floor((ph + dph + pi) / (2 * pi) * np));
        }
        for (int a = ic0; a <= ic1; ++a)
            for (int b = ip0; b <= ip1; ++b) {
                int bb = ((b % int(np)) + int(np)) % int(np);
                cells[size_t(a) * np + bb].push_back(i);
            }
    }
    start.assign(size_t(nc) * np + 1, 0);
    items.clear();
    for (size_t k = 0; k < cells.size(); ++k) {
        start[k] = uint32_t(items.size());
        // a star can land twice in one cell when its azimuth window wraps
        std::sort(cells[k].begin(), cells[k].end());
        cells[k].erase(std::unique(cells[k].begin(), cells[k].end()), cells[k].end());
        items.insert(items.end(), cells[k].begin(), cells[k].end());
    }
    start.back() = uint32_t(items.size());
}

e_magik_result_types ensure_stars(mgk_context* c, const mgk_scene_t* sc) {
    if (!(sc->flags & MGK_SCENE_STARS) || sc->n_star == 0) return MAGIK_SUCCESS;
    if (c->star_key == sc->star_seed && c->star_n == sc->n_star && c->star_rho == sc->star_radius_rad &&
        c->star_scale == sc->star_scale)
        return MAGIK_SUCCESS;
    std::vector<float4> st;
    star_catalogue(sc->star_seed, sc->n_star, sc->star_scale, st);
    std::vector<uint32_t> start, items;
    const uint32_t nc = 256, np = 512;
    star_grid(st, sc->star_radius_rad, nc, np, start, items);
    HIP_TRY(c->stars.reserve(st.size() * sizeof(float4)));
    HIP_TRY(c->cell_start.reserve(start.size() * sizeof(uint32_t)));
    HIP_TRY(c->cell_items.reserve(std::max<size_t>(1, items.size()) * sizeof(uint32_t)));
    HIP_TRY(hipMemcpy(c->stars.ptr, st.data(), st.size() * sizeof(float4), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->cell_start.ptr, start.data(), start.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    if (!items.empty()) HIP_TRY(hipMemcpy(c->cell_items.ptr, items.data(), items.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    c->grid_nc = nc;
    c->grid_np = np;
    c->star_key = sc->star_seed;
    c->star_n = sc->n_star;
    c->star_rho = sc->star_radius_rad;
    c->star_scale = sc->star_scale;
    return MAGIK_SUCCESS;
}

e_magik_result_types ensure_star_band(mgk_context* c, const mgk_scene_t* sc, float rho) {
    e_magik_result_types r = ensure_stars(c, sc);
    if (r != MAGIK_SUCCESS) return r;
    if (c->band_key == sc->star_seed && c->band_n == sc->n_star && c->band_rho == rho) return MAGIK_SUCCESS;
    std::vector<float4> st;
    star_catalogue(sc->star_seed, sc->n_star, sc->star_scale, st);
    std::vector<uint32_t> start, items;
    const uint32_t nc = 256, np = 512;
    star_grid(st, rho, nc, np, start, items);
    HIP_TRY(c->band_start.reserve(start.size() * sizeof(uint32_t)));
    HIP_TRY(c->band_items.reserve(std::max<size_t>(1, items.size()) * sizeof(uint32_t)));
    HIP_TRY(hipMemcpy(c->band_start.ptr, start.data(), start.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    if (!items.empty()) HIP_TRY(hipMemcpy(c->band_items.ptr, items.data(), items.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    c->band_nc = nc;
    c->band_np = np;
    c->band_key = sc->star_seed;
    c->band_n = sc->n_star;
    c->band_rho = rho;
    return MAGIK_SUCCESS;
}

e_magik_result_types fill_shade_params(mgk_context* c, const mgk_scene_t* sc, const mgk_spectral_t* sp, const mgk_camera_t* cam,
                                       uint64_t seed, uint32_t sample_begin, uint32_t spp, const mgk_hit_t* hits, float* xyz,
                                       ShadeParams& P) {
    if (sp->n_bin == 0 || sp->n_bin > uint32_t(kMaxBins) || sp->n_hero == 0 || sp->n_hero > 4 || !(sp->sigma_nm > 0.0f) ||
        sp->lambda_sampling > MGK_LAMBDA_CMF || (sp->lambda_sampling == MGK_LAMBDA_CMF && !(sp->hero_max_nm > sp->hero_min_nm)))
        return MGK_ERROR_INVALID_ARGUMENT;
    e_magik_result_types r = ensure_stars(c, sc);
    if (r != MAGIK_SUCCESS) return r;
    std::memset(&P, 0, sizeof(P));
    P.W = cam->x_resolution;
    P.H = cam->y_resolution;
    P.spp = spp;
    P.sample_begin = sample_begin;
    P.camera_id = cam->camera_id;
    P.seed = seed;
    P.hits = hits;
    P.xyz = xyz;
    P.flags = sc->flags;
    P.disk_r_in = sc->disk_r_in;
    P.disk_T_in = sc->disk_T_in;
    P.sphere_T = sc->sphere_T;
    P.sphere_checker = sc->sphere_checker ? sc->sphere_checker : 1;
    P.background_scale = sc->background_scale;
    P.background_T = sc->background_T;
    P.star_T = sc->star_T;
    {
        float ch = 2.0f * std::sin(0.5f * sc->star_radius_rad);
        P.chord2_star_r = ch * ch;
    }
    if ((sc->flags & MGK_SCENE_STARS) && sc->n_star > 0) {
        P.stars = reinterpret_cast<const float4*>(c->stars.ptr);
        P.cell_start = reinterpret_cast<const uint32_t*>(c->cell_start.ptr);
        P.cell_items = reinterpret_cast<const uint32_t*>(c->cell_items.ptr);
        P.grid_nc = c->grid_nc;
        P.grid_np = c->grid_np;
    } else {
        P.flags &= ~uint32_t(MGK_SCENE_STARS);
    }
    P.n_hero = sp->n_hero;
    P.hero_min = sp->hero_min_nm;
    P.hero_span = sp->hero_max_nm - sp->hero_min_nm;
    P.n_bin = sp->n_bin;
    P.bin_min = sp->bin_min_nm;
    P.bin_width = (sp->bin_max_nm - sp->bin_min_nm) / float(sp->n_bin);
    P.sigma = sp->sigma_nm;
    // radiance_to_irradiance (device_color_utilities.cuh:22-29) for throughput span / n_hero
    float inv_b = cam->sensor_width_mm / std::sqrt(float(cam->x_resolution * cam->x_resolution + cam->y_resolution * cam->y_resolution));
    float sw = float(cam->x_resolution) * inv_b, sh = float(cam->y_resolution) * inv_b;
    float area = (sw * sh) / 1e6f;
    float npx = float(cam->x_resolution * cam->y_resolution);
    float thr = P.hero_span / float(sp->n_hero);
    P.irr_scale = thr * (area / npx) * sp->iso * 2.0f * kPi / ((360.0f * sp->fps) / sp->shutter_angle);
    P.irr_base = (area / npx) * sp->iso * 2.0f * kPi / ((360.0f * sp->fps) / sp->shutter_angle);
    P.lambda_cmf = sp->lambda_sampling == MGK_LAMBDA_CMF ? 1u : 0u;
    if (P.lambda_cmf) cmf_inverse_cdf(sp->hero_min_nm, sp->hero_max_nm, P.inv_cdf);
    P.led_scale = sc->led_scale;
    bin_weights(sp, P.weights);
    HIP_TRY(c->shade_resp.reserve(size_t(kRespN) * sizeof(float4)));
    P.resp = reinterpret_cast<const float4*>(c->shade_resp.ptr);
    return MAGIK_SUCCESS;
}

e_magik_result_types fill_raygen(const mgk_integrator_t* in, const mgk_camera_t* cam, uint64_t seed, uint32_t sample_begin,
                                 uint32_t spp, void* rays, uint64_t stride, RayGenParams& P) {
    std::memset(&P, 0, sizeof(P));
    P.metric = in->metric_type;
    P.precision = in->precision;
    P.time_reversed = in->time_reversed;
    P.spin = signed_spin(in);
    P.mass_sign = mass_sign_of(in);
    P.length_scale = length_scale_of(in);
    for (int k = 0; k < 3; ++k) P.position[k] = cam->position[k];
    for (int k = 0; k < 9; ++k) P.orientation[k] = cam->orientation[k];
    P.focal_mm = cam->focal_length_mm;
    P.sensor_width_mm = cam->sensor_width_mm;
    P.W = cam->x_resolution;
    P.H = cam->y_resolution;
    P.camera_id = cam->camera_id;
    for (int k = 0; k < 3; ++k) P.observer_velocity[k] = cam->observer_velocity[k];
    P.seed = seed;
    P.sample_begin = sample_begin;
    P.spp = spp;
    P.rays = rays;
    P.stride = stride;
    if (P.W == 0 || P.H == 0) return MGK_ERROR_INVALID_ARGUMENT;
    if (uint64_t(P.W) * P.H * spp > stride) return MGK_ERROR_CAPACITY_EXCEEDED;
    return MAGIK_SUCCESS;
}

size_t ray_elem(const mgk_integrator_t* in) { return in->precision == MGK_PRECISION_FP32_REFERENCE ? 4 : 8; }

}  // namespace

extern "C" {

MGK_API e_magik_result_types magik_get_version(uint32_t* major, uint32_t* minor, uint32_t* revision, const char** as_char) {
    if (major) *major = 0;
    if (minor) *minor = 7;
    if (revision) *revision = 0;
    if (as_char) *as_char = "magik_amd 0.7.0 - gfx950";
    return MAGIK_SUCCESS;
}

MGK_API e_magik_result_types mgk_context_create(int device, mgk_context_handle* out) {
    if (!out) return MAGIK_ERROR_INVALID_POINTER;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return MAGIK_INVALID_CUDA_DEVICE;
    mgk_context* c = new (std::nothrow) mgk_context();
    if (!c) return MAGIK_ERROR_HOST_OUT_OF_MEMORY;
    c->device = device;
    auto fail = [&](e_magik_result_types r) {
        mgk_context_destroy(c);
        return r;
    };
    if (set_device(c) != MAGIK_SUCCESS) return fail(MGK_ERROR_HIP);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return fail(MGK_ERROR_HIP);
    c->n_cu = prop.multiProcessorCount;
    if (hipMalloc(&c->d_counters, kCtrTotal * sizeof(unsigned long long)) != hipSuccess) return fail(MAGIK_ERROR_DEVICE_ALLOCATION_FAILED);
    {
        const hipError_t e = scratch_address(device, &c->scratch);
        if (e != hipSuccess) {
            std::fprintf(stderr, "magik_amd: hipGetSymbolAddress(scratch) failed: %s\n", hipGetErrorString(e));
            return fail(MGK_ERROR_HIP);
        }
    }
    if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) return fail(MGK_ERROR_HIP);
    *out = c;
    return MAGIK_SUCCESS;
}

MGK_API e_magik_result_types mgk_context_destroy(mgk_context_handle c) {
    if (!c) return MAGIK_ERROR_INVALID_POINTER;
    (void)hipSetDevice(c->device);
    (void)hipDeviceSynchronize();
    if (c->d_counters) (void)hipFree(c->d_counters);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    c->stars.release();
    c->cell_start.release();
    c->cell_items.release();
    c->rays.release();
    c->hits.release();
    c->batch_params.release();
    c->ic_camera.release();
    c->shade_resp.release();
    c->mesh_nodes.release();
    c->mesh_tris.release();
    c->mesh_normals.release();
    c->mesh_inst.release();
    for (int k = 0; k < 2; ++k) {
        c->zone_state[k].release();
        c->zone_steps[k].release();
    }
    c->zone_q0.release();
    c->zone_q1.release();
    delete c;
    return MAGIK_SUCCESS;
}

MGK_API e_magik_result_types mgk_set_scene(mgk_context_handle c, const mgk_scene_t* sc) {
    if (!c || !sc) return MAGIK_ERROR_INVALID_POINTER;
    e_magik_result_types r = set_device(c);
    if (r != MAGIK_SUCCESS) return r;
    return ensure_stars(c, sc);
}

MGK_API e_magik_result_types mgk_set_meshes(mgk_context_handle c, const mgk_mesh_t* meshes, uint32_t n_mesh,
                                            const mgk_instance_t* instances, uint32_t n_instance) {
    if (!c) return MAGIK_ERROR_INVALID_POINTER;
    e_magik_result_types r = set_device(c);
    if (r != MAGIK_SUCCESS) return r;
    MeshBuild b;
    if ((r = build_meshes(meshes, n_mesh, instances, n_instance, b)) != MAGIK_SUCCESS) return r;
    c->mesh_valid = false;
    c->mesh_diffuse = false;
    if (n_mesh == 0) return MAGIK_SUCCESS;
    HIP_TRY(hipDeviceSynchronize());   // earlier traces may still read the old hierarchy
    struct Up { DeviceBuffer* buf; const std::vector<float4>* v; } ups[4] = {
        {&c->mesh_nodes, &b.nodes}, {&c->mesh_tris, &b.tris}, {&c->mesh_normals, &b.normals}, {&c->mesh_inst, &b.inst}};
    for (const Up& u : ups) {
        const size_t bytes = u.v->size() * sizeof(float4);
        HIP_TRY(u.buf->reserve(bytes));
        HIP_TRY(hipMemcpy(u.buf->ptr, u.v->data(), bytes, hipMemcpyHostToDevice));
    }
    c->mesh_root = b.root;
    for (int k = 0; k < 3; ++k) { c->mesh_lo[k] = b.lo[k]; c->mesh_hi[k] = b.hi[k]; }
    c->mesh_diffuse = b.any_diffuse;
    c->mesh_valid = true;
    return MAGIK_SUCCESS;
}

MGK_API e_magik_result_types mgk_generate_rays(mgk_context_handle c, const mgk_integrator_t* in, const mgk_camera_t* cam, uint64_t seed,
                                               uint32_t sample_begin, uint32_t spp, void* d_rays, uint64_t stride, void* stream) {
    if (!c || !cam || !d_rays) return MAGIK_ERROR_INVALID_POINTER;
    e_magik_result_types r = check_integrator(in);
    if (r != MAGIK_SUCCESS) return r;
    if ((r = set_device(c)) != MAGIK_SUCCESS) return r;
    RayGenParams P;
    if ((r = fill_raygen(in, cam, seed, sample_begin, spp, d_rays, stride, P)) != MAGIK_SUCCESS) return r;
    HIP_TRY(c->ic_camera.reserve(sizeof(IcCamera)));
    HIP_TRY(launch_generate_rays(P, reinterpret_cast<IcCamera*>(c->ic_camera.ptr), hipStream_t(stream)));
    return MAGIK_SUCCESS;
}

MGK_API e_magik_result_types mgk_trace_rays(mgk_context_handle c, const mgk_integrator_t* in, const mgk_scene_t* sc, const float cam_pos[3],
                                            const void* d_rays, uint64_t stride, uint64_t n, const uint32_t* d_queue, mgk_hit_t* d_hits,
                                            uint32_t* d_steps, void* stream) {
    if (!c || !sc || !cam_pos || !d_hits || (n > 0 && !d_rays)) return MAGIK_ERROR_INVALID_POINTER;
    e_magik_result_types r = check_integrator(in);
    if (r != MAGIK_SUCCESS) return r;
    if ((r = check_scene(sc)) != MAGIK_SUCCESS) return r;
    if (!d_queue && n > stride) return MGK_ERROR_CAPACITY_EXCEEDED;
    if ((r = set_device(c)) != MAGIK_SUCCESS) return r;
    TraceParams P;
    fill_trace_params(in, sc, cam_pos, P);
    P.rays = d_rays;
    P.stride = stride;
    P.n = n;
    P.queue = d_queue;
    P.hits = d_hits;
    P.steps_out = d_steps;
    return do_trace(c, P, hipStream_t(stream));
}

MGK_API e_magik_result_types mgk_trace_rays_noalloc(mgk_context_handle c, const mgk_integrator_t* in, const mgk_scene_t* sc,
                                                    const float cam_pos[3], const void* d_rays, uint64_t stride, uint64_t n,
                                                    const uint32_t* d_queue, mgk_hit_t* d_hits, uint32_t* d_steps, void* stream) {
    if (!c || !sc || !cam_pos || !d_hits || (n > 0 && !d_rays)) return MAGIK_ERROR_INVALID_POINTER;
    e_magik_result_types r = check_integrator(in);
    if (r != MAGIK_SUCCESS) return r;
    if ((r = check_scene(sc)) != MAGIK_SUCCESS) return r;
    if (!d_queue && n > stride) return MGK_ERROR_CAPACITY_EXCEEDED;
    if (in->precision == MGK_PRECISION_MIXED && in->metric_type != MINKOWSKI && n > kNoAllocZoneRays) {
        std::fprintf(stderr, "magik_amd: mgk_trace_rays_noalloc: %llu rays exceed the %llu a mixed-precision call may take\n",
                     (unsigned long long)n, (unsigned long long)kNoAllocZoneRays);
        return MGK_ERROR_CAPACITY_EXCEEDED;
    }
    if (!c->scratch) return MGK_ERROR_HIP;
    if ((r = set_device(c)) != MAGIK_SUCCESS) return r;
    TraceParams P;
    fill_trace_params(in, sc, cam_pos, P);
    P.rays = d_rays;
    P.stride = stride;
    P.n = n;
    P.queue = d_queue;
    P.hits = d_hits;
    P.steps_out = d_steps;
    const Workspace ws = scratch_workspace(c->scratch);
    return do_trace(c, P, hipStream_t(stream), nullptr, 0, &ws);
}

MGK_API e_magik_result_types mgk_trace_rays_final(mgk_context_handle c, const mgk_integrator_t* in, const mgk_scene_t* sc,
                                                  const float cam_pos[3], const void* d_rays, uint64_t stride, uint64_t n,
                                                  const uint32_t* d_queue, mgk_hit_t* d_hits, double* d_final, void* stream) {
    if (!c || !sc || !cam_pos || !d_hits || !d_final || (n > 0 && !d_rays)) return MAGIK_ERROR_INVALID_POINTER;
    e_magik_result_types r = check_integrator(in);
    if (r != MAGIK_SUCCESS) return r;
    if ((r = check_scene(sc)) != MAGIK_SUCCESS) return r;
    if (!d_queue && n > stride) return MGK_ERROR_CAPACITY_EXCEEDED;
    if ((r = set_device(c)) != MAGIK_SUCCESS) return r;
    TraceParams P;
    fill_trace_params(in, sc, cam_pos, P);
    P.rays = d_rays;
    P.stride = stride;
    P.n = n;
    P.queue = d_queue;
    P.hits = d_hits;
    P.final_state = d_final;
    return do_trace(c, P, hipStream_t(stream));
}

MGK_API e_magik_result_types mgk_trace_paths(mgk_context_handle c, const mgk_integrator_t* in, const mgk_scene_t* sc,
                                             const mgk_camera_t* cam, uint64_t seed, uint32_t sample_begin, uint32_t spp, void* d_rays,
                                             uint64_t stride, mgk_hit_t* d_hits, uint32_t* d_steps, double* d_final, void* stream) {
    if (!c || !sc || !cam || !d_hits || !d_rays) return MAGIK_ERROR_INVALID_POINTER;
    e_magik_result_types r = check_integrator(in);
    if (r != MAGIK_SUCCESS) return r;
    if ((r = check_scene(sc)) != MAGIK_SUCCESS) return r;
    if (spp == 0) return MGK_ERROR_INVALID_ARGUMENT;
    const uint64_t n = uint64_t(cam->x_resolution) * cam->y_resolution * spp;
    if (n > stride) return MGK_ERROR_CAPACITY_EXCEEDED;
    if ((r = set_device(c)) != MAGIK_SUCCESS) return r;
    TraceParams P;
    fill_trace_params(in, sc, cam->position, P);
    P.rays = d_rays;
    P.stride = stride;
    P.n = n;
    P.hits = d_hits;
    P.steps_out = d_steps;
    P.final_state = d_final;
    const PathInfo paths{seed, cam->camera_id, cam->x_resolution * cam->y_resolution, sample_begin};
    return do_trace(c, P, hipStream_t(stream), &paths, sc->max_reflections);
}

}  // extern "C"

e_magik_result_types mgk_trace_rays_per_ray(mgk_context_handle c, const mgk_integrator_t* in, const mgk_scene_t* sc, const float origin[3],
                                            const double* d_rays, uint64_t stride, uint64_t n, mgk_hit_t* d_hits, hipStream_t stream) {
    if (!c || !sc || !d_hits || (n > 0 && !d_rays)) return MAGIK_ERROR_INVALID_POINTER;
    e_magik_result_types r = check_integrator(in);
    if (r != MAGIK_SUCCESS) return r;
    if (in->precision == MGK_PRECISION_FP32_REFERENCE || n > stride) return MGK_ERROR_INVALID_ARGUMENT;
    if ((r = set_device(c)) != MAGIK_SUCCESS) return r;
    TraceParams P;
    fill_trace_params(in, sc, origin, P);
    P.rays = d_rays;
    P.stride = stride;
    P.n = n;
    P.hits = d_hits;
    P.per_ray = 1;
    P.disk_isco = 1;
    return do_trace(c, P, stream);
}

extern "C" {

MGK_API e_magik_result_types mgk_generate_rays_batch(mgk_context_handle c, const mgk_integrator_t* in, const mgk_batch_camera_t* cams,
                                                     uint32_t n_cam, uint64_t seed, uint32_t spp, double* d_rays, uint64_t stride,
                                                     uint64_t* offsets, void* stream) {
    if (!c || !cams || !offsets || (n_cam > 0 && !d_rays)) return MAGIK_ERROR_INVALID_POINTER;
    e_magik_result_types r = check_integrator(in);
    if (r != MAGIK_SUCCESS) return r;
    if (in->precision == MGK_PRECISION_FP32_REFERENCE || spp == 0 || n_cam > kMaxBatchCameras) return MGK_ERROR_INVALID_ARGUMENT;
    if ((r = set_device(c)) != MAGIK_SUCCESS) return r;
    std::vector<RayGenParams> P(n_cam);
    std::vector<uint64_t> off(size_t(n_cam) + 1, 0);
    for (uint32_t k = 0; k < n_cam; ++k) {
        mgk_integrator_t ik = *in;
        ik.spin = cams[k].spin;
        const uint64_t nk = uint64_t(cams[k].camera.x_resolution) * cams[k].camera.y_resolution * spp;
        off[k + 1] = off[k] + nk;
        if ((r = fill_raygen(&ik, &cams[k].camera, seed, 0, spp, d_rays, stride, P[k])) != MAGIK_SUCCESS) return r;
        P[k].per_ray_spin = 1;
    }
    for (uint32_t k = 0; k <= n_cam; ++k) offsets[k] = off[k];
    if (off[n_cam] > stride) return MGK_ERROR_CAPACITY_EXCEEDED;
    if (n_cam == 0 || off[n_cam] == 0) return MAGIK_SUCCESS;
    hipStream_t s = hipStream_t(stream);
    // layout: RayGenParams[n_cam] | offsets[n_cam + 1] | IcCamera[n_cam] (256-byte aligned)
    const size_t ic_at = (P.size() * sizeof(RayGenParams) + off.size() * sizeof(uint64_t) + 255) & ~size_t(255);
    HIP_TRY(c->batch_params.reserve(ic_at + size_t(n_cam) * sizeof(IcCamera)));
    RayGenParams* d_p = reinterpret_cast<RayGenParams*>(c->batch_params.ptr);
    uint64_t* d_off = reinterpret_cast<uint64_t*>(d_p + n_cam);
    IcCamera* d_ic = reinterpret_cast<IcCamera*>(static_cast<char*>(c->batch_params.ptr) + ic_at);
    HIP_TRY(hipMemcpyAsync(d_p, P.data(), P.size() * sizeof(RayGenParams), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d_off, off.data(), off.size() * sizeof(uint64_t), hipMemcpyHostToDevice, s));
    HIP_TRY(launch_generate_rays_batch(d_p, d_off, n_cam, off[n_cam], d_ic, s));
    HIP_TRY(hipStreamSynchronize(s));   // the host staging vectors go out of scope
    return MAGIK_SUCCESS;
}

MGK_API e_magik_result_types mgk_trace_rays_batch(mgk_context_handle c, const mgk_integrator_t* in, const mgk_scene_t* sc,
                                                  const double* d_rays, uint64_t stride, uint64_t n, const uint32_t* d_queue,
                                                  mgk_hit_t* d_hits, uint32_t* d_steps, void* stream) {
    if (!c || !sc || !d_hits || (n > 0 && !d_rays)) return MAGIK_ERROR_INVALID_POINTER;
    e_magik_result_types r = check_integrator(in);
    if (r != MAGIK_SUCCESS) return r;
    if (in->precision == MGK_PRECISION_FP32_REFERENCE) return MGK_ERROR_INVALID_ARGUMENT;
    if (!d_queue && n > stride) return MGK_ERROR_CAPACITY_EXCEEDED;
    if ((r = set_device(c)) != MAGIK_SUCCESS) return r;
    const float origin[3] = {0.0f, 0.0f, 0.0f};
    TraceParams P;
    fill_trace_params(in, sc, origin, P);
    P.rays = d_rays;
    P.stride = stride;
    P.n = n;
    P.queue = d_queue;
    P.hits = d_hits;
    P.steps_out = d_steps;
    P.per_ray = 1;
    P.disk_isco = 1;
    return do_trace(c, P, hipStream_t(stream));
}

MGK_API e_magik_result_types mgk_last_trace_stats(mgk_context_handle c, mgk_trace_stats_t* out) {
    if (!c || !out) return MAGIK_ERROR_INVALID_POINTER;
    std::memset(out, 0, sizeof(*out));
    if (!c->timed) return MAGIK_SUCCESS;
    e_magik_result_types r = set_device(c);
    if (r != MAGIK_SUCCESS) return r;
    HIP_TRY(hipEventSynchronize(c->ev1));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    unsigned long long h[3];
    HIP_TRY(hipMemcpy(h, c->d_counters + kCtrStats, sizeof(h), hipMemcpyDeviceToHost));
    out->n_rays = c->last_n;
    out->n_steps = h[0];
    out->n_trials = h[1];
    out->n_trials_fp64 = h[2];
    out->ms = double(ms);
    out->n_reflections = c->last_reflections;
    return MAGIK_SUCCESS;
}

MGK_API e_magik_result_types mgk_shade(mgk_context_handle c, const mgk_scene_t* sc, const mgk_spectral_t* sp, const mgk_camera_t* cam,
                                       uint64_t seed, uint32_t sample_begin, uint32_t spp, const mgk_hit_t* d_hits, float* d_xyz,
                                       void* stream) {
    if (!c || !sc || !sp || !cam || !d_hits || !d_xyz) return MAGIK_ERROR_INVALID_POINTER;
    if (spp == 0) return MGK_ERROR_INVALID_ARGUMENT;
    e_magik_result_types r = set_device(c);
    if (r != MAGIK_SUCCESS) return r;
    ShadeParams P;
    if ((r = fill_shade_params(c, sc, sp, cam, seed, sample_begin, spp, d_hits, d_xyz, P)) != MAGIK_SUCCESS) return r;
    HIP_TRY(launch_shade(P, hipStream_t(stream)));
    return MAGIK_SUCCESS;
}

MGK_API e_magik_result_types mgk_render_frame(mgk_context_handle c, const mgk_integrator_t* in, const mgk_scene_t* sc,
                                              const mgk_spectral_t* sp, const mgk_camera_t* cam, uint64_t seed, uint32_t spp,
                                              float* d_xyz, void* stream) {
    return mgk::render_samples(c, in, sc, sp, cam, seed, 0, spp, d_xyz, hipStream_t(stream));
}

}  // extern "C"

// Samples [sample_begin, sample_begin + spp) of a frame: generate + trace
// (whole paths) + shade, XYZ / spp per pixel.
e_magik_result_types mgk::render_samples(mgk_context_handle c, const mgk_integrator_t* in, const mgk_scene_t* sc,
                                         const mgk_spectral_t* sp, const mgk_camera_t* cam, uint64_t seed, uint32_t sample_begin,
                                         uint32_t spp, float* d_xyz, hipStream_t s, StepHistogram* hist) {
    if (!c || !sc || !sp || !cam || !d_xyz) return MAGIK_ERROR_INVALID_POINTER;
    e_magik_result_types r = check_integrator(in);
    if (r != MAGIK_SUCCESS) return r;
    if ((r = check_scene(sc)) != MAGIK_SUCCESS) return r;
    if (spp == 0) return MGK_ERROR_INVALID_ARGUMENT;
    if ((r = set_device(c)) != MAGIK_SUCCESS) return r;
    const uint64_t n = uint64_t(cam->x_resolution) * cam->y_resolution * spp;
    HIP_TRY(c->rays.reserve(n * 10 * ray_elem(in)));
    HIP_TRY(c->hits.reserve(n * sizeof(mgk_hit_t)));
    RayGenParams G;
    if ((r = fill_raygen(in, cam, seed, sample_begin, spp, c->rays.ptr, n, G)) != MAGIK_SUCCESS) return r;
    HIP_TRY(c->ic_camera.reserve(sizeof(IcCamera)));
    HIP_TRY(launch_generate_rays(G, reinterpret_cast<IcCamera*>(c->ic_camera.ptr), s));
    TraceParams P;
    fill_trace_params(in, sc, cam->position, P);
    P.rays = c->rays.ptr;
    P.stride = n;
    P.n = n;
    P.hits = reinterpret_cast<mgk_hit_t*>(c->hits.ptr);
    const PathInfo paths{seed, cam->camera_id, cam->x_resolution * cam->y_resolution, sample_begin};
    if ((r = do_trace(c, P, s, &paths, sc->max_reflections, nullptr, hist)) != MAGIK_SUCCESS) return r;
    ShadeParams S;
    if ((r = fill_shade_params(c, sc, sp, cam, seed, sample_begin, spp, reinterpret_cast<mgk_hit_t*>(c->hits.ptr), d_xyz, S)) !=
        MAGIK_SUCCESS)
        return r;
    HIP_TRY(launch_shade(S, s));
    return MAGIK_SUCCESS;
}

e_magik_result_types mgk_camera_constants(mgk_context_handle c, const mgk_integrator_t* in, const mgk_camera_t* cam, uint64_t seed,
                                          uint32_t spp, RayGenParams* params, const IcCamera** d_ic, void* stream) {
    if (!c || !cam || !params || !d_ic) return MAGIK_ERROR_INVALID_POINTER;
    e_magik_result_types r = check_integrator(in);
    if (r != MAGIK_SUCCESS) return r;
    if (in->precision == MGK_PRECISION_FP32_REFERENCE) return MGK_ERROR_INVALID_ARGUMENT;
    if ((r = set_device(c)) != MAGIK_SUCCESS) return r;
    const uint64_t n = uint64_t(cam->x_resolution) * cam->y_resolution * spp;
    if ((r = fill_raygen(in, cam, seed, 0, spp, nullptr, n, *params)) != MAGIK_SUCCESS) return r;
    HIP_TRY(c->ic_camera.reserve(sizeof(IcCamera)));
    HIP_TRY(launch_ic_camera(*params, reinterpret_cast<IcCamera*>(c->ic_camera.ptr), hipStream_t(stream)));
    *d_ic = reinterpret_cast<const IcCamera*>(c->ic_camera.ptr);
    return MAGIK_SUCCESS;
}

e_magik_result_types mgk_star_band_view(mgk_context_handle c, const mgk_scene_t* sc, float margin, StarView* out) {
    if (!c || !sc || !out) return MAGIK_ERROR_INVALID_POINTER;
    std::memset(out, 0, sizeof(*out));
    if (!(sc->flags & MGK_SCENE_STARS) || sc->n_star == 0) return MAGIK_SUCCESS;
    if (!(margin >= 0.0f) || !(sc->star_radius_rad > 0.0f)) return MGK_ERROR_INVALID_ARGUMENT;
    e_magik_result_types r = set_device(c);
    if (r != MAGIK_SUCCESS) return r;
    const float rho = sc->star_radius_rad + margin;
    if ((r = ensure_star_band(c, sc, rho)) != MAGIK_SUCCESS) return r;
    out->stars = reinterpret_cast<const float4*>(c->stars.ptr);
    out->cell_start = reinterpret_cast<const uint32_t*>(c->band_start.ptr);
    out->cell_items = reinterpret_cast<const uint32_t*>(c->band_items.ptr);
    out->nc = c->band_nc;
    out->np = c->band_np;
    out->rho = rho;
    return MAGIK_SUCCESS;
}

extern "C" {

MGK_API e_magik_result_types mgk_color_weights(const mgk_spectral_t* sp, float* weights_xyz, float* xyz_to_scene_out) {
    if (!sp) return MAGIK_ERROR_INVALID_POINTER;
    if (sp->n_bin == 0 || sp->n_bin > uint32_t(kMaxBins)) return MGK_ERROR_INVALID_ARGUMENT;
    if (weights_xyz) bin_weights(sp, weights_xyz);
    if (xyz_to_scene_out) {
        m3 m = xyz_to_scene(sp->white_point_K > 0.0f ? sp->white_point_K : 5600.0f);
        std::memcpy(xyz_to_scene_out, m.m, sizeof(m.m));
    }
    return MAGIK_SUCCESS;
}

MGK_API e_magik_result_types mgk_cmf_inverse_cdf(const mgk_spectral_t* sp, float* inverse_cdf_nm) {
    if (!sp || !inverse_cdf_nm) return MAGIK_ERROR_INVALID_POINTER;
    if (!(sp->hero_max_nm > sp->hero_min_nm)) return MGK_ERROR_INVALID_ARGUMENT;
    cmf_inverse_cdf(sp->hero_min_nm, sp->hero_max_nm, inverse_cdf_nm);
    return MAGIK_SUCCESS;
}

MGK_API e_magik_result_types mgk_display_transform(mgk_context_handle c, const mgk_display_t* d, const float* d_xyz, uint64_t n,
                                                   float* d_out, void* stream) {
    if (!c || !d || (n > 0 && (!d_xyz || !d_out))) return MAGIK_ERROR_INVALID_POINTER;
    if (n > (uint64_t(1) << 60) / 12) return MGK_ERROR_INVALID_ARGUMENT;
    // in place (d_out == d_xyz) or disjoint: a partial overlap would race between pixels
    const uintptr_t a = uintptr_t(d_xyz), b = uintptr_t(d_out), bytes = uintptr_t(n) * 12u;
    if (a != b && a < b + bytes && b < a + bytes) return MGK_ERROR_INVALID_ARGUMENT;
    e_magik_result_types r;
    if (!c->disp_valid || std::memcmp(&c->disp_key, d, sizeof(*d)) != 0) {
        c->disp_valid = false;
        if ((r = display_params(d, c->disp_params)) != MAGIK_SUCCESS) return r;
        c->disp_key = *d;
        c->disp_valid = true;
    }
    if ((r = set_device(c)) != MAGIK_SUCCESS) return r;
    HIP_TRY(launch_display(c->disp_params, d_xyz, n, d_out, hipStream_t(stream)));
    return MAGIK_SUCCESS;
}

MGK_API e_magik_result_types mgk_camera_boyer_lindquist(float r_obs, float theta_deg, float spin, float mass_kg, float fov_deg,
                                                        uint32_t x_res, uint32_t y_res, mgk_camera_t* out) {
    if (!out) return MAGIK_ERROR_INVALID_POINTER;
    if (!(r_obs > 0.0f) || !(fov_deg > 0.0f && fov_deg < 180.0f) || x_res == 0 || y_res == 0) return MGK_ERROR_INVALID_ARGUMENT;
    std::memset(out, 0, sizeof(*out));
    const double d2r = 3.14159265358979323846 / 180.0;
    double th = double(theta_deg) * d2r, r = r_obs, a = spin;
    // KS Cartesian point of BL (r, theta) at the azimuth where x = 0:
    // y = r cos(theta), z = sqrt(r^2 + a^2) sin(theta)
    double P[3] = {0.0, r * std::cos(th), std::sqrt(r * r + a * a) * std::sin(th)};
    double L = mass_kg == 0.0f ? 1.0 : double(kG) * double(mass_kg) / (double(kC) * double(kC));
    double n = std::sqrt(P[0] * P[0] + P[1] * P[1] + P[2] * P[2]);
    double zc[3] = {P[0] / n, P[1] / n, P[2] / n};   // camera looks down -z_cam, i.e. at the origin
    double up[3] = {0.0, 1.0, 0.0};
    double d = up[0] * zc[0] + up[1] * zc[1] + up[2] * zc[2];
    double yc[3] = {up[0] - d * zc[0], up[1] - d * zc[1], up[2] - d * zc[2]};
    double yl = std::sqrt(yc[0] * yc[0] + yc[1] * yc[1] + yc[2] * yc[2]);
    if (yl < 1e-9) return MGK_ERROR_INVALID_ARGUMENT;   // on the spin axis
    for (int k = 0; k < 3; ++k) yc[k] /= yl;
    double xc[3] = {yc[1] * zc[2] - yc[2] * zc[1], yc[2] * zc[0] - yc[0] * zc[2], yc[0] * zc[1] - yc[1] * zc[0]};
    for (int k = 0; k < 3; ++k) out->position[k] = float(P[k] * L);
    for (int rrow = 0; rrow < 3; ++rrow) {
        out->orientation[3 * rrow + 0] = float(xc[rrow]);
        out->orientation[3 * rrow + 1] = float(yc[rrow]);
        out->orientation[3 * rrow + 2] = float(zc[rrow]);
    }
    out->sensor_width_mm = 24.0f;
    out->focal_length_mm = float(12.0 / std::tan(0.5 * double(fov_deg) * d2r));
    out->x_resolution = x_res;
    out->y_resolution = y_res;
    return MAGIK_SUCCESS;
}

MGK_API e_magik_result_types magik_trace_geodesic(magik_trace_geodesic_container_t* container,
                                                  magik_trace_geodesic_inital_conditions_t ic,
                                                  magik_trace_geodesic_metric_parameters_t mp,
                                                  magik_trace_geodesic_integrator_parameters_t ip) {
    if (!container || !container->steps) return MAGIK_ERROR_INVALID_POINTER;
    if (container->capacity != sizeof(magik_trace_geodesic_step_t) * container->n_allocated_step)
        return MAGIK_ERROR_HOST_TRACE_ALLOCATED_STEPS_AND_CAPACITY_MISMATCH;
    if (!valid_metric(mp.metric_type)) return MAGIK_ERROR_UNKNOWN;
    container->n_written = 0;
    if (container->n_allocated_step == 0) return MAGIK_SUCCESS;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return MAGIK_INVALID_CUDA_DEVICE;
    static std::mutex mu;   // one in-flight host-API trace at a time
    std::lock_guard<std::mutex> lock(mu);
    HIP_TRY(hipSetDevice(0));
    GeodesicStepParams P;
    std::memset(&P, 0, sizeof(P));
    P.metric = mp.metric_type;
    P.time_reversed = mp.time_reversed ? 1u : 0u;
    P.n_alloc = container->n_allocated_step;
    float mass = mp.metric_type == KERR ? mp.kerr_mass : mp.schwarzschild_mass;
    P.spin = mp.metric_type == KERR ? (mp.time_reversed ? -mp.kerr_spin : mp.kerr_spin) : 0.0f;
    P.mass_sign = (mp.metric_type == SCHWARZSCHILD && mass <= 0.0f) ? -1.0f : 1.0f;
    P.length_scale = mp.metric_type == MINKOWSKI ? 1.0f : kG * (mp.metric_type == SCHWARZSCHILD ? std::fabs(mass) : mass) / (kC * kC);
    P.spin_in_range = std::fabs(P.spin) > 1.0f ? 0 : 1;
    P.horizon_r = mp.metric_type == KERR ? (1.0f + std::sqrt(1.0f - P.spin * P.spin)) * mp.event_horizon_scalar
                                         : 2.0f * mp.event_horizon_scalar;
    P.h_max = 1000.0f;
    P.abs_tol = ip.integrator_absolute_error_tolerance;
    P.rel_tol = ip.integrator_relative_error_tolerance;
    P.n_max = ip.n_integrator_max_substep;
    P.position[0] = ic.spatial_position_x;
    P.position[1] = ic.spatial_position_y;
    P.position[2] = ic.spatial_position_z;
    // rotate(x_rot, y_rot, z_rot) = m_x(z_rot) m_y(y_rot) m_z(x_rot) (host_trace_geodesic.cu:17)
    mgk::rotate_matrix(ic.x_rotation, ic.y_rotation, ic.z_rotation, P.rot);
    P.lpv[0] = ic.local_proper_velocity_x;
    P.lpv[1] = ic.local_proper_velocity_y;
    P.lpv[2] = ic.local_proper_velocity_z;
    P.massive = ic.test_particle_type == MASSIVE ? 1 : 0;
    P.speed = ic.massive_particle_speed;
    magik_trace_geodesic_step_t* d_steps = nullptr;
    uint32_t* d_n = nullptr;
    size_t bytes = sizeof(magik_trace_geodesic_step_t) * container->n_allocated_step;
    if (hipMalloc(&d_steps, bytes) != hipSuccess) return MAGIK_ERROR_DEVICE_ALLOCATION_FAILED;
    if (hipMalloc(&d_n, sizeof(uint32_t)) != hipSuccess) {
        (void)hipFree(d_steps);
        return MAGIK_ERROR_DEVICE_ALLOCATION_FAILED;
    }
    P.steps = d_steps;
    P.n_written = d_n;
    e_magik_result_types res = MAGIK_SUCCESS;
    uint32_t n = 0;
    if (launch_geodesic_steps(P, nullptr) != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(&n, d_n, sizeof(n), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(container->steps, d_steps, sizeof(magik_trace_geodesic_step_t) * n, hipMemcpyDeviceToHost) != hipSuccess)
        res = MGK_ERROR_HIP;
    else
        container->n_written = n;
    (void)hipFree(d_steps);
    (void)hipFree(d_n);
    return res;
}

}  // extern "C"
