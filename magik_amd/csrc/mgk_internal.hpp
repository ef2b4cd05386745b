// Launch-parameter structs shared by the kernels (mgk_kernels.hip) and the
// C-ABI layer (mgk_api.hip).
#pragma once

#include <cmath>
#include <vector>

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "magik_amd.h"

namespace mgk {

constexpr int kMaxBins = 64;
constexpr int kRespN = 8192;   // samples of the tabulated spectral response

struct RayGenParams {
    int metric, precision;
    uint32_t time_reversed;
    float spin, mass_sign, length_scale;
    float position[3];
    float orientation[9];
    float focal_mm, sensor_width_mm;
    uint32_t W, H, camera_id;
    float observer_velocity[3];
    uint64_t seed;
    uint32_t sample_begin, spp;
    void* rays;
    uint64_t stride;
    int per_ray_spin;            // also write the spin of the equations as component 10
};

constexpr uint32_t kMaxBatchCameras = 4096;

// FP64 initial conditions shared by all rays of one pinhole camera: every ray
// starts at the lens point, so the position, the metric there, the observer
// and the camera tetrad are per camera and a ray only adds its direction.
struct IcCamera {
    double q[4];      // x, y, z, t (dimensionless)
    double M[4][4];   // covariant momentum p_mu = sum_k M[mu][k] l_k, mu = x, y, z, t; l = (d_x, d_y, d_z, |d|) / L
    double w[4];      // u_obs . p = sum_k w[k] l_k
};

// Triangle meshes (mgk_mesh.hpp): hierarchy sentinels and record sizes.
constexpr uint32_t kBvhEnd = 0xFFFFFFFFu;   // `skip` past the last node of a hierarchy
constexpr uint32_t kLeafShift = 24;         // leaf info = (count << 24) | first
constexpr uint32_t kLeafMask = 0xFFFFFFu;
constexpr int kInstF4 = 6;                  // float4 per instance record
// Work counters of the persistent trace kernels: kHeads of them, kHeadStride
// u64 apart (128 bytes), at the start of the context's counter block.
constexpr uint32_t kHeads = 8;
constexpr uint32_t kHeadStride = 16;
constexpr int kHermiteSegments = 4;         // straight sub-chords that follow a Hermite chord
constexpr int kHermiteFine = 8;             // finer sub-chords within the first coarse one that hits
constexpr float kBaryEps = 1e-4f;           // barycentric slack of a Hermite crossing on its triangle

struct TraceParams {
    int metric, precision;
    const void* rays;
    uint64_t stride, n;
    const uint32_t* queue;
    mgk_hit_t* hits;
    uint32_t* steps_out;
    const uint32_t* steps_in;    // optional: step counts a ray starts from (paths continued after a reflection)
    double* final_state;         // optional: the state (q, p) at termination, 8 x stride doubles; for a
                                 // box hit, the chord point of the hit and the step's end momentum
    unsigned long long* head;    // work counter (zeroed by the launcher)
    unsigned long long* stats;   // [0] steps [1] trials [2] fp64 trials
    float spin, mass_sign, length_scale;
    float h_max, abs_tol, rel_tol;
    uint32_t n_max, step_limit;
    float horizon_r_f;           // (1 + sqrt(1 - a^2)) * scalar, FP32 as the reference computes it
    double horizon_r_d;
    int spin_in_range;           // |a| <= 1 (metric_kerr.cuh:171)
    float cam[3];
    float cel_r2_f;              // celestial radius^2 (SI)
    double cel_r2_d;
    uint32_t flags;
    float disk_r_in, disk_r_out;
    float sphere_c[3], sphere_r;
    double zone_rho2;            // mixed precision: FP64 trials while x^2+y^2+z^2 < this
    int per_ray;                 // spin in ray component 10; escape sphere about each ray's origin
    float horizon_scalar;
    int disk_isco;               // per-ray rays: the disk starts at the ISCO of the ray's spin
    uint32_t n_box;              // axis-aligned boxes (SI metres) and their matter
    float box_min[MGK_MAX_BOX][3], box_max[MGK_MAX_BOX][3];
    uint32_t box_matter[MGK_MAX_BOX];
    // triangle meshes (MGK_SCENE_MESHES): the two-level hierarchy, its world bounds, the chord mode
    const float4* mesh_nodes;
    const float4* mesh_tris;
    const float4* mesh_inst;
    uint32_t mesh_root;          // top-level root node
    uint32_t chord;              // mgk_chord_t
    float mesh_lo[3], mesh_hi[3];
};

// Reflection of the paths that ended on diffuse box matter: new initial
// conditions in place in the ray buffer, and the reflected ray ids appended
// to a queue for the next trace.
struct ReflectParams {
    int metric;
    const mgk_hit_t* hits;
    const double* state;             // the trace's final_state (hit point, momentum), 8 x stride
    void* rays;                      // 10 x stride, float (FP32 reference path) or double
    uint64_t stride;
    const uint32_t* queue_in;        // the rays just traced (NULL: 0 .. n_in-1)
    uint64_t n_in;
    uint32_t* queue_out;
    unsigned long long* count;
    uint64_t seed;
    uint32_t camera_id, npix, sample_begin, block;   // Philox stream of reflection k: block 1 + k
    float spin, mass_sign;
    const float4* mesh_inst;         // mesh hits: instance records and object-space unit normals
    const float4* mesh_normals;
};

// Doubles parked per handed-off ray: q (4), p (4), h, u.p, L0, Q0, and the
// step controller's memory (last accepted h and ln error).
constexpr int kParked = 14;

// Hand-off queue slots are reserved kOutBlock at a time per wave: one atomic
// per block of hand-offs instead of one per wave and iteration.  Slots of a
// wave's last block that stay unused hold kHole and are skipped by the next
// pass; a pass reserves at most n_in + (waves of the launch) * kOutBlock slots.
constexpr uint32_t kOutBlock = 256;
constexpr uint32_t kHole = 0xFFFFFFFFu;

// One pass of the zone-split mixed-precision integrator.
struct ZoneParams {
    int from_ic;                     // load rays from the IC buffer (first pass) instead of the parked state
    int no_handoff;                  // integrate to termination regardless of the zone
    const uint32_t* in_queue;        // ray ids of this pass (NULL: 0 .. n_in-1); kHole entries are skipped
    uint64_t n_in;
    uint32_t* out_queue;             // rays handed to the other zone
    unsigned long long* out_count;
    uint64_t cap;                    // capacity of the queues and the parked-state arrays
    const double* state_in;          // states parked by the previous pass, kParked x cap doubles, at the queue slot
    const uint32_t* steps_in;        // their step counts
    double* state_out;               // states parked by this pass, at their out-queue slot
    uint32_t* steps_out;
};

constexpr int kCmfCdfN = 256;   // entries of the CMF inverse CDF (host_color_utilities.cuh:10)

struct ShadeParams {
    uint32_t W, H, spp, sample_begin, camera_id;
    uint64_t seed;
    const mgk_hit_t* hits;
    float* xyz;
    uint32_t flags;
    float disk_r_in, disk_T_in;
    float sphere_T;
    uint32_t sphere_checker;
    float background_scale, background_T;
    float star_T, chord2_star_r;
    const float4* stars;
    const uint32_t* cell_start;
    const uint32_t* cell_items;
    uint32_t grid_nc, grid_np;
    uint32_t n_hero;
    float hero_min, hero_span;
    uint32_t n_bin;
    float bin_min, bin_width, sigma;
    float irr_scale;             // radiance_to_irradiance factor per unit L, uniform sampling's throughput included
    float irr_base;              // the same factor per unit L * throughput (inverse-CDF sampling)
    float led_scale;             // illuminant boxes: led_scale x white LED spectrum
    uint32_t lambda_cmf;         // hero wavelengths through the CMF inverse CDF (else uniform)
    float inv_cdf[kCmfCdfN];     // its 256 wavelengths at u = k / 255
    float weights[3 * kMaxBins];
    const float4* resp;          // spectral response R(lambda), kRespN samples over [hero_min, hero_min + hero_span]
};

struct GeodesicStepParams {
    int metric;
    uint32_t time_reversed, n_alloc;
    float spin, mass_sign, length_scale, horizon_r, h_max, abs_tol, rel_tol;
    uint32_t n_max;
    int spin_in_range;
    float position[3];
    float rot[9];
    float lpv[3];
    int massive;
    float speed;
    magik_trace_geodesic_step_t* steps;
    uint32_t* n_written;
};

// Device view of a star catalogue's sky grid (256 bins in cos(theta) x 512 in
// phi, CSR): every star is listed in each cell its disc of radius `rho` meets.
struct StarView {
    const float4* stars;          // unit direction (x, y = cos(theta), z) and brightness
    const uint32_t* cell_start;
    const uint32_t* cell_items;
    uint32_t nc, np;
    float rho;
};

namespace disp { struct Params; }
hipError_t launch_display(const disp::Params& p, const float* xyz, uint64_t n, float* out, hipStream_t s);

// d_ic: device space for the per-camera constants (one IcCamera per camera)
hipError_t launch_generate_rays(const RayGenParams& p, IcCamera* d_ic, hipStream_t s);
hipError_t launch_ic_camera(const RayGenParams& p, IcCamera* d_ic, hipStream_t s);
hipError_t launch_generate_rays_batch(const RayGenParams* d_cams, const uint64_t* d_offsets, uint32_t n_cam, uint64_t n, IcCamera* d_ic,
                                      hipStream_t s);
hipError_t launch_trace(const TraceParams& p, int grid, hipStream_t s);
hipError_t launch_trace_zone(const TraceParams& p, const ZoneParams& z, bool near, int grid, hipStream_t s);
hipError_t launch_shade(const ShadeParams& p, hipStream_t s);
hipError_t launch_reflect(const ReflectParams& p, int precision, hipStream_t s);
// Adds each path (all n, or queue[0..n)) to bin clamp(steps, 1, n_bin - 1) of
// limit_out (hit class MGK_HIT_STEP_LIMIT) or end_out (any other, or hits NULL).
hipError_t launch_step_hist(const uint32_t* steps, const mgk_hit_t* hits, const uint32_t* queue, uint64_t n, uint32_t n_bin,
                            uint32_t* end_out, uint32_t* limit_out, int n_cu, hipStream_t s);

// Per-path step counts of one render, binned by the step at which each event
// happened: end[k] paths whose geodesic ended at step k (a hit, escape or
// horizon), limit[k] paths stopped there by the step limit, reflect[k]
// reflections taken there.  Bin 0 stays empty; steps past the last bin count
// in it.  All three have the same length.
struct StepHistogram {
    std::vector<uint32_t> end, limit, reflect;
};

// Samples [sample_begin, sample_begin + spp) of a frame into d_xyz (XYZ / spp
// per pixel): the body of mgk_render_frame.  With hist, also the render's step
// histograms (one more kernel and a synchronous copy).
e_magik_result_types render_samples(mgk_context_handle c, const mgk_integrator_t* in, const mgk_scene_t* sc, const mgk_spectral_t* sp,
                                    const mgk_camera_t* cam, uint64_t seed, uint32_t sample_begin, uint32_t spp, float* d_xyz,
                                    hipStream_t s, StepHistogram* hist = nullptr);

// mat3 rotate(a, b, g) of internal_math.h:1784-1810: m_x(g) m_y(b) m_z(a), row-major.
inline void rotate_matrix(float a, float b, float g, float out[9]) {
    const float ca = std::cos(a), sa = std::sin(a), cb = std::cos(b), sb = std::sin(b), cg = std::cos(g), sg = std::sin(g);
    const float mz[9] = {ca, -sa, 0, sa, ca, 0, 0, 0, 1}, my[9] = {cb, 0, sb, 0, 1, 0, -sb, 0, cb}, mx[9] = {1, 0, 0, 0, cg, -sg, 0, sg, cg};
    float t[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) t[3 * i + j] = mx[3 * i] * my[j] + mx[3 * i + 1] * my[3 + j] + mx[3 * i + 2] * my[6 + j];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) out[3 * i + j] = t[3 * i] * mz[j] + t[3 * i + 1] * mz[3 + j] + t[3 * i + 2] * mz[6 + j];
}
hipError_t launch_geodesic_steps(const GeodesicStepParams& p, hipStream_t s);


// Host build of the mesh hierarchy (mgk_mesh.hip): nodes, triangle slots,
// normals and instance records as laid out in mgk_mesh.hpp, all meshes'
// bottom levels first, then the top level.
struct MeshBuild {
    std::vector<float4> nodes, tris, normals, inst;
    std::vector<uint32_t> mesh_roots;   // each mesh's bottom-level root, in mesh order
    uint32_t root = kBvhEnd;
    float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    bool any_diffuse = false;
};
e_magik_result_types build_meshes(const mgk_mesh_t* meshes, uint32_t n_mesh, const mgk_instance_t* inst, uint32_t n_inst,
                                  MeshBuild& out);

}  // namespace mgk

// Internal (not exported): a sky grid of the scene's stars whose discs are
// widened to star_radius_rad + margin (the surrogate's star-edge test); the
// grid stays valid until the next call on the context.
e_magik_result_types mgk_star_band_view(mgk_context_handle ctx, const mgk_scene_t* scene, float margin, mgk::StarView* out);

// Internal (not exported): the FP64 per-camera constants of mgk_generate_rays
// (FP64 and mixed precision) computed into the context, without generating the
// rays: *params and *d_ic then give any path's initial conditions through
// ray_ic_fp64 (mgk_raygen.hpp), bit for bit those mgk_generate_rays writes.
// d_ic stays valid until the next ray generation on the context.
e_magik_result_types mgk_camera_constants(mgk_context_handle ctx, const mgk_integrator_t* in, const mgk_camera_t* cam, uint64_t seed,
                                          uint32_t spp, mgk::RayGenParams* params, const mgk::IcCamera** d_ic, void* stream);
