/*
 * magik_amd.h - C ABI of the MI355X-native wavefront spectral geodesic tracer.
 *
 * Two groups of entry points:
 *
 *  1. A layout-compatible subset of the reference's public header
 *     (include/magik/magik.h of ErikHall1801/Magik): result codes, metric and
 *     particle enums, the magik_trace_geodesic_* structs and
 *     magik_trace_geodesic() itself, which here runs on the GPU.
 *
 *  2. Batched device entry points for the render hot path (ray generation +
 *     initial conditions, persistent RKF45 geodesic integration, spectral
 *     shading, the MLP surrogate).  They replace the internal seam
 *     wavefront_sample() -> launch_kernel_<stage>_async() of the reference
 *     (src/renderer/wavefront/wavefront.cpp:79-120, kernels/kernel_*.cuh:9).
 *     All pointers named d_* are device pointers; streams are hipStream_t
 *     passed as void* so this header needs no HIP include.
 *
 * Every function returns e_magik_result_types; none calls exit().
 */
#ifndef MAGIK_AMD_H
#define MAGIK_AMD_H

#include <stdint.h>
#include <stddef.h>
#ifndef __cplusplus
#include <stdbool.h>
#endif

#define MGK_API __attribute__((visibility("default")))

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------------
 * 1. Reference-compatible subset (values and layouts as magik.h:38-110,
 *    :583-656).
 * ------------------------------------------------------------------------- */
typedef enum e_magik_result_types {
    MAGIK_SUCCESS = 0,
    MAGIK_ERROR_NOT_RUNNING = 1,
    MAGIK_ERROR_INVALID_POINTER = 2,
    MAGIK_UNKNOWN_ENUM_TYPE = 3,
    MAGIK_INVALID_CUDA_DEVICE = 4,
    MAGIK_ERROR_HOST_OUT_OF_MEMORY = 100,
    MAGIK_ERROR_HOST_MEMORY_ALLOCATION_FAILED = 101,
    MAGIK_ERROR_HOST_TRACE_ALLOCATED_STEPS_AND_CAPACITY_MISMATCH = 102,
    MAGIK_ERROR_DEVICE_OUT_OF_MEMORY = 200,
    MAGIK_ERROR_DEVICE_ALLOCATION_FAILED = 201,
    MAGIK_ERROR_INVALID_ID = 400,
    MAGIK_ERROR_PROVIDED_ID_NOT_FOUND = 401,
    MAGIK_ERROR_COMMAND_BUFFER_OVERFLOW = 402,
    MAGIK_ERROR_INVALID_COMMAND = 403,
    MAGIK_ERROR_PACKED_COMMAND_NOT_ALLIGNED = 404,
    MAGIK_ERROR_COMMAND_DROPPED = 405,
    MAGIK_ERROR_COMMAND_BUFFER_ALLOCATION_FAILED = 406,
    MAGIK_ERROR_COMMAND_SIZE_NOT_A_MULTIPLE_OF_4 = 407,
    MAGIK_ERROR_OUT_OF_BOUNDS_COMMAND_BUFFER_READ = 408,
    MAGIK_ERROR_UNKNOWN_METRIC = 707,
    MAGIK_ERROR_AOV_INCORRECT_EXTRACT_CALL = 900,
    MAGIK_ERROR_AOV_ALLOCATED_BEFORE_INITIALIZATION = 901,
    MAGIK_ERROR_AOV_SWAPCHAIN_NOT_INITALIZED = 902,
    MAGIK_ERROR_GL_LOADER_FAILED = 1000,
    MAGIK_ERROR_UNKNOWN = 10000,
    /* Additions of this library (outside the reference's ranges). */
    MGK_ERROR_INVALID_ARGUMENT = 20000,
    MGK_ERROR_CAPACITY_EXCEEDED = 20001,
    MGK_ERROR_HIP = 20002,
    MAGIK_PROHIBITED_FORCE_SIZE = 0x7FFFFFFF
} e_magik_result_types;

typedef enum e_magik_spacetime_metric_types { MINKOWSKI = 0, SCHWARZSCHILD = 1, KERR = 2 } e_magik_spacetime_metric_types;

typedef enum e_magik_test_particle_types { MASSIVE = 0, MASSLESS = 1 } e_magik_test_particle_types;

typedef struct magik_trace_geodesic_metric_parameters_t {
    e_magik_spacetime_metric_types metric_type;
    bool time_reversed;
    float event_horizon_scalar;
    float schwarzschild_mass; /* kg */
    float kerr_mass;          /* kg */
    float kerr_spin;
} magik_trace_geodesic_metric_parameters_t;

typedef struct magik_trace_geodesic_integrator_parameters_t {
    unsigned int n_integrator_max_substep;
    float integrator_absolute_error_tolerance;
    float integrator_relative_error_tolerance;
} magik_trace_geodesic_integrator_parameters_t;

typedef struct magik_trace_geodesic_inital_conditions_t {
    e_magik_test_particle_types test_particle_type;
    float spatial_position_x;
    float spatial_position_y;
    float spatial_position_z;
    float x_rotation;
    float y_rotation;
    float z_rotation;
    float local_proper_velocity_x;
    float local_proper_velocity_y;
    float local_proper_velocity_z;
    float massive_particle_speed;
} magik_trace_geodesic_inital_conditions_t;

typedef struct magik_trace_geodesic_step_t {
    float four_position[4];
    float four_momentum[4];
    float moment_derivative[4];
    float four_vector_contraction;
    float affine_parameter;
    float delta_affine_parameter;
} magik_trace_geodesic_step_t;

typedef struct magik_trace_geodesic_container_t {
    uint64_t capacity;
    uint32_t n_allocated_step;
    uint32_t n_written;
    magik_trace_geodesic_step_t* steps;
} magik_trace_geodesic_container_t;

/* Replaces magik_trace_geodesic (magik.h:656, src/magik.cpp:196-207,
 * relativity/host_trace_geodesic.cu:132-144): same arguments, same
 * capacity check, same per-step records (FP32, SI positions); the geodesic is
 * integrated by a HIP kernel on device 0 instead of the caller's CPU. */
MGK_API e_magik_result_types magik_trace_geodesic(magik_trace_geodesic_container_t* container,
                                                  magik_trace_geodesic_inital_conditions_t initial_conditions,
                                                  magik_trace_geodesic_metric_parameters_t metric_parameters,
                                                  magik_trace_geodesic_integrator_parameters_t integrator_parameters);

/* Colour pipeline enums (values as magik.h:330-383). */
typedef enum e_magik_scene_colorspace_types {
    MAGIK_SCENE_COLORSPACE_CIE_XYZ = 0,
    MAGIK_SCENE_COLORSPACE_CIE_RGB = 1,
    MAGIK_SCENE_COLORSPACE_LINEAR_REC709 = 2,
    MAGIK_SCENE_COLORSPACE_LINEAR_REC2020 = 3,
    MAGIK_SCENE_COLORSPACE_ACES_CG = 4,
    MAGIK_SCENE_COLORSPACE_ADOBE_RGB_1998 = 5
} e_magik_scene_colorspace_types;

typedef enum e_magik_display_colorspace_types {
    MAGIK_DISPLAY_COLORSPACE_SRGB = 0,
    MAGIK_DISPLAY_COLORSPACE_REC709 = 1,
    MAGIK_DISPLAY_COLORSPACE_REC2100_PQ = 2,
    MAGIK_DISPLAY_COLORSPACE_REC2100_HLG = 3,
    MAGIK_DISPLAY_COLORSPACE_DCI_P3_THEATRICAL_WHITE_POINT = 4,
    MAGIK_DISPLAY_COLORSPACE_DISPLAY_P3 = 5,
    MAGIK_DISPLAY_COLORSPACE_ADOBE_RGB = 6
} e_magik_display_colorspace_types;

typedef enum e_magik_view_transform_types {
    MAGIK_VIEW_TRANSFORM_NONE = 0,
    MAGIK_VIEW_TRANSFORM_REINHARD_SIMPLE = 1,
    MAGIK_VIEW_TRANSFORM_REINHARD_EXTENDED = 2,
    MAGIK_VIEW_TRANSFORM_GRAN_TURISMO_7 = 3,
    MAGIK_VIEW_TRANSFORM_AGX = 4,
    MAGIK_VIEW_TRANSFORM_KHRONOS_PBR_NEUTRAL = 5
} e_magik_view_transform_types;

typedef enum e_magik_gran_turismo_7_mode_types {
    MAGIK_GRAN_TURISMO_7_MODE_ICTCP = 0,
    MAGIK_GRAN_TURISMO_7_MODE_JZAZBZ = 1
} e_magik_gran_turismo_7_mode_types;

typedef enum e_magik_agx_look_types { MAGIK_AGX_LOOK_NONE = 0, MAGIK_AGX_LOOK_GOLDEN = 1, MAGIK_AGX_LOOK_PUNCHY = 2 } e_magik_agx_look_types;

typedef enum e_magik_output_format_types {
    MAGIK_OUTPUT_FORMAT_UNORM8 = 0,
    MAGIK_OUTPUT_FORMAT_UNORM10 = 1,
    MAGIK_OUTPUT_FORMAT_UNORM12 = 2,
    MAGIK_OUTPUT_FORMAT_UNORM16 = 3,
    MAGIK_OUTPUT_FORMAT_FLOAT16 = 4,
    MAGIK_OUTPUT_FORMAT_FLOAT32 = 5
} e_magik_output_format_types;

/* magik.h:262 (version query). */
MGK_API e_magik_result_types magik_get_version(uint32_t* major, uint32_t* minor, uint32_t* revision, const char** as_char);

/* ---------------------------------------------------------------------------
 * 2. Batched render path.
 * ------------------------------------------------------------------------- */

/* Arithmetic of the geodesic integrator. */
typedef enum mgk_precision_t {
    /* All FP32 in the reference's operation order and controller
     * (internal_math.h:2160-2205): parity with the reference integrator. */
    MGK_PRECISION_FP32_REFERENCE = 0,
    /* FP32 right-hand side and stage states; FP64 state accumulation, error
     * estimate and accept/reject; full FP64 trials inside the near zone
     * (r < max(1.5 r_photon, r_+ + 0.1)); FP64 null-constraint
     * renormalisation. */
    MGK_PRECISION_MIXED = 1,
    /* All FP64: the GPU twin of the FP64 oracle. */
    MGK_PRECISION_FP64 = 2
} mgk_precision_t;

/* Termination classes of a path (the reference maps escape and horizon both
 * to matter slot 3, kernel_intersect.cu:111,:119; they are kept apart here). */
typedef enum mgk_hit_class_t {
    MGK_HIT_NONE = 0,
    MGK_HIT_HORIZON = 1,
    MGK_HIT_ESCAPE = 2,
    MGK_HIT_DISK = 3,
    MGK_HIT_SPHERE = 4,
    MGK_HIT_STEP_LIMIT = 5,
    /* diffuse box matter still hit after the last allowed reflection
     * (kernel_enforce_ray_limits.cu:48-50, n_reflected > limit): no radiance */
    MGK_HIT_BOX_DIFFUSE = 6,
    /* illuminant box matter (matter_examples.cuh:63-76) */
    MGK_HIT_BOX_EMITTER = 7,
    /* triangle-mesh matter (MGK_SCENE_MESHES), diffuse and illuminant; the hit
     * record's a0 is the triangle's index in its mesh, a1 the instance index */
    MGK_HIT_MESH_DIFFUSE = 8,
    MGK_HIT_MESH_EMITTER = 9
} mgk_hit_class_t;

/* Integrator / spacetime settings (relativity_settings.h:11-25,
 * render_settings.h:16, metric_*::make_params). */
typedef struct mgk_integrator_t {
    int32_t metric_type;         /* e_magik_spacetime_metric_types */
    int32_t precision;           /* mgk_precision_t */
    uint32_t time_reversed;      /* 1 on the render path (kernel_advance_geodesic.cu:87) */
    uint32_t n_max_substep;      /* RKF45 trials per step (56) */
    float mass_kg;               /* Kerr or Schwarzschild mass; < 0 = negative-mass Schwarzschild */
    float spin;                  /* dimensionless a (Kerr) as given by the user; negated when time_reversed */
    float event_horizon_scalar;  /* 1.01 */
    float h_max;                 /* 1000 */
    float abs_tol;
    float rel_tol;
    uint32_t step_limit;         /* terminate when n_step > step_limit */
    uint32_t reserved;
} mgk_integrator_t;

/* Pinhole camera (optical_system_settings.h:20-34, device_sensor_utilities.cuh,
 * thin_lens.cuh).  sensor_width_mm is the reference's sensor_diagonal_mm,
 * which acts as the sensor width. */
typedef struct mgk_camera_t {
    float position[3];           /* SI metres */
    float orientation[9];        /* row-major camera -> world */
    float focal_length_mm;
    float sensor_width_mm;
    uint32_t x_resolution;
    uint32_t y_resolution;
    float observer_velocity[3];  /* local proper velocity, m/s */
    uint32_t camera_id;          /* RNG stream */
} mgk_camera_t;

enum {
    MGK_SCENE_DISK = 1u,
    MGK_SCENE_SPHERE = 2u,
    MGK_SCENE_STARS = 4u,
    MGK_SCENE_BACKGROUND = 8u,
    MGK_SCENE_BOXES = 16u,
    MGK_SCENE_MESHES = 32u,      /* the context's triangle meshes (mgk_set_meshes) */
    /* With MGK_SCENE_DISK: locate the disk crossing on the geodesic (one RKF45
     * trial from the step's start to the chord's crossing, then a Newton step
     * along the tangent to the plane) instead of on the step's chord, whose
     * crossing is off by O(h^2) of the step's curvature. */
    MGK_SCENE_DISK_ON_GEODESIC = 64u
};

/* How a completed step's chord meets triangle meshes.  AFFINE: the straight
 * segment between the step's end points, as the reference's box test
 * (kernel_intersect.cu:63-70).  HERMITE: the cubic Hermite curve through the
 * end points with the geodesic's tangents dx/dlambda = g^{mu nu} p_nu there,
 * the curved chord the reference reserves (wavefront.cpp:94-100).  Boxes keep
 * the affine chord. */
typedef enum mgk_chord_t {
    MGK_CHORD_AFFINE = 0,
    MGK_CHORD_HERMITE = 1
} mgk_chord_t;

/* Matter of a box: the reference's registered matter slots 1 and 2
 * (kernel_bxdf.cu:126, matter_examples.cuh:36-76). */
typedef enum mgk_matter_t {
    MGK_MATTER_DIFFUSE = 1,      /* specular on faces with an x normal, Lambertian on the others */
    MGK_MATTER_ILLUMINANT = 2    /* white LED spectrum x led_scale */
} mgk_matter_t;

#define MGK_MAX_BOX 4

/* A triangle mesh in object space, SI metres, in host memory.  Triangles are
 * two-sided; their geometric normal is (v1 - v0) x (v2 - v0). */
typedef struct mgk_mesh_t {
    const float* vertices;       /* n_vertex x 3 */
    uint32_t n_vertex;
    uint32_t n_triangle;
    const uint32_t* indices;     /* n_triangle x 3 vertex indices */
} mgk_mesh_t;

/* A placed mesh: its matter and the row-major 3 x 4 affine object -> world map
 * (invertible). */
typedef struct mgk_instance_t {
    uint32_t mesh;               /* index into the meshes of mgk_set_meshes */
    uint32_t matter;             /* mgk_matter_t */
    float object_to_world[12];
} mgk_instance_t;

/* Triangles per mesh and over all meshes, and instances: hit records carry the
 * triangle and instance indices as floats, exact below 2^24. */
#define MGK_MAX_MESH_TRIANGLES 16777215u
#define MGK_MAX_INSTANCES 16777215u

/* An axis-aligned box in SI metres (stage.h:19-22; traceables.cuh:17-74). */
typedef struct mgk_box_t {
    float min[3];
    float max[3];
    uint32_t matter;             /* mgk_matter_t */
    uint32_t reserved;
} mgk_box_t;

/* Traceables and emitters. */
typedef struct mgk_scene_t {
    float celestial_radius_m;    /* escape when |x - camera| >= this (traceables.cuh:76-86) */
    uint32_t flags;              /* MGK_SCENE_* */
    float disk_r_in, disk_r_out; /* thin disk in the y = 0 plane, units of GM/c^2 */
    float disk_T_in;             /* T(r) = disk_T_in (r / disk_r_in)^-3/4 */
    float sphere_center[3];      /* units of GM/c^2 (metres for Minkowski) */
    float sphere_radius;
    float sphere_T;
    uint32_t sphere_checker;     /* N x N checkerboard in (theta, phi) */
    float background_scale;      /* uniform background: scale * B(lambda, background_T) */
    float background_T;
    float star_scale;            /* star: star_scale 10^(-0.4 m) B(lambda, star_T) */
    float star_T;
    float star_radius_rad;
    uint32_t n_star;
    uint64_t star_seed;
    /* Boxes (MGK_SCENE_BOXES), tested in order after the disk and the sphere;
     * a later box wins only when strictly closer (kernel_intersect.cu:83-105). */
    uint32_t n_box;
    uint32_t max_reflections;    /* render_settings.h:17 n_reflected_limit (10) */
    float led_scale;             /* illuminant radiance scale (matter_examples.cuh:70: 7e5) */
    uint32_t chord;              /* mgk_chord_t for the meshes */
    mgk_box_t box[MGK_MAX_BOX];
} mgk_scene_t;

/* Wavelength sampling of a path's hero wavelengths. */
typedef enum mgk_lambda_sampling_t {
    MGK_LAMBDA_UNIFORM = 0,      /* stratified over [hero_min, hero_max], weight span / n_hero */
    MGK_LAMBDA_CMF = 1           /* stratified primary samples through the 256-bin inverse CDF of x+y+z with a
                                    5% floor over [hero_min, hero_max] (host_color_utilities.cuh:188-233,
                                    device_sensor_utilities.cuh:25-33), weight 1 / (n_hero pdf)
                                    (kernel_generate_thin_lens_ray.cu:73-74) */
} mgk_lambda_sampling_t;

/* Spectral accumulation (color_settings.h:11-14, optical_system_settings.h:22-27). */
typedef struct mgk_spectral_t {
    uint32_t n_hero;             /* hero wavelengths per path (1..4), stratified */
    float hero_min_nm, hero_max_nm;
    uint32_t n_bin;              /* intrinsic spectral irradiance bins (<= 64) */
    float bin_min_nm, bin_max_nm;
    float sigma_nm;              /* Gaussian redistribution width */
    float iso, fps, shutter_angle;
    float white_point_K;         /* scene white balance (color_settings.h:27) */
    uint32_t lambda_sampling;    /* mgk_lambda_sampling_t */
} mgk_spectral_t;

/* One terminated path (16 bytes).  DISK: a0 = r_hit, a1 = azimuth;
 * ESCAPE: a0 = cos(theta) of the escape direction about +y, a1 = atan2(z, x);
 * SPHERE: a0, a1 = the same angles of the surface normal at the hit.
 * g = (u.p)_camera / (u.p)_emitter (relativity_utilities.cuh:62-67). */
typedef struct mgk_hit_t {
    uint32_t cls;
    float g;
    float a0;
    float a1;
} mgk_hit_t;

typedef struct mgk_trace_stats_t {
    uint64_t n_rays;
    uint64_t n_steps;            /* accepted (or exhausted) steps */
    uint64_t n_trials;           /* RKF45 trials */
    uint64_t n_trials_fp64;      /* trials run entirely in FP64 */
    double ms;                   /* device time of the integration kernel(s) */
    uint64_t n_reflections;      /* reflections off diffuse box matter (paths re-traced) */
} mgk_trace_stats_t;

typedef struct mgk_context* mgk_context_handle;

/* Library state per device: counters, star index, cached work buffers. */
MGK_API e_magik_result_types mgk_context_create(int device, mgk_context_handle* out);
MGK_API e_magik_result_types mgk_context_destroy(mgk_context_handle ctx);

/* Uploads the scene's star catalogue (generated from star_seed) and its
 * sky-grid index.  Must precede tracing or shading of a scene with stars. */
MGK_API e_magik_result_types mgk_set_scene(mgk_context_handle ctx, const mgk_scene_t* scene);

/* Builds the two-level bounding volume hierarchy of the meshes (one bottom
 * level per mesh, a top level over the instances' world bounds) and uploads
 * it; scenes with MGK_SCENE_MESHES trace against it until the next call.
 * n_mesh = 0 clears it.  Has no counterpart in the reference, whose top- and
 * bottom-level intersection stages are placeholders (wavefront.cpp:94-100). */
MGK_API e_magik_result_types mgk_set_meshes(mgk_context_handle ctx, const mgk_mesh_t* meshes, uint32_t n_mesh,
                                            const mgk_instance_t* instances, uint32_t n_instance);

/* The hierarchy mgk_set_meshes would build, built and checked on the host
 * only (no device is touched): the same input checks and result codes, then
 * a walk of every level that verifies the skip links move forward, children
 * lie inside their parents and every triangle sits in exactly one leaf.
 * Returns MGK_ERROR_INVALID_ARGUMENT if a check fails.  No counterpart in the
 * reference; for callers that validate geometry before uploading it. */
typedef struct mgk_mesh_hierarchy_info_t {
    uint64_t n_nodes;            /* all levels */
    uint64_t n_triangles;        /* triangle slots over all meshes */
    uint32_t max_depth;          /* deepest bottom level */
    uint32_t max_leaf;           /* most triangles in a leaf */
    float world_min[3], world_max[3];
} mgk_mesh_hierarchy_info_t;
MGK_API e_magik_result_types mgk_mesh_hierarchy_info(const mgk_mesh_t* meshes, uint32_t n_mesh, const mgk_instance_t* instances,
                                                     uint32_t n_instance, mgk_mesh_hierarchy_info_t* info);

/* Ray generation + initial conditions (kernel_generate_thin_lens_ray.cu:51-77,
 * kernel_generate_initial_conditions.cu:54-74) for every pixel and the samples
 * [sample_begin, sample_begin + spp).  Ray i = s * W * H + py * W + px is
 * written as SoA: d_rays[c * ray_stride + i] for c = 0..9 =
 * x, y, z, t, p_x, p_y, p_z, p_t, h0, u.p.  Element type is float for
 * MGK_PRECISION_FP32_REFERENCE and double for MIXED and FP64. */
MGK_API e_magik_result_types mgk_generate_rays(mgk_context_handle ctx, const mgk_integrator_t* integ, const mgk_camera_t* cam,
                                               uint64_t seed, uint32_t sample_begin, uint32_t spp, void* d_rays,
                                               uint64_t ray_stride, void* stream);

/* Persistent geodesic integration (replaces kernel_advance_geodesic,
 * kernel_intersect, kernel_enforce_ray_limits and their queues).  Rays are
 * pulled from [0, n_rays), or from d_queue[0..n_rays) when d_queue is not
 * NULL (entries below ray_stride and below 0xFFFFFFFF); hit i (or d_queue[i]'s
 * slot) receives the result.  d_steps may be NULL. */
MGK_API e_magik_result_types mgk_trace_rays(mgk_context_handle ctx, const mgk_integrator_t* integ, const mgk_scene_t* scene,
                                            const float camera_position[3], const void* d_rays, uint64_t ray_stride,
                                            uint64_t n_rays, const uint32_t* d_queue, mgk_hit_t* d_hits,
                                            uint32_t* d_steps, void* stream);

/* mgk_trace_rays that allocates nothing: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * checks and results as mgk_trace_rays; its work counters, statistics and
 * (mixed precision) zone hand-off queues and parked states live in a fixed
 * ~20 MB device array of the library, one per device, resolved when the
 * context is created.  A mixed-precision call takes at most 16384 rays
 * (MGK_ERROR_CAPACITY_EXCEEDED beyond); the FP32-reference and FP64
 * precisions take any count mgk_trace_rays takes.  Reflections are not
 * followed, and mgk_last_trace_stats keeps describing the last other trace.
 * Calls that overlap on one device share that array and race: serialise them
 * (one stream, or events). */
MGK_API e_magik_result_types mgk_trace_rays_noalloc(mgk_context_handle ctx, const mgk_integrator_t* integ, const mgk_scene_t* scene,
                                                    const float camera_position[3], const void* d_rays, uint64_t ray_stride,
                                                    uint64_t n_rays, const uint32_t* d_queue, mgk_hit_t* d_hits,
                                                    uint32_t* d_steps, void* stream);

/* mgk_trace_rays that also writes each ray's state at termination (the end
 * of its last step; for box matter, the chord point of the hit with the
 * step's end momentum): d_final[c * ray_stride + i] for c = 0..7 = x, y,
 * z, t, p_x, p_y, p_z, p_t (doubles, dimensionless).  For endpoint-direction
 * and invariant-drift measurements against the FP64 path. */
MGK_API e_magik_result_types mgk_trace_rays_final(mgk_context_handle ctx, const mgk_integrator_t* integ, const mgk_scene_t* scene,
                                                  const float camera_position[3], const void* d_rays, uint64_t ray_stride,
                                                  uint64_t n_rays, const uint32_t* d_queue, mgk_hit_t* d_hits, double* d_final,
                                                  void* stream);

/* Whole paths of a camera's rays (W x H x spp, ray i = sample i / (W H) of
 * pixel i % (W H), as mgk_generate_rays lays them out): trace, and at diffuse
 * box matter reflect and trace again, up to scene->max_reflections times
 * (kernel_bxdf.cu:69-142 with bxdf_utilities.cuh:56-151, matter_examples.cuh:36-60,
 * brdf.cuh:12-29; the limit of kernel_enforce_ray_limits.cu:48-50).  The
 * step count runs on across reflections.  Reflection k draws its uniforms from
 * Philox block 1 + k of (seed, camera_id, pixel, sample_begin + sample).
 * d_rays (FP32 for the FP32 reference precision, else FP64) is overwritten in
 * place with the reflected rays' initial conditions.  d_hits receives each
 * path's last hit; d_steps (optional) its total steps; d_final (optional,
 * 8 x ray_stride doubles) its terminal state as mgk_trace_rays_final writes it. */
MGK_API e_magik_result_types mgk_trace_paths(mgk_context_handle ctx, const mgk_integrator_t* integ, const mgk_scene_t* scene,
                                             const mgk_camera_t* camera, uint64_t seed, uint32_t sample_begin, uint32_t spp,
                                             void* d_rays, uint64_t ray_stride, mgk_hit_t* d_hits, uint32_t* d_steps,
                                             double* d_final, void* stream);

/* Batched cameras in one persistent launch (several scenes' viewpoints with
 * their own spin, resolution and pose).  mgk_generate_rays_batch writes the
 * rays of camera k to [offsets[k], offsets[k + 1]) (offsets: n_cam + 1
 * host entries, filled in), each camera's block laid out as in
 * mgk_generate_rays, as 11 doubles per ray: the 10 components above plus the
 * camera's spin of the equations.  integ supplies everything but the spin
 * (precision MIXED or FP64).  mgk_trace_rays_batch integrates such rays with
 * each ray's own spin, the escape sphere centred on each ray's origin and the
 * disk from each ray's ISCO to scene->disk_r_out. */
typedef struct mgk_batch_camera_t {
    mgk_camera_t camera;
    float spin;                  /* dimensionless a as given by the user */
    uint32_t reserved;
} mgk_batch_camera_t;

MGK_API e_magik_result_types mgk_generate_rays_batch(mgk_context_handle ctx, const mgk_integrator_t* integ,
                                                     const mgk_batch_camera_t* cameras, uint32_t n_cameras, uint64_t seed,
                                                     uint32_t spp, double* d_rays, uint64_t ray_stride, uint64_t* offsets,
                                                     void* stream);
MGK_API e_magik_result_types mgk_trace_rays_batch(mgk_context_handle ctx, const mgk_integrator_t* integ, const mgk_scene_t* scene,
                                                  const double* d_rays, uint64_t ray_stride, uint64_t n_rays, const uint32_t* d_queue,
                                                  mgk_hit_t* d_hits, uint32_t* d_steps, void* stream);

/* Statistics of the last mgk_trace_rays / mgk_render_frame (synchronises). */
MGK_API e_magik_result_types mgk_last_trace_stats(mgk_context_handle ctx, mgk_trace_stats_t* out);

/* Spectral shading + accumulation (kernel_bxdf.cu emit path,
 * kernel_redistribute_radiance.cu:55-92, kernel_splat_to_renderbuffer.cu:37-57
 * up to the linear colour): per pixel, samples [0, spp) of d_hits (sample-major,
 * W*H per sample) are shaded at their hero wavelengths, redistributed into
 * n_bin Gaussian bins and folded with the bin CMF weights into CIE XYZ / spp.
 * d_xyz receives 3 floats per pixel. */
MGK_API e_magik_result_types mgk_shade(mgk_context_handle ctx, const mgk_scene_t* scene, const mgk_spectral_t* spec,
                                       const mgk_camera_t* cam, uint64_t seed, uint32_t sample_begin, uint32_t spp,
                                       const mgk_hit_t* d_hits, float* d_xyz, void* stream);

/* Whole frame: generate + trace + shade with library-owned ray and hit
 * buffers (grown on demand). */
MGK_API e_magik_result_types mgk_render_frame(mgk_context_handle ctx, const mgk_integrator_t* integ, const mgk_scene_t* scene,
                                              const mgk_spectral_t* spec, const mgk_camera_t* cam, uint64_t seed,
                                              uint32_t spp, float* d_xyz, void* stream);

/* Host helpers of the colour pipeline (host_color_utilities.cuh:145-178,
 * :29-68; color_pipeline.cuh:21-48): bin CMF weights (n_bin x 3, XYZ) and the
 * white-balanced XYZ -> linear Rec.709 matrix (row-major). */
MGK_API e_magik_result_types mgk_color_weights(const mgk_spectral_t* spec, float* weights_xyz, float* xyz_to_scene);
/* The 256-entry inverse CDF (nm) that MGK_LAMBDA_CMF samples over [hero_min, hero_max]
 * (host_color_utilities.cuh:188-233 with the 5% floor of color_pipeline.cuh:36); host only. */
MGK_API e_magik_result_types mgk_cmf_inverse_cdf(const mgk_spectral_t* spec, float* inverse_cdf_nm);

/* Display settings: the colour part of the reference's stage
 * (color_settings.h:9-30, with its defaults in the comments). */
typedef struct mgk_display_t {
    uint32_t scene_colorspace;          /* e_magik_scene_colorspace_types (LINEAR_REC709) */
    uint32_t view_transform;            /* e_magik_view_transform_types (GRAN_TURISMO_7) */
    uint32_t display_colorspace;        /* e_magik_display_colorspace_types (REC709) */
    uint32_t output_format;             /* e_magik_output_format_types (FLOAT32); values stay FP32 */
    float white_point_K;                /* white balance: blackbody (5600) */
    float tint;                         /* ... shifted along the locus normal (0) */
    float reinhard_white_luminance;     /* (4) */
    float khronos_start_compression;    /* (0.8) */
    float khronos_desaturation;         /* (0.15) */
    uint32_t gt7_mode;                  /* e_magik_gran_turismo_7_mode_types (JZAZBZ) */
    float gt7_sdr_paper_white;          /* cd/m^2 (250) */
    float gt7_reference_luminance;      /* cd/m^2 per framebuffer unit (100) */
    float gt7_jzazbz_exponent_scale;    /* (1.7) */
    uint32_t agx_look;                  /* e_magik_agx_look_types (GOLDEN) */
} mgk_display_t;

/* The display side of the splat (kernel_splat_to_renderbuffer.cu:37-57,
 * device_color_utilities.cuh:56-72 with color_pipeline.cuh:21-48): per pixel,
 * linear XYZ (as mgk_shade writes it) -> white-balanced scene colour space ->
 * view transform -> display primaries (Bradford adaptation between the white
 * points) -> display transfer function.  d_xyz and d_out hold 3 floats per
 * pixel and may be the same buffer. */
MGK_API e_magik_result_types mgk_display_transform(mgk_context_handle ctx, const mgk_display_t* display, const float* d_xyz,
                                                   uint64_t n_pixels, float* d_out, void* stream);

/* Places a camera at Boyer-Lindquist (r, theta) of a Kerr black hole with spin
 * a (as given by the user), in Kerr-Schild Cartesian coordinates with the spin
 * along +y, looking at the origin with +y up; horizontal field of view fov_deg
 * with the given sensor width.  mass_kg sets the length scale. */
MGK_API e_magik_result_types mgk_camera_boyer_lindquist(float r_obs, float theta_deg, float spin, float mass_kg,
                                                        float fov_deg, uint32_t x_res, uint32_t y_res, mgk_camera_t* out);


/* ---------------------------------------------------------------------------
 * 3. Learned surrogate (no counterpart in the reference: the MLP of the
 *    tracer's Kerr path).  5 features (spin a, observer r and theta in
 *    Boyer-Lindquist, screen alpha, beta) -> 4 x 128 ReLU -> outputs
 *    (horizon / disk logits, normalised log r of the disk hit, escape
 *    direction cos(theta), sin(phi), cos(phi)); bf16 weights and activations,
 *    FP32 accumulation on the matrix cores.
 * ------------------------------------------------------------------------- */
typedef struct mgk_surrogate* mgk_surrogate_handle;

typedef struct mgk_surrogate_ranges_t {
    float spin_min, spin_max;             /* physical spin a (before time reversal) */
    float r_min, r_max;                   /* observer BL radius, units of M */
    float theta_min_deg, theta_max_deg;   /* observer inclination */
    float screen_max;                     /* |alpha|, |beta| <= screen_max (units of M) */
    float disk_r_max;                     /* largest disk radius the r output must span */
} mgk_surrogate_ranges_t;

typedef struct mgk_train_config_t {
    uint32_t batch;
    uint32_t epochs;
    float lr, lr_final;                   /* Adam, cosine decay */
    float w_r, w_d;                       /* loss weights: disk radius, escape direction */
} mgk_train_config_t;

/* Fallback rules of the surrogate render path: a path is traced by RKF45
 * instead of being answered by the network when any rule selects it. */
typedef struct mgk_fallback_t {
    float min_confidence;                 /* max class probability below this (0.9) */
    float critical_margin;                /* within this screen distance (M) of the critical curve (0.05) */
    float direction_margin;               /* escape: predicted direction within this angle (rad) of a star's edge */
    float disk_edge_margin;               /* disk: predicted r within this fraction of r_in or r_out */
    uint32_t judge_stride;                /* every judge_stride-th path is also traced by RKF45 (0 = none) */
    uint32_t trace_disk;                  /* 1: every disk-class path is traced */
    /* 1: every answered path is checked, and a disk path's radius refined, by
     * the Mino-time integrals of its separated radial and polar motion
     * (FP64, seeded by the network's radius); a path whose class they do not
     * confirm, or cannot resolve (near the critical curve), is traced.  The
     * disk_edge_margin rule is then not needed and not applied. */
    uint32_t analytic;
} mgk_fallback_t;

typedef struct mgk_surrogate_stats_t {
    uint64_t n_paths;
    uint64_t n_fallback;                  /* paths traced by RKF45 (judge paths included) */
    uint64_t n_fallback_confidence;       /* ... selected by the first rule that fired */
    uint64_t n_fallback_critical;
    uint64_t n_fallback_disk;
    uint64_t n_fallback_star;
    uint64_t n_fallback_domain;           /* screen point outside the trained range */
    uint64_t n_judge;                     /* judge paths the network answered */
    uint64_t judge_class_agree;           /* ... whose class matched RKF45 */
    uint64_t judge_n_disk, judge_n_escape;
    double judge_mean_rel_r;              /* mean |r_s - r| / r over judge paths that are disk in both */
    double judge_mean_rel_g;              /* mean |g_s - g| / g over the same */
    double judge_mean_dir, judge_max_dir; /* escape direction error (rad) over judge paths that escape in both */
    uint64_t n_trials;                    /* RKF45 trials of the fallback trace */
    double ms_raygen, ms_mlp, ms_trace, ms_shade, ms_total;   /* device time of the stages */
    uint64_t n_fallback_analytic;         /* fallback.analytic: class not confirmed by, or unresolved in, the Mino-time integrals */
} mgk_surrogate_stats_t;

MGK_API e_magik_result_types mgk_surrogate_create(int device, uint64_t seed, const mgk_surrogate_ranges_t* ranges,
                                                  mgk_surrogate_handle* out);
MGK_API e_magik_result_types mgk_surrogate_destroy(mgk_surrogate_handle s);
/* Parameter counts of the padded layout (weights 16x128, 3 x 128x128, 128x32). */
MGK_API e_magik_result_types mgk_surrogate_n_params(uint64_t* n_weights, uint64_t* n_biases);
MGK_API e_magik_result_types mgk_surrogate_get_weights(mgk_surrogate_handle s, float* weights, float* biases);
MGK_API e_magik_result_types mgk_surrogate_set_weights(mgk_surrogate_handle s, const float* weights, const float* biases);
/* Inference of n feature rows (n x 16 normalised features) -> n x 8 outputs. */
MGK_API e_magik_result_types mgk_surrogate_forward(mgk_surrogate_handle s, const float* d_x, uint64_t n, float* d_y, void* stream);
/* Forward (training kernels) + backward from a given dL/dy (n x 32): the
 * gradients land in host arrays (weights, biases layout); d_y (optional, n x 32)
 * receives the forward outputs.  For gradient checks. */
MGK_API e_magik_result_types mgk_surrogate_gradients(mgk_surrogate_handle s, const float* d_x, const float* d_dy, uint32_t n, float* d_y,
                                                     float* weight_grad, float* bias_grad);
/* One Adam step on a batch (targets n x 8); losses[4] = mean CE, disk, dir, accuracy. */
MGK_API e_magik_result_types mgk_surrogate_train_step(mgk_surrogate_handle s, const float* d_x, const float* d_t, uint32_t n, float lr,
                                                      float w_r, float w_d, float* losses);
MGK_API e_magik_result_types mgk_surrogate_train(mgk_surrogate_handle s, const float* d_x, const float* d_t, uint64_t n,
                                                 const mgk_train_config_t* cfg, float* epoch_losses);
/* Training data: n seeded rays uniform over the ranges, traced by the FP64
 * RKF45 path (integ must be Kerr / FP64); d_rays n x 11 doubles (SoA),
 * features n x 16, targets n x 8, hits n. */
MGK_API e_magik_result_types mgk_surrogate_generate_data(mgk_context_handle ctx, mgk_surrogate_handle s, const mgk_integrator_t* integ,
                                                         const mgk_scene_t* scene, uint64_t seed, uint64_t n, double* d_rays,
                                                         float* d_x, float* d_t, mgk_hit_t* d_hits, void* stream);
/* Whole frame with the surrogate: MLP inference for every path, RKF45 (the
 * integ settings) for the paths the fallback rules select and for the judge
 * sample, then the same spectral shading.  The camera must be a
 * Boyer-Lindquist camera (mgk_camera_boyer_lindquist) inside the trained
 * ranges, the scene's disk the trained one (r_ISCO .. disk_r_max).  d_hits
 * (optional, W*H*spp records, sample-major) receives every path's hit. */
MGK_API e_magik_result_types mgk_render_frame_surrogate(mgk_context_handle ctx, mgk_surrogate_handle s, const mgk_integrator_t* integ,
                                                        const mgk_scene_t* scene, const mgk_spectral_t* spec, const mgk_camera_t* cam,
                                                        uint64_t seed, uint32_t spp, const mgk_fallback_t* fb, float* d_xyz,
                                                        mgk_hit_t* d_hits, mgk_surrogate_stats_t* stats, void* stream);
/* The analytic check of fallback.analytic on the host: the disk crossing,
 * capture or escape of n Kerr rays from the Mino-time integrals of their
 * separated radial and polar motion, in FP64.  rays: HOST memory in the SoA
 * layout of mgk_generate_rays (FP64, component c of ray i at
 * rays[c * ray_stride + i]); integ gives the spin and time_reversed, scene the
 * disk's r_in .. r_out.  kind[i]: 0 unresolved (near the critical curve, or
 * not starting inward), 1 captured, 2 escapes, 3 disk.  With r_seed (an
 * estimate of the crossing radius, > 0) r_hit[i] is the disk radius, solved
 * for from that seed; without it only the class is found (r_hit[i] = 0).  No
 * counterpart in the reference. */
MGK_API e_magik_result_types mgk_mino_disk(const mgk_integrator_t* integ, const mgk_scene_t* scene, const double* rays,
                                           uint64_t ray_stride, uint64_t n, const double* r_seed, int32_t* kind, double* r_hit);

#ifdef __cplusplus
}
#endif

#endif /* MAGIK_AMD_H */
