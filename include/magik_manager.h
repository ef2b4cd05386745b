/* The render manager and its command queue system: the reference's public
 * manager API (magik.h:215-313 manager, :405-428 statistics, :430-575
 * framebuffer collections, :676-1275 commands) with the same symbols, struct
 * layouts and enum values, so that a DCC written against magik.h drives the
 * HIP path unchanged.  The structs are declared without the reference's C++
 * default member initialisers (this header is C); the defaults are given in
 * the comments.  Layouts are checked against magik.h by the tests.
 *
 * Differences of behaviour from the reference (each a choice of this library):
 *  - MAGIK_COMMAND_SET_STATE_PAUSE stops sampling and MAGIK_COMMAND_SET_STATE_RENDER
 *    resumes it; MAGIK_COMMAND_CLEAR_RENDER_BUFFER restarts the accumulation
 *    (the reference accepts all three and does nothing, command.cpp:152-165).
 *  - Only host framebuffer collections: MAGIK_FRAMEBUFFER_COLLECTION_CONFIG_HOST.
 *    The OpenGL / Vulkan / CUDA configurations are not built (no interop on this
 *    platform); configuring one returns MAGIK_UNKNOWN_ENUM_TYPE.
 *  - The scene is the reference stage's diffuse box and LED panel
 *    (stage.h:19-22), which no command changes, traced with whole paths. */
#ifndef MAGIK_MANAGER_H
#define MAGIK_MANAGER_H

#include "magik_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- render manager (magik.h:240-313) ---- */
typedef enum e_magik_manager_display_types {
    MAGIK_DISPLAY_SWAPCHAIN = 0,
    MAGIK_DISPLAY_HEADLESS = 1
} e_magik_manager_display_types;

typedef struct magik_manager_descriptor_t {
    e_magik_manager_display_types display_type; /* MAGIK_DISPLAY_SWAPCHAIN */
    uint32_t user_device_id;                    /* 0 */
    void* user_stream;                          /* NULL: the manager creates its own (hipStream_t) */
    uint32_t cqs_n_reserved_chunk;              /* 4096 four-byte chunks */
    bool cqs_drop_overflow;                     /* false */
} magik_manager_descriptor_t;

typedef struct magik_render_manager* magik_render_manager_t;

MGK_API e_magik_result_types magik_render_manager_create(magik_render_manager_t* manager, magik_manager_descriptor_t descriptor);
MGK_API e_magik_result_types magik_render_manager_destroy(magik_render_manager_t manager);
MGK_API e_magik_result_types magik_render_manager_fetch_error(magik_render_manager_t manager);

/* ---- colour (magik.h:319-321; the other colour enums are in magik_amd.h) ---- */
typedef enum e_magik_color_matching_function_types { MAGIK_CMF_CIE_1931 = 0 } e_magik_color_matching_function_types;

/* ---- statistics (magik.h:397-428) ---- */
typedef struct magik_renderer_statistics_t {
    size_t x_sample;                                 /* samples accumulated in the render buffer */
    double total_render_time_seconds;
    double latest_render_loop_iteration_time_seconds;
    double total_wavefront_sample_time_seconds;
    double latest_wavefront_sample_time_seconds;
    /* Queue counts per wavefront iteration of the latest sample (iteration k
     * takes every live path to step k + 1): advance_A = intersect = paths
     * with more than k steps, advance_B = those going on to iteration k + 1,
     * bxdf = paths meeting matter at step k + 1 (ending or reflecting),
     * redistribute = splat = paths ending there.  Paths stopped by the step
     * limit leave advance_A only.  The six arrays are one host allocation
     * made by a fresh fetch and freed by magik_destroy_renderer_statistics. */
    size_t latest_wavefront_history_length;
    uint32_t* h_latest_advance_A_queue_count_history;
    uint32_t* h_latest_advance_B_queue_count_history;
    uint32_t* h_latest_intersect_queue_count_history;
    uint32_t* h_latest_bxdf_queue_count_history;
    uint32_t* h_latest_redistribute_queue_count_history;
    uint32_t* h_latest_splat_queue_count_history;
    size_t total_framebuffer_swapchain_size;
    size_t total_wavefront_buffer_size;
} magik_renderer_statistics_t;

MGK_API e_magik_result_types magik_fetch_renderer_statistics(magik_render_manager_t manager, bool* is_new_fetch,
                                                               magik_renderer_statistics_t* dcc_statistics);
MGK_API e_magik_result_types magik_destroy_renderer_statistics(magik_renderer_statistics_t* statistics);

/* ---- framebuffer collections (magik.h:430-575) ---- */
typedef enum e_magik_framebuffer_collection_config_types {
    MAGIK_FRAMEBUFFER_COLLECTION_CONFIG_HOST = 0,
    MAGIK_FRAMEBUFFER_COLLECTION_CONFIG_CUDA = 1,
    MAGIK_FRAMEBUFFER_COLLECTION_CONFIG_OPENGL_INTEROP = 2,
    MAGIK_FRAMEBUFFER_COLLECTION_CONFIG_VULKAN_INTEROP = 3
} e_magik_framebuffer_collection_config_types;

typedef struct magik_framebuffer_collection_external* magik_framebuffer_collection_external_t;

typedef struct magik_framebuffer_container_config_host_t {
    uint32_t x_resolution;
    uint32_t y_resolution;
    uint32_t channels;
    size_t size_of_data;   /* bytes */
    float* h_data;         /* owned by the collection; valid until its next fetch or destroy */
} magik_framebuffer_container_config_host_t;

MGK_API e_magik_result_types magik_configure_framebuffer_collection(magik_framebuffer_collection_external_t* framebuffer,
                                                                      e_magik_framebuffer_collection_config_types config_type);
MGK_API e_magik_result_types magik_framebuffer_collection_fetch(magik_render_manager_t manager, bool* is_new_fetch,
                                                                  magik_framebuffer_collection_external_t framebuffer);
MGK_API e_magik_result_types magik_extract_config_host(magik_render_manager_t manager, magik_framebuffer_container_config_host_t* container,
                                                         magik_framebuffer_collection_external_t framebuffer, const char* name);
MGK_API e_magik_result_types magik_destroy_framebuffer_collection(magik_framebuffer_collection_external_t framebuffer);

/* ---- command queue system (magik.h:676-1275) ---- */
typedef enum e_magik_cqs_command_types {
    MAGIK_COMMAND_SET_STATE_RENDER = 1000,
    MAGIK_COMMAND_SET_STATE_PAUSE = 1001,
    MAGIK_COMMAND_CLEAR_RENDER_BUFFER = 1002,
    MAGIK_COMMAND_ADD_AOV = 1003,
    MAGIK_COMMAND_REMOVE_AOV = 1004,
    MAGIK_COMMAND_PRINTF = 2000,
    MAGIK_COMMAND_SET_METRIC_MINKOWSKI = 4000,
    MAGIK_COMMAND_SET_METRIC_SCHWARZSCHILD = 4001,
    MAGIK_COMMAND_SET_METRIC_KERR = 4002,
    MAGIK_COMMAND_SET_ABSOLUTE_ERROR_TOLERANCE = 4100,
    MAGIK_COMMAND_SET_RELATIVE_ERROR_TOLERANCE = 4101,
    MAGIK_COMMAND_SET_MAX_SUBSTEPS = 4102,
    MAGIK_COMMAND_SET_OBSERVER_LOCAL_PROPER_VELOCITY = 4200,
    MAGIK_COMMAND_SET_CELESTIAL_SPHERE_RADIUS = 4300,
    MAGIK_COMMAND_SET_SCHWARZSCHILD_EVENT_HORIZON_SCALAR = 4400,
    MAGIK_COMMAND_SET_SCHWARZSCHILD_MASS_PARAMETER = 4401,
    MAGIK_COMMAND_SET_KERR_EVENT_HORIZON_SCALAR = 4402,
    MAGIK_COMMAND_SET_KERR_MASS_PARAMETER = 4403,
    MAGIK_COMMAND_SET_KERR_DIMENSIONLESS_SPIN_PARAMETER = 4404,
    MAGIK_COMMAND_SET_ACTIVE_CAMERA = 5000,
    MAGIK_COMMAND_SET_CAMERA_POSITION = 5001,
    MAGIK_COMMAND_SET_CAMERA_ORIENTATION = 5002,
    MAGIK_COMMAND_SET_CAMERA_RESOLUTION = 5003,
    MAGIK_COMMAND_SET_N_INTRINSIC_SPECTRAL_IRRADIANCE_BIN = 6000,
    MAGIK_COMMAND_SET_SAMPLE_DOMAIN_MAXIMUM_WAVELENGTH_NM = 6001,
    MAGIK_COMMAND_SET_SAMPLE_DOMAIN_MINIMUM_WAVELENGTH_NM = 6002,
    MAGIK_COMMAND_SET_REDISTRIBUTION_STANDARD_DEVIATION_NM = 6003,
    MAGIK_COMMAND_SET_COLOR_MATCHING_FUNCTION = 6100,
    MAGIK_COMMAND_SET_SCENE_COLORSPACE = 6101,
    MAGIK_COMMAND_SET_VIEW_TRANSFORM = 6102,
    MAGIK_COMMAND_SET_DISPLAY_COLORSPACE = 6103,
    MAGIK_COMMAND_SET_OUTPUT_FORMAT = 6104,
    MAGIK_COMMAND_SET_WHITE_BALANCE_WHITE_POINT = 6200,
    MAGIK_COMMAND_SET_WHITE_BALANCE_TINT = 6201,
    MAGIK_COMMAND_SET_VIEW_TRANSFORM_REINHARD_EXTENDED_WHITE_LUMINANCE = 6300,
    MAGIK_COMMAND_SET_VIEW_TRANSFORM_KHRONOS_PBR_NEUTRAL_START_COMPRESSION = 6301,
    MAGIK_COMMAND_SET_VIEW_TRANSFORM_KHRONOS_PBR_NEUTRAL_DESATURATION = 6302,
    MAGIK_COMMAND_SET_VIEW_TRANSFORM_GT7_MODE = 6304,
    MAGIK_COMMAND_SET_VIEW_TRANSFORM_GT7_SDR_PAPER_WHITE_BALANCE = 6305,
    MAGIK_COMMAND_SET_VIEW_TRANSFORM_GT7_REFERENCE_LUMINANCE = 6306,
    MAGIK_COMMAND_SET_VIEW_TRANSFORM_GT7_JZAZBZ_EXPONENT_SCALE_FACTOR = 6407,
    MAGIK_COMMAND_SET_VIEW_TRANSFORM_AGX_LOOK = 6308,
    MAGIK_COMMAND_SET_N_GEODESIC_STEP_LIMIT = 7000,
    MAGIK_COMMAND_SET_N_REFLECTED_LIMIT = 7001,
    MAGIK_COMMAND_PROHIBITED_FORCE_SIZE = 0x7FFFFFFF
} e_magik_cqs_command_types;

/* Every command starts with its own type (embedded_type); a struct of four-byte
 * members, pushed by address (4-byte aligned). */
#define MGK_COMMAND_0(name) typedef struct name { e_magik_cqs_command_types embedded_type; } name;
#define MGK_COMMAND_1(name, t1, f1) typedef struct name { e_magik_cqs_command_types embedded_type; t1 f1; } name;
MGK_COMMAND_0(magik_command_set_state_render_t)
MGK_COMMAND_0(magik_command_set_state_pause_t)
MGK_COMMAND_0(magik_command_clear_render_buffer_t)
typedef struct magik_command_add_aov_t {
    e_magik_cqs_command_types embedded_type;
    uint32_t length_of_name;
    char name[512];
    uint32_t channels;
} magik_command_add_aov_t;
typedef struct magik_command_remove_aov_t {
    e_magik_cqs_command_types embedded_type;
    uint32_t length_of_name;
    char name[512];
} magik_command_remove_aov_t;
typedef struct magik_command_printf_t {
    e_magik_cqs_command_types embedded_type;
    uint32_t length_of_text;
    char text[512];
} magik_command_printf_t;
MGK_COMMAND_0(magik_command_set_metric_minkowski_t)
MGK_COMMAND_0(magik_command_set_metric_schwarzschild_t)
MGK_COMMAND_0(magik_command_set_metric_kerr_t)
MGK_COMMAND_1(magik_command_set_absolute_error_tolerance_t, float, absolute_error_tolerance)            /* 1e-3 */
MGK_COMMAND_1(magik_command_set_relative_error_tolerance_t, float, relative_error_tolerance)            /* 1e-10 */
MGK_COMMAND_1(magik_command_set_max_substeps_t, uint32_t, max_substeps)                                 /* 56 */
typedef struct magik_command_set_observer_local_proper_velocity_t {
    e_magik_cqs_command_types embedded_type;
    float velocity_x, velocity_y, velocity_z;
} magik_command_set_observer_local_proper_velocity_t;
MGK_COMMAND_1(magik_command_set_celestial_sphere_radius_t, float, celestial_sphere_radius)              /* 1000 m */
MGK_COMMAND_1(magik_command_set_schwarzschild_event_horizon_scalar_t, float, scalar)                    /* 1.01 */
MGK_COMMAND_1(magik_command_set_schwarzschild_mass_parameter_t, float, mass_parameter)                  /* 6.7329546075e26 kg */
MGK_COMMAND_1(magik_command_set_kerr_event_horizon_scalar_t, float, scalar)                             /* 1.01 */
MGK_COMMAND_1(magik_command_set_kerr_mass_parameter_t, float, mass_parameter)                           /* 6.7329546075e26 kg */
MGK_COMMAND_1(magik_command_set_kerr_dimensionless_spin_parameter_t, float, dimensionless_spin_parameter) /* 0.999 */
MGK_COMMAND_1(magik_command_set_active_camera_t, uint32_t, id)
typedef struct magik_command_set_camera_position_t {
    e_magik_cqs_command_types embedded_type;
    uint32_t id;
    float position_x, position_y, position_z;   /* SI metres */
} magik_command_set_camera_position_t;
typedef struct magik_command_set_camera_orientation_t {
    e_magik_cqs_command_types embedded_type;
    uint32_t id;
    float rotation_x, rotation_y, rotation_z;   /* radians, applied as rotate(z, y, x) (command.cpp:386) */
} magik_command_set_camera_orientation_t;
typedef struct magik_command_set_camera_resolution_t {
    e_magik_cqs_command_types embedded_type;
    uint32_t x_resolution, y_resolution;        /* 2195 x 1860 */
} magik_command_set_camera_resolution_t;
MGK_COMMAND_1(magik_command_set_n_intrinsic_spectral_irradiance_bin_t, uint32_t, n_isi_bin)             /* 24 */
MGK_COMMAND_1(magik_command_set_sample_domain_maximum_wavelength_nm_t, float, maximum_wavelength_nm)     /* 830 */
MGK_COMMAND_1(magik_command_set_sample_domain_minimum_wavelength_nm_t, float, minimum_wavelength_nm)     /* 360 */
MGK_COMMAND_1(magik_command_set_redistribution_standard_deviation_nm_t, float, standard_deviation_nm)    /* 10 */
MGK_COMMAND_1(magik_command_set_color_matching_function_t, e_magik_color_matching_function_types, cmf_type)
MGK_COMMAND_1(magik_command_set_scene_colorspace_t, e_magik_scene_colorspace_types, scene_colorspace)
MGK_COMMAND_1(magik_command_set_view_transform_t, e_magik_view_transform_types, view_transform)
MGK_COMMAND_1(magik_command_set_display_colorspace_t, e_magik_display_colorspace_types, display_colorspace)
MGK_COMMAND_1(magik_command_set_output_format_t, e_magik_output_format_types, output_format)
MGK_COMMAND_1(magik_command_set_white_balance_white_point_t, float, white_point_kelvin)                 /* 5600 */
MGK_COMMAND_1(magik_command_set_white_balance_tint_t, float, tint)                                      /* 0 */
MGK_COMMAND_1(magik_command_set_view_transform_reinhard_extended_white_luminance_t, float, white_luminance)
MGK_COMMAND_1(magik_command_set_view_transform_khronos_pbr_neutral_start_compression_t, float, start_compression)
MGK_COMMAND_1(magik_command_set_view_transform_khronos_pbr_neutral_desaturation_t, float, desaturation)
MGK_COMMAND_1(magik_command_set_view_transform_gt7_mode_t, e_magik_gran_turismo_7_mode_types, mode)
MGK_COMMAND_1(magik_command_set_view_transform_gt7_sdr_paper_white_balance_t, float, sdr_paper_white_balance)
MGK_COMMAND_1(magik_command_set_view_transform_gt7_reference_luminance_t, float, reference_luminance)
MGK_COMMAND_1(magik_command_set_view_transform_gt7_jzazbz_exponent_scale_factor_t, float, jzazbz_exponent_scale_factor)
MGK_COMMAND_1(magik_command_set_view_transform_agx_look_t, e_magik_agx_look_types, look)
MGK_COMMAND_1(magik_command_set_n_geodesic_step_limit_t, uint32_t, n_geodesic_step_limit)               /* 512 */
MGK_COMMAND_1(magik_command_set_n_reflected_limit_t, uint32_t, n_reflected_limit)                       /* 10 */
#undef MGK_COMMAND_0
#undef MGK_COMMAND_1

/* Copies the command into the DCC-side buffer (command.cpp:683-716). */
MGK_API e_magik_result_types magik_cqs_push_command(magik_render_manager_t manager, const void* command);
/* Hands the buffer to the worker if it has consumed the previous one
 * (command.cpp:718-737); *is_dispatched tells which. */
MGK_API e_magik_result_types magik_cqs_dispatch_command_buffer(magik_render_manager_t manager, bool* is_dispatched);

#ifdef __cplusplus
}
#endif

#endif
