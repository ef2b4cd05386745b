// The render manager and its command queue system (include/magik_manager.h):
// a worker thread per manager that applies the dispatched commands to its
// stage and renders one sample per iteration into the "test" AOV, and a
// three-buffer swap (back / ready / front) through which the DCC fetches the
// latest finished frame into host memory.  Restates the behaviour of
// renderer.cpp:23-144, command/command.cpp:13-790 and
// framebuffer/framebuffer.cpp:6-440 over the library's render path: a sample
// is mgk_render_frame's sample n (Philox keyed by pixel and sample), so the
// progressive frame after n iterations equals the n-sample frame.
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "magik_manager.h"
#include "mgk_internal.hpp"

namespace {

using clk = std::chrono::steady_clock;

// Bytes of a command of each type (command.cpp:13-140); 0: not a command.
size_t command_size(uint32_t type) {
    switch (type) {
#define C(E, T) case E: return sizeof(T);
        C(MAGIK_COMMAND_SET_STATE_RENDER, magik_command_set_state_render_t)
        C(MAGIK_COMMAND_SET_STATE_PAUSE, magik_command_set_state_pause_t)
        C(MAGIK_COMMAND_CLEAR_RENDER_BUFFER, magik_command_clear_render_buffer_t)
        C(MAGIK_COMMAND_ADD_AOV, magik_command_add_aov_t)
        C(MAGIK_COMMAND_REMOVE_AOV, magik_command_remove_aov_t)
        C(MAGIK_COMMAND_PRINTF, magik_command_printf_t)
        C(MAGIK_COMMAND_SET_METRIC_MINKOWSKI, magik_command_set_metric_minkowski_t)
        C(MAGIK_COMMAND_SET_METRIC_SCHWARZSCHILD, magik_command_set_metric_schwarzschild_t)
        C(MAGIK_COMMAND_SET_METRIC_KERR, magik_command_set_metric_kerr_t)
        C(MAGIK_COMMAND_SET_ABSOLUTE_ERROR_TOLERANCE, magik_command_set_absolute_error_tolerance_t)
        C(MAGIK_COMMAND_SET_RELATIVE_ERROR_TOLERANCE, magik_command_set_relative_error_tolerance_t)
        C(MAGIK_COMMAND_SET_MAX_SUBSTEPS, magik_command_set_max_substeps_t)
        C(MAGIK_COMMAND_SET_OBSERVER_LOCAL_PROPER_VELOCITY, magik_command_set_observer_local_proper_velocity_t)
        C(MAGIK_COMMAND_SET_CELESTIAL_SPHERE_RADIUS, magik_command_set_celestial_sphere_radius_t)
        C(MAGIK_COMMAND_SET_SCHWARZSCHILD_EVENT_HORIZON_SCALAR, magik_command_set_schwarzschild_event_horizon_scalar_t)
        C(MAGIK_COMMAND_SET_SCHWARZSCHILD_MASS_PARAMETER, magik_command_set_schwarzschild_mass_parameter_t)
        C(MAGIK_COMMAND_SET_KERR_EVENT_HORIZON_SCALAR, magik_command_set_kerr_event_horizon_scalar_t)
        C(MAGIK_COMMAND_SET_KERR_MASS_PARAMETER, magik_command_set_kerr_mass_parameter_t)
        C(MAGIK_COMMAND_SET_KERR_DIMENSIONLESS_SPIN_PARAMETER, magik_command_set_kerr_dimensionless_spin_parameter_t)
        C(MAGIK_COMMAND_SET_ACTIVE_CAMERA, magik_command_set_active_camera_t)
        C(MAGIK_COMMAND_SET_CAMERA_POSITION, magik_command_set_camera_position_t)
        C(MAGIK_COMMAND_SET_CAMERA_ORIENTATION, magik_command_set_camera_orientation_t)
        C(MAGIK_COMMAND_SET_CAMERA_RESOLUTION, magik_command_set_camera_resolution_t)
        C(MAGIK_COMMAND_SET_N_INTRINSIC_SPECTRAL_IRRADIANCE_BIN, magik_command_set_n_intrinsic_spectral_irradiance_bin_t)
        C(MAGIK_COMMAND_SET_SAMPLE_DOMAIN_MAXIMUM_WAVELENGTH_NM, magik_command_set_sample_domain_maximum_wavelength_nm_t)
        C(MAGIK_COMMAND_SET_SAMPLE_DOMAIN_MINIMUM_WAVELENGTH_NM, magik_command_set_sample_domain_minimum_wavelength_nm_t)
        C(MAGIK_COMMAND_SET_REDISTRIBUTION_STANDARD_DEVIATION_NM, magik_command_set_redistribution_standard_deviation_nm_t)
        C(MAGIK_COMMAND_SET_COLOR_MATCHING_FUNCTION, magik_command_set_color_matching_function_t)
        C(MAGIK_COMMAND_SET_SCENE_COLORSPACE, magik_command_set_scene_colorspace_t)
        C(MAGIK_COMMAND_SET_VIEW_TRANSFORM, magik_command_set_view_transform_t)
        C(MAGIK_COMMAND_SET_DISPLAY_COLORSPACE, magik_command_set_display_colorspace_t)
        C(MAGIK_COMMAND_SET_OUTPUT_FORMAT, magik_command_set_output_format_t)
        C(MAGIK_COMMAND_SET_WHITE_BALANCE_WHITE_POINT, magik_command_set_white_balance_white_point_t)
        C(MAGIK_COMMAND_SET_WHITE_BALANCE_TINT, magik_command_set_white_balance_tint_t)
        C(MAGIK_COMMAND_SET_VIEW_TRANSFORM_REINHARD_EXTENDED_WHITE_LUMINANCE, magik_command_set_view_transform_reinhard_extended_white_luminance_t)
        C(MAGIK_COMMAND_SET_VIEW_TRANSFORM_KHRONOS_PBR_NEUTRAL_START_COMPRESSION,
          magik_command_set_view_transform_khronos_pbr_neutral_start_compression_t)
        C(MAGIK_COMMAND_SET_VIEW_TRANSFORM_KHRONOS_PBR_NEUTRAL_DESATURATION, magik_command_set_view_transform_khronos_pbr_neutral_desaturation_t)
        C(MAGIK_COMMAND_SET_VIEW_TRANSFORM_GT7_MODE, magik_command_set_view_transform_gt7_mode_t)
        C(MAGIK_COMMAND_SET_VIEW_TRANSFORM_GT7_SDR_PAPER_WHITE_BALANCE, magik_command_set_view_transform_gt7_sdr_paper_white_balance_t)
        C(MAGIK_COMMAND_SET_VIEW_TRANSFORM_GT7_REFERENCE_LUMINANCE, magik_command_set_view_transform_gt7_reference_luminance_t)
        C(MAGIK_COMMAND_SET_VIEW_TRANSFORM_GT7_JZAZBZ_EXPONENT_SCALE_FACTOR, magik_command_set_view_transform_gt7_jzazbz_exponent_scale_factor_t)
        C(MAGIK_COMMAND_SET_VIEW_TRANSFORM_AGX_LOOK, magik_command_set_view_transform_agx_look_t)
        C(MAGIK_COMMAND_SET_N_GEODESIC_STEP_LIMIT, magik_command_set_n_geodesic_step_limit_t)
        C(MAGIK_COMMAND_SET_N_REFLECTED_LIMIT, magik_command_set_n_reflected_limit_t)
#undef C
        default: return 0;
    }
}

// The stage: the reference's settings with their defaults (relativity_settings.h:11-25,
// optical_system_settings.h:17-34, color_settings.h:9-30, render_settings.h:16-17, stage.h:19-22).
struct Stage {
    int32_t metric = KERR;
    float abs_tol = 1e-3f, rel_tol = 1e-10f;
    uint32_t max_substeps = 56;
    float velocity[3] = {0.0f, 0.0f, 0.0f};
    float celestial_radius = 1000.0f;
    float schwarzschild_scalar = 1.01f, schwarzschild_mass = 6.7329546075e26f;
    float kerr_scalar = 1.01f, kerr_mass = 6.7329546075e26f, kerr_spin = 0.999f;
    float position[3] = {2.0f, 0.0f, 9.0f};
    float orientation[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    uint32_t W = 2195, H = 1860;
    float focal_mm = 12.0f, sensor_mm = 24.0f, iso = 500.0f, fps = 24.0f, shutter = 180.0f;
    uint32_t n_bin = 24;
    float wl_max = 830.0f, wl_min = 360.0f, sigma = 10.0f;
    mgk_display_t display{MAGIK_SCENE_COLORSPACE_LINEAR_REC709, MAGIK_VIEW_TRANSFORM_GRAN_TURISMO_7, MAGIK_DISPLAY_COLORSPACE_REC709,
                          MAGIK_OUTPUT_FORMAT_FLOAT32, 5600.0f, 0.0f, 4.0f, 0.8f, 0.15f, MAGIK_GRAN_TURISMO_7_MODE_JZAZBZ, 250.0f, 100.0f,
                          1.7f, MAGIK_AGX_LOOK_GOLDEN};
    uint32_t step_limit = 512, reflected_limit = 10;

    mgk_integrator_t integrator() const {
        mgk_integrator_t in{};
        in.metric_type = metric;
        in.precision = MGK_PRECISION_MIXED;
        in.time_reversed = 1;
        in.n_max_substep = max_substeps;
        in.mass_kg = metric == KERR ? kerr_mass : schwarzschild_mass;
        in.spin = kerr_spin;
        in.event_horizon_scalar = metric == KERR ? kerr_scalar : schwarzschild_scalar;
        in.h_max = 1000.0f;
        in.abs_tol = abs_tol;
        in.rel_tol = rel_tol;
        in.step_limit = step_limit;
        return in;
    }
    mgk_scene_t scene() const {
        mgk_scene_t sc{};
        sc.celestial_radius_m = celestial_radius;
        sc.flags = MGK_SCENE_BOXES;
        sc.n_box = 2;
        sc.max_reflections = reflected_limit;
        sc.led_scale = 7e5f;
        const float lo[2][3] = {{-20.0f, -10.0f, -20.0f}, {-10.0f, 9.5f, -10.0f}}, hi[2][3] = {{20.0f, 10.0f, 20.0f}, {10.0f, 10.5f, 10.0f}};
        for (int b = 0; b < 2; ++b)
            for (int k = 0; k < 3; ++k) {
                sc.box[b].min[k] = lo[b][k];
                sc.box[b].max[k] = hi[b][k];
            }
        sc.box[0].matter = MGK_MATTER_DIFFUSE;
        sc.box[1].matter = MGK_MATTER_ILLUMINANT;
        return sc;
    }
    mgk_spectral_t spectral() const {
        // the sample domain is the wavelength range of the bins and of the hero wavelengths
        mgk_spectral_t sp{};
        sp.n_hero = 4;
        sp.hero_min_nm = wl_min;
        sp.hero_max_nm = wl_max;
        sp.n_bin = n_bin;
        sp.bin_min_nm = wl_min;
        sp.bin_max_nm = wl_max;
        sp.sigma_nm = sigma;
        sp.iso = iso;
        sp.fps = fps;
        sp.shutter_angle = shutter;
        sp.white_point_K = display.white_point_K;
        sp.lambda_sampling = MGK_LAMBDA_CMF;   // the stage's sampler (kernel_generate_thin_lens_ray.cu:73-74)
        return sp;
    }
    mgk_camera_t camera() const {
        mgk_camera_t cam{};
        std::memcpy(cam.position, position, sizeof(position));
        std::memcpy(cam.orientation, orientation, sizeof(orientation));
        cam.focal_length_mm = focal_mm;
        cam.sensor_width_mm = sensor_mm;
        cam.x_resolution = W;
        cam.y_resolution = H;
        std::memcpy(cam.observer_velocity, velocity, sizeof(velocity));
        return cam;
    }
};

struct DeviceAov {
    uint32_t W = 0, H = 0, channels = 0;
    float* d = nullptr;
};
struct Collection {
    std::map<std::string, DeviceAov> aov;
};

struct CommandBuffer {
    std::vector<uint32_t> data;
    uint32_t n_occupied = 0;
};

// acc += x; out = acc / n (the progressive mean of the XYZ samples)
__global__ void __launch_bounds__(256) k_accumulate(const float* __restrict__ x, float* __restrict__ acc, float* __restrict__ out,
                                                    uint64_t n, float inv_samples) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x) {
        const float a = acc[i] + x[i];
        acc[i] = a;
        out[i] = a * inv_samples;
    }
}

double seconds(clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); }

// The statistics' queue histories, in this order.
enum QueueHistory { kHistAdvanceA, kHistAdvanceB, kHistIntersect, kHistBxdf, kHistRedistribute, kHistSplat };

// The reference's wavefront loop stores each queue's count once per iteration
// (ray.h:68-76, kernel_advance_geodesic.cu:97, kernel_intersect.cu:157,
// kernel_bxdf.cu:132, kernel_splat_to_renderbuffer.cu:76); iteration k takes
// every live path one step, to step k + 1.  The persistent kernels run each
// path through all its steps at once, so the same counts follow from the step
// at which each path ended or reflected:
//   advance_A[k] = intersect[k] = paths with more than k steps
//   bxdf[k]      = paths that met matter at step k + 1 (ended on it or reflected)
//   advance_B[k] = paths that go on to iteration k + 1 (advance_A[k + 1])
//   redistribute[k] = splat[k] = paths that ended at step k + 1
// A path stopped by the step limit leaves advance_A with no bxdf, redistribute
// or splat entry (kernel_enforce_ray_limits.cu:48-50 drops it).  The length
// is the largest step of any path.
void queue_histories(const mgk::StepHistogram& h, std::vector<uint32_t> (&q)[6]) {
    size_t len = 0;
    uint64_t n_paths = 0;
    for (size_t k = 0; k < h.end.size(); ++k) {
        n_paths += uint64_t(h.end[k]) + h.limit[k];
        if (h.end[k] || h.limit[k] || h.reflect[k]) len = k;
    }
    for (auto& v : q) v.assign(len, 0u);
    uint64_t alive = n_paths;   // paths with more than k steps
    for (size_t k = 0; k < len; ++k) {
        alive -= uint64_t(h.end[k]) + h.limit[k];
        const uint64_t next = alive - h.end[k + 1] - h.limit[k + 1];
        q[kHistAdvanceA][k] = q[kHistIntersect][k] = uint32_t(alive);
        q[kHistAdvanceB][k] = uint32_t(next);
        q[kHistBxdf][k] = h.end[k + 1] + h.reflect[k + 1];
        q[kHistRedistribute][k] = q[kHistSplat][k] = h.end[k + 1];
    }
}

}  // namespace

struct magik_render_manager {
    e_magik_manager_display_types display = MAGIK_DISPLAY_SWAPCHAIN;
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    // command queue: the DCC writes front, the worker consumes back
    uint32_t n_reserved = 4096;
    bool drop_overflow = false;
    std::unique_ptr<CommandBuffer> front_cmd, back_cmd;
    std::atomic<bool> swap_ready{true};
    // worker
    std::thread worker;
    std::atomic<bool> running{false}, stop{false};
    std::atomic<int> last_error{MAGIK_SUCCESS};
    Stage stage;
    bool clear = true, paused = false;
    std::map<std::string, DeviceAov> render;   // the render framebuffer
    // swapchain: the worker owns back, the DCC owns front, and the ready slot
    // is one word, coll index | kReadyNew, exchanged whole so that the index
    // and its new bit never part (the reference's ready_slot_t, m_swapchain.h)
    static constexpr uint32_t kReadyNew = 4u;
    Collection coll[3];
    uint32_t back = 2, front = 0;
    std::atomic<uint32_t> ready{1u};
    // statistics; history[q] holds queue q's count per iteration of the latest
    // sample (the order of kQueueHistories), copied out by each fresh fetch
    std::mutex stats_mutex;
    magik_renderer_statistics_t stats{};
    bool stats_new = false;
    std::vector<uint32_t> history[6];

    void fail(e_magik_result_types r) {
        int expected = MAGIK_SUCCESS;
        last_error.compare_exchange_strong(expected, int(r));
    }
};

struct magik_framebuffer_collection_external {
    e_magik_framebuffer_collection_config_types config = MAGIK_FRAMEBUFFER_COLLECTION_CONFIG_HOST;
    struct HostAov {
        uint32_t W = 0, H = 0, channels = 0;
        std::vector<float> data;
    };
    std::map<std::string, HostAov> aov;
};

namespace {

void free_aov(DeviceAov& a) {
    if (a.d) (void)hipFree(a.d);
    a.d = nullptr;
}

// One command (command.cpp:145-680): the reference's ranges, and a change of
// a setting restarts the accumulation.
void apply_command(magik_render_manager* m, const uint32_t* at, uint32_t type, size_t size) {
    Stage& st = m->stage;
    auto take = [&](auto& cmd) { std::memcpy(&cmd, at, size); };
    bool changed = true;
    switch (type) {
        case MAGIK_COMMAND_SET_STATE_RENDER: m->paused = false; changed = false; break;
        case MAGIK_COMMAND_SET_STATE_PAUSE: m->paused = true; changed = false; break;
        case MAGIK_COMMAND_CLEAR_RENDER_BUFFER: break;
        case MAGIK_COMMAND_ADD_AOV: {
            magik_command_add_aov_t c;
            take(c);
            const std::string name(c.name, c.length_of_name < 512 ? c.length_of_name : 512);
            auto it = m->render.find(name);
            if (it == m->render.end()) m->render.emplace(name, DeviceAov{0, 0, c.channels, nullptr});
            break;
        }
        case MAGIK_COMMAND_REMOVE_AOV: {
            magik_command_remove_aov_t c;
            take(c);
            const std::string name(c.name, c.length_of_name < 512 ? c.length_of_name : 512);
            auto it = m->render.find(name);
            if (it != m->render.end()) {
                free_aov(it->second);
                m->render.erase(it);
            }
            break;
        }
        case MAGIK_COMMAND_PRINTF: {
            magik_command_printf_t c;
            take(c);
            if (c.length_of_text < 512) std::printf("%.*s \n", int(c.length_of_text), c.text);
            changed = false;
            break;
        }
        case MAGIK_COMMAND_SET_METRIC_MINKOWSKI: st.metric = MINKOWSKI; break;
        case MAGIK_COMMAND_SET_METRIC_SCHWARZSCHILD: st.metric = SCHWARZSCHILD; break;
        case MAGIK_COMMAND_SET_METRIC_KERR: st.metric = KERR; break;
#define SET_IF(T, field, cond, target) \
    {                                  \
        T c;                           \
        take(c);                       \
        if (cond) target = c.field;    \
        else changed = false;          \
        break;                         \
    }
        case MAGIK_COMMAND_SET_ABSOLUTE_ERROR_TOLERANCE:
            SET_IF(magik_command_set_absolute_error_tolerance_t, absolute_error_tolerance, c.absolute_error_tolerance > 0.0f, st.abs_tol)
        case MAGIK_COMMAND_SET_RELATIVE_ERROR_TOLERANCE:
            SET_IF(magik_command_set_relative_error_tolerance_t, relative_error_tolerance, c.relative_error_tolerance > 0.0f, st.rel_tol)
        case MAGIK_COMMAND_SET_MAX_SUBSTEPS:
            SET_IF(magik_command_set_max_substeps_t, max_substeps, c.max_substeps > 0, st.max_substeps)
        case MAGIK_COMMAND_SET_OBSERVER_LOCAL_PROPER_VELOCITY: {
            magik_command_set_observer_local_proper_velocity_t c;
            take(c);
            st.velocity[0] = c.velocity_x; st.velocity[1] = c.velocity_y; st.velocity[2] = c.velocity_z;
            break;
        }
        case MAGIK_COMMAND_SET_CELESTIAL_SPHERE_RADIUS:
            SET_IF(magik_command_set_celestial_sphere_radius_t, celestial_sphere_radius, c.celestial_sphere_radius > 0.0f, st.celestial_radius)
        case MAGIK_COMMAND_SET_SCHWARZSCHILD_EVENT_HORIZON_SCALAR:
            SET_IF(magik_command_set_schwarzschild_event_horizon_scalar_t, scalar, c.scalar > 1.0f, st.schwarzschild_scalar)
        case MAGIK_COMMAND_SET_SCHWARZSCHILD_MASS_PARAMETER:
            SET_IF(magik_command_set_schwarzschild_mass_parameter_t, mass_parameter, true, st.schwarzschild_mass)
        case MAGIK_COMMAND_SET_KERR_EVENT_HORIZON_SCALAR:
            SET_IF(magik_command_set_kerr_event_horizon_scalar_t, scalar, c.scalar > 1.0f, st.kerr_scalar)
        case MAGIK_COMMAND_SET_KERR_MASS_PARAMETER:
            SET_IF(magik_command_set_kerr_mass_parameter_t, mass_parameter, true, st.kerr_mass)
        case MAGIK_COMMAND_SET_KERR_DIMENSIONLESS_SPIN_PARAMETER:
            SET_IF(magik_command_set_kerr_dimensionless_spin_parameter_t, dimensionless_spin_parameter, true, st.kerr_spin)
        case MAGIK_COMMAND_SET_ACTIVE_CAMERA: changed = false; break;   // one camera
        case MAGIK_COMMAND_SET_CAMERA_POSITION: {
            magik_command_set_camera_position_t c;
            take(c);
            st.position[0] = c.position_x; st.position[1] = c.position_y; st.position[2] = c.position_z;
            break;
        }
        case MAGIK_COMMAND_SET_CAMERA_ORIENTATION: {
            magik_command_set_camera_orientation_t c;
            take(c);
            mgk::rotate_matrix(c.rotation_z, c.rotation_y, c.rotation_x, st.orientation);   // command.cpp:386
            break;
        }
        case MAGIK_COMMAND_SET_CAMERA_RESOLUTION: {
            magik_command_set_camera_resolution_t c;
            take(c);
            st.W = c.x_resolution > 0 ? c.x_resolution : 2;
            st.H = c.y_resolution > 0 ? c.y_resolution : 2;
            break;
        }
        case MAGIK_COMMAND_SET_N_INTRINSIC_SPECTRAL_IRRADIANCE_BIN:
            SET_IF(magik_command_set_n_intrinsic_spectral_irradiance_bin_t, n_isi_bin, c.n_isi_bin > 0 && c.n_isi_bin <= 64, st.n_bin)
        case MAGIK_COMMAND_SET_SAMPLE_DOMAIN_MAXIMUM_WAVELENGTH_NM:
            SET_IF(magik_command_set_sample_domain_maximum_wavelength_nm_t, maximum_wavelength_nm, c.maximum_wavelength_nm > 0.0f, st.wl_max)
        case MAGIK_COMMAND_SET_SAMPLE_DOMAIN_MINIMUM_WAVELENGTH_NM:
            SET_IF(magik_command_set_sample_domain_minimum_wavelength_nm_t, minimum_wavelength_nm, c.minimum_wavelength_nm > 0.0f, st.wl_min)
        case MAGIK_COMMAND_SET_REDISTRIBUTION_STANDARD_DEVIATION_NM:
            SET_IF(magik_command_set_redistribution_standard_deviation_nm_t, standard_deviation_nm, c.standard_deviation_nm > 0.0f, st.sigma)
        case MAGIK_COMMAND_SET_COLOR_MATCHING_FUNCTION: changed = false; break;   // CIE 1931 only
        case MAGIK_COMMAND_SET_SCENE_COLORSPACE:
            SET_IF(magik_command_set_scene_colorspace_t, scene_colorspace, true, st.display.scene_colorspace)
        case MAGIK_COMMAND_SET_VIEW_TRANSFORM:
            SET_IF(magik_command_set_view_transform_t, view_transform, true, st.display.view_transform)
        case MAGIK_COMMAND_SET_DISPLAY_COLORSPACE:
            SET_IF(magik_command_set_display_colorspace_t, display_colorspace, true, st.display.display_colorspace)
        case MAGIK_COMMAND_SET_OUTPUT_FORMAT:
            SET_IF(magik_command_set_output_format_t, output_format, true, st.display.output_format)
        case MAGIK_COMMAND_SET_WHITE_BALANCE_WHITE_POINT:
            SET_IF(magik_command_set_white_balance_white_point_t, white_point_kelvin, true, st.display.white_point_K)
        case MAGIK_COMMAND_SET_WHITE_BALANCE_TINT:
            SET_IF(magik_command_set_white_balance_tint_t, tint, true, st.display.tint)
        case MAGIK_COMMAND_SET_VIEW_TRANSFORM_REINHARD_EXTENDED_WHITE_LUMINANCE:
            SET_IF(magik_command_set_view_transform_reinhard_extended_white_luminance_t, white_luminance, true,
                   st.display.reinhard_white_luminance)
        case MAGIK_COMMAND_SET_VIEW_TRANSFORM_KHRONOS_PBR_NEUTRAL_START_COMPRESSION:
            SET_IF(magik_command_set_view_transform_khronos_pbr_neutral_start_compression_t, start_compression, true,
                   st.display.khronos_start_compression)
        case MAGIK_COMMAND_SET_VIEW_TRANSFORM_KHRONOS_PBR_NEUTRAL_DESATURATION:
            SET_IF(magik_command_set_view_transform_khronos_pbr_neutral_desaturation_t, desaturation, true, st.display.khronos_desaturation)
        case MAGIK_COMMAND_SET_VIEW_TRANSFORM_GT7_MODE:
            SET_IF(magik_command_set_view_transform_gt7_mode_t, mode, true, st.display.gt7_mode)
        case MAGIK_COMMAND_SET_VIEW_TRANSFORM_GT7_SDR_PAPER_WHITE_BALANCE:
            SET_IF(magik_command_set_view_transform_gt7_sdr_paper_white_balance_t, sdr_paper_white_balance, c.sdr_paper_white_balance > 0.0f,
                   st.display.gt7_sdr_paper_white)
        case MAGIK_COMMAND_SET_VIEW_TRANSFORM_GT7_REFERENCE_LUMINANCE:
            SET_IF(magik_command_set_view_transform_gt7_reference_luminance_t, reference_luminance, c.reference_luminance > 0.0f,
                   st.display.gt7_reference_luminance)
        case MAGIK_COMMAND_SET_VIEW_TRANSFORM_GT7_JZAZBZ_EXPONENT_SCALE_FACTOR:
            SET_IF(magik_command_set_view_transform_gt7_jzazbz_exponent_scale_factor_t, jzazbz_exponent_scale_factor,
                   c.jzazbz_exponent_scale_factor > 1.0f, st.display.gt7_jzazbz_exponent_scale)
        case MAGIK_COMMAND_SET_VIEW_TRANSFORM_AGX_LOOK:
            SET_IF(magik_command_set_view_transform_agx_look_t, look, true, st.display.agx_look)
        case MAGIK_COMMAND_SET_N_GEODESIC_STEP_LIMIT:
            SET_IF(magik_command_set_n_geodesic_step_limit_t, n_geodesic_step_limit, c.n_geodesic_step_limit > 0, st.step_limit)
        case MAGIK_COMMAND_SET_N_REFLECTED_LIMIT:
            SET_IF(magik_command_set_n_reflected_limit_t, n_reflected_limit, c.n_reflected_limit > 0, st.reflected_limit)
#undef SET_IF
        default: changed = false; break;
    }
    if (changed) m->clear = true;
}

// Applies the back buffer (command.cpp:740-784), with the bound in chunks.
e_magik_result_types consume_commands(magik_render_manager* m) {
    CommandBuffer& b = *m->back_cmd;
    uint32_t pos = 0;
    e_magik_result_types r = MAGIK_SUCCESS;
    while (pos < b.n_occupied) {
        const uint32_t type = b.data[pos];
        const size_t size = command_size(type);
        if (size == 0 || size % 4 != 0) { r = MAGIK_ERROR_INVALID_COMMAND; break; }
        if (pos + size / 4 > b.n_occupied) { r = MAGIK_ERROR_OUT_OF_BOUNDS_COMMAND_BUFFER_READ; break; }
        apply_command(m, b.data.data() + pos, type, size);
        pos += uint32_t(size / 4);
    }
    b.n_occupied = 0;
    m->swap_ready.store(true, std::memory_order_release);
    return r;
}

#define W_TRY(expr)                                         \
    do {                                                    \
        if ((expr) != hipSuccess) {                         \
            m->fail(MGK_ERROR_HIP);                         \
            goto done;                                      \
        }                                                   \
    } while (0)

// The worker loop (renderer.cpp:23-144).
void run(magik_render_manager* m) {
    mgk_context_handle ctx = nullptr;
    float *d_xyz = nullptr, *d_acc = nullptr;
    uint64_t px_cap = 0, n_sample = 0;
    mgk::StepHistogram step_hist;   // of the latest sample, for the queue histories
    double total_sample_s = 0.0;
    const auto t_start = clk::now();
    if (hipSetDevice(m->device) != hipSuccess) {
        m->fail(MAGIK_INVALID_CUDA_DEVICE);
        return;
    }
    if (!m->stream) {
        if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) {
            m->fail(MGK_ERROR_HIP);
            return;
        }
        m->own_stream = true;
    }
    if (mgk_context_create(m->device, &ctx) != MAGIK_SUCCESS) {
        m->fail(MAGIK_INVALID_CUDA_DEVICE);
        return;
    }
    m->running.store(true, std::memory_order_release);
    while (!m->stop.load(std::memory_order_acquire)) {
        const auto t_it = clk::now();
        e_magik_result_types r = consume_commands(m);
        if (r != MAGIK_SUCCESS) {
            m->fail(r);
            break;
        }
        Stage& st = m->stage;
        const uint64_t npx = uint64_t(st.W) * st.H;
        // render framebuffers at the stage's resolution (framebuffer.cpp:177-200)
        bool dirty = false;
        for (auto& [name, a] : m->render) {
            if (a.W != st.W || a.H != st.H || !a.d) {
                dirty = true;
                free_aov(a);
                a.W = st.W;
                a.H = st.H;
                W_TRY(hipMalloc(&a.d, npx * a.channels * sizeof(float) + 4));
                W_TRY(hipMemsetAsync(a.d, 0, npx * a.channels * sizeof(float), m->stream));
            }
        }
        if (dirty) m->clear = true;
        if (m->display == MAGIK_DISPLAY_SWAPCHAIN) {
            // render -> back (framebuffer.cpp:202-241), then back <-> ready (:243-250)
            if (!dirty) {
                Collection& b = m->coll[m->back];
                for (auto it = b.aov.begin(); it != b.aov.end();) {
                    if (m->render.find(it->first) == m->render.end()) {
                        free_aov(it->second);
                        it = b.aov.erase(it);
                    } else {
                        ++it;
                    }
                }
                for (auto& [name, a] : m->render) {
                    DeviceAov& d = b.aov[name];
                    if (d.W != a.W || d.H != a.H || d.channels != a.channels || !d.d) {
                        free_aov(d);
                        d.W = a.W;
                        d.H = a.H;
                        d.channels = a.channels;
                        W_TRY(hipMalloc(&d.d, uint64_t(a.W) * a.H * a.channels * sizeof(float) + 4));
                    }
                    W_TRY(hipMemcpyAsync(d.d, a.d, uint64_t(a.W) * a.H * a.channels * sizeof(float), hipMemcpyDeviceToDevice, m->stream));
                }
                W_TRY(hipStreamSynchronize(m->stream));
            }
            m->back = m->ready.exchange(m->back | magik_render_manager::kReadyNew, std::memory_order_acq_rel) & 3u;
        }
        if (m->clear) {
            n_sample = 0;
            if (d_acc && px_cap) W_TRY(hipMemsetAsync(d_acc, 0, px_cap * 3 * sizeof(float), m->stream));
            m->clear = false;
        }
        auto test = m->render.find("test");
        if (test != m->render.end() && test->second.channels == 3 && !m->paused) {
            const auto t_s = clk::now();
            if (npx > px_cap) {
                if (d_xyz) (void)hipFree(d_xyz);
                if (d_acc) (void)hipFree(d_acc);
                d_xyz = d_acc = nullptr;
                W_TRY(hipMalloc(&d_xyz, npx * 3 * sizeof(float)));
                W_TRY(hipMalloc(&d_acc, npx * 3 * sizeof(float)));
                px_cap = npx;
                n_sample = 0;
            }
            if (n_sample == 0) W_TRY(hipMemsetAsync(d_acc, 0, npx * 3 * sizeof(float), m->stream));
            const mgk_integrator_t in = st.integrator();
            const mgk_scene_t sc = st.scene();
            const mgk_spectral_t sp = st.spectral();
            const mgk_camera_t cam = st.camera();
            r = mgk::render_samples(ctx, &in, &sc, &sp, &cam, 0, uint32_t(n_sample), 1, d_xyz, m->stream, &step_hist);
            if (r != MAGIK_SUCCESS) {
                m->fail(r);
                break;
            }
            const uint64_t n = npx * 3;
            const unsigned blocks = unsigned(std::min<uint64_t>((n + 255) / 256, 65536));
            hipLaunchKernelGGL(k_accumulate, dim3(blocks), dim3(256), 0, m->stream, d_xyz, d_acc, test->second.d, n,
                               1.0f / float(n_sample + 1));
            W_TRY(hipGetLastError());
            r = mgk_display_transform(ctx, &st.display, test->second.d, npx, test->second.d, m->stream);
            if (r != MAGIK_SUCCESS) {
                m->fail(r);
                break;
            }
            W_TRY(hipStreamSynchronize(m->stream));
            ++n_sample;
            const double dt = seconds(t_s, clk::now());
            total_sample_s += dt;
            std::vector<uint32_t> q[6];
            queue_histories(step_hist, q);
            std::lock_guard<std::mutex> lk(m->stats_mutex);
            m->stats.latest_wavefront_sample_time_seconds = dt;
            for (int i = 0; i < 6; ++i) m->history[i].swap(q[i]);
        } else {
            std::this_thread::sleep_for(std::chrono::milliseconds(m->paused ? 2 : 10));
        }
        {
            std::lock_guard<std::mutex> lk(m->stats_mutex);
            m->stats.x_sample = size_t(n_sample);
            m->stats.total_render_time_seconds = seconds(t_start, clk::now());
            m->stats.latest_render_loop_iteration_time_seconds = seconds(t_it, clk::now());
            m->stats.total_wavefront_sample_time_seconds = total_sample_s;
            size_t fb = 0;
            for (auto& [name, a] : m->render) fb += size_t(a.W) * a.H * a.channels * sizeof(float);
            m->stats.total_framebuffer_swapchain_size = fb * 4;
            m->stats.total_wavefront_buffer_size = size_t(px_cap) * 6 * sizeof(float);
            m->stats_new = true;
        }
    }
done:
    (void)hipStreamSynchronize(m->stream);
    if (d_xyz) (void)hipFree(d_xyz);
    if (d_acc) (void)hipFree(d_acc);
    mgk_context_destroy(ctx);
    m->running.store(false, std::memory_order_release);
}
#undef W_TRY

}  // namespace

extern "C" {

MGK_API e_magik_result_types magik_render_manager_create(magik_render_manager_t* manager, magik_manager_descriptor_t descriptor) {
    if (!manager) return MAGIK_ERROR_INVALID_POINTER;
    *manager = nullptr;
    if (descriptor.display_type != MAGIK_DISPLAY_SWAPCHAIN && descriptor.display_type != MAGIK_DISPLAY_HEADLESS) return MAGIK_UNKNOWN_ENUM_TYPE;
    magik_render_manager* m = new (std::nothrow) magik_render_manager();
    if (!m) return MAGIK_ERROR_HOST_MEMORY_ALLOCATION_FAILED;
    m->display = descriptor.display_type;
    m->device = int(descriptor.user_device_id);
    m->stream = hipStream_t(descriptor.user_stream);
    m->n_reserved = descriptor.cqs_n_reserved_chunk;
    m->drop_overflow = descriptor.cqs_drop_overflow;
    try {
        m->front_cmd = std::make_unique<CommandBuffer>();
        m->back_cmd = std::make_unique<CommandBuffer>();
        m->front_cmd->data.resize(m->n_reserved);
        m->back_cmd->data.resize(m->n_reserved);
        m->worker = std::thread(run, m);
    } catch (...) {
        delete m;
        return MAGIK_ERROR_COMMAND_BUFFER_ALLOCATION_FAILED;
    }
    *manager = m;
    return MAGIK_SUCCESS;
}

MGK_API e_magik_result_types magik_render_manager_destroy(magik_render_manager_t m) {
    if (!m) return MAGIK_ERROR_INVALID_POINTER;
    m->stop.store(true, std::memory_order_release);
    if (m->worker.joinable()) m->worker.join();
    if (hipSetDevice(m->device) == hipSuccess) {
        for (auto& [name, a] : m->render) free_aov(a);
        for (auto& c : m->coll)
            for (auto& [name, a] : c.aov) free_aov(a);
        if (m->own_stream && m->stream) (void)hipStreamDestroy(m->stream);
    }
    delete m;
    return MAGIK_SUCCESS;
}

MGK_API e_magik_result_types magik_render_manager_fetch_error(magik_render_manager_t m) {
    if (!m) return MAGIK_ERROR_INVALID_POINTER;
    return e_magik_result_types(m->last_error.load());
}

MGK_API e_magik_result_types magik_cqs_push_command(magik_render_manager_t m, const void* command) {
    if (!m || !command) return MAGIK_ERROR_INVALID_POINTER;
    if (reinterpret_cast<uintptr_t>(command) % sizeof(uint32_t) != 0) return MAGIK_ERROR_PACKED_COMMAND_NOT_ALLIGNED;
    uint32_t type;
    std::memcpy(&type, command, sizeof(type));
    const size_t size = command_size(type);
    if (size == 0) return MAGIK_ERROR_INVALID_COMMAND;
    if (size % sizeof(uint32_t) != 0) return MAGIK_ERROR_COMMAND_SIZE_NOT_A_MULTIPLE_OF_4;
    CommandBuffer& f = *m->front_cmd;
    if (size_t(f.n_occupied) * 4 + size > size_t(m->n_reserved) * 4)
        return m->drop_overflow ? MAGIK_SUCCESS : MAGIK_ERROR_COMMAND_BUFFER_OVERFLOW;
    std::memcpy(f.data.data() + f.n_occupied, command, size);
    f.n_occupied += uint32_t(size / 4);
    return MAGIK_SUCCESS;
}

MGK_API e_magik_result_types magik_cqs_dispatch_command_buffer(magik_render_manager_t m, bool* is_dispatched) {
    if (!m) {
        if (is_dispatched) *is_dispatched = false;
        return MAGIK_ERROR_INVALID_POINTER;
    }
    bool done = false;
    if (m->swap_ready.load(std::memory_order_acquire)) {
        m->front_cmd.swap(m->back_cmd);
        m->swap_ready.store(false, std::memory_order_release);
        done = true;
    }
    if (is_dispatched) *is_dispatched = done;
    return MAGIK_SUCCESS;
}

MGK_API e_magik_result_types magik_fetch_renderer_statistics(magik_render_manager_t m, bool* is_new_fetch,
                                                             magik_renderer_statistics_t* out) {
    if (!m || !out) {
        if (is_new_fetch) *is_new_fetch = false;
        return MAGIK_ERROR_INVALID_POINTER;
    }
    std::lock_guard<std::mutex> lk(m->stats_mutex);
    const bool fresh = m->stats_new;
    if (fresh) {
        // the histories go out in one host allocation owned by the caller
        // until magik_destroy_renderer_statistics
        const size_t len = m->history[0].size();
        uint32_t* block = nullptr;
        if (len > 0) {
            block = static_cast<uint32_t*>(std::malloc(6 * len * sizeof(uint32_t)));
            if (!block) {
                if (is_new_fetch) *is_new_fetch = false;
                return MAGIK_ERROR_HOST_MEMORY_ALLOCATION_FAILED;
            }
            for (int i = 0; i < 6; ++i) std::memcpy(block + i * len, m->history[i].data(), len * sizeof(uint32_t));
        }
        *out = m->stats;
        out->latest_wavefront_history_length = len;
        uint32_t** dst[6] = {&out->h_latest_advance_A_queue_count_history, &out->h_latest_advance_B_queue_count_history,
                             &out->h_latest_intersect_queue_count_history, &out->h_latest_bxdf_queue_count_history,
                             &out->h_latest_redistribute_queue_count_history, &out->h_latest_splat_queue_count_history};
        static_assert(kHistAdvanceA == 0 && kHistAdvanceB == 1 && kHistIntersect == 2 && kHistBxdf == 3 &&
                          kHistRedistribute == 4 && kHistSplat == 5,
                      "history order");
        for (int i = 0; i < 6; ++i) *dst[i] = block ? block + i * len : nullptr;
        m->stats_new = false;
    }
    if (is_new_fetch) *is_new_fetch = fresh;
    return MAGIK_SUCCESS;
}

MGK_API e_magik_result_types magik_destroy_renderer_statistics(magik_renderer_statistics_t* s) {
    if (!s) return MAGIK_ERROR_INVALID_POINTER;
    std::free(s->h_latest_advance_A_queue_count_history);   // the start of the fetch's one allocation
    std::memset(s, 0, sizeof(*s));
    return MAGIK_SUCCESS;
}

MGK_API e_magik_result_types magik_configure_framebuffer_collection(magik_framebuffer_collection_external_t* fb,
                                                                    e_magik_framebuffer_collection_config_types config) {
    if (!fb) return MAGIK_ERROR_INVALID_POINTER;
    *fb = nullptr;
    if (config != MAGIK_FRAMEBUFFER_COLLECTION_CONFIG_HOST) return MAGIK_UNKNOWN_ENUM_TYPE;
    magik_framebuffer_collection_external* c = new (std::nothrow) magik_framebuffer_collection_external();
    if (!c) return MAGIK_ERROR_HOST_MEMORY_ALLOCATION_FAILED;
    c->config = config;
    *fb = c;
    return MAGIK_SUCCESS;
}

// Takes the newest ready frame as the front buffer and copies each of its
// AOVs into the collection's host memory (framebuffer.cpp:22-52, :272-312).
MGK_API e_magik_result_types magik_framebuffer_collection_fetch(magik_render_manager_t m, bool* is_new_fetch,
                                                                magik_framebuffer_collection_external_t fb) {
    if (is_new_fetch) *is_new_fetch = false;
    if (!m || !fb) return MAGIK_ERROR_INVALID_POINTER;
    if (!m->running.load(std::memory_order_acquire)) return MAGIK_SUCCESS;
    if (m->display == MAGIK_DISPLAY_HEADLESS) return MAGIK_ERROR_AOV_SWAPCHAIN_NOT_INITALIZED;
    // take the ready slot only while it holds a frame not fetched yet
    uint32_t slot = m->ready.load(std::memory_order_acquire);
    do {
        if (!(slot & magik_render_manager::kReadyNew)) return MAGIK_SUCCESS;
    } while (!m->ready.compare_exchange_weak(slot, m->front, std::memory_order_acq_rel, std::memory_order_acquire));
    m->front = slot & 3u;
    const Collection& front = m->coll[m->front];
    int old = 0;
    (void)hipGetDevice(&old);
    if (hipSetDevice(m->device) != hipSuccess) return MAGIK_INVALID_CUDA_DEVICE;
    e_magik_result_types r = MAGIK_SUCCESS;
    for (auto it = fb->aov.begin(); it != fb->aov.end();) {
        if (front.aov.find(it->first) == front.aov.end()) it = fb->aov.erase(it);
        else ++it;
    }
    for (auto& [name, d] : front.aov) {
        auto& h = fb->aov[name];
        const size_t n = size_t(d.W) * d.H * d.channels;
        h.W = d.W;
        h.H = d.H;
        h.channels = d.channels;
        h.data.resize(n);
        if (n && hipMemcpy(h.data.data(), d.d, n * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) r = MGK_ERROR_HIP;
    }
    (void)hipSetDevice(old);
    if (is_new_fetch) *is_new_fetch = r == MAGIK_SUCCESS;
    return r;
}

MGK_API e_magik_result_types magik_extract_config_host(magik_render_manager_t m, magik_framebuffer_container_config_host_t* container,
                                                       magik_framebuffer_collection_external_t fb, const char* name) {
    if (!m || !container || !fb || !name) return MAGIK_ERROR_INVALID_POINTER;
    if (fb->config != MAGIK_FRAMEBUFFER_COLLECTION_CONFIG_HOST) return MAGIK_ERROR_AOV_INCORRECT_EXTRACT_CALL;
    auto it = fb->aov.find(name);
    if (it == fb->aov.end()) return MAGIK_SUCCESS;   // framebuffer.cpp:71-74: nothing yet
    container->x_resolution = it->second.W;
    container->y_resolution = it->second.H;
    container->channels = it->second.channels;
    container->size_of_data = it->second.data.size() * sizeof(float);
    container->h_data = it->second.data.data();
    return MAGIK_SUCCESS;
}

MGK_API e_magik_result_types magik_destroy_framebuffer_collection(magik_framebuffer_collection_external_t fb) {
    delete fb;
    return MAGIK_SUCCESS;
}

}  // extern "C"
