"""ctypes mirror of include/magik_amd.h and the loader of the HIP library.

The Python side is plumbing for tests and the benchmark: the product is the
C-ABI shared library ``magik_amd/libmagik_amd.so`` built from
``magik_amd/csrc``.  Loading fails loudly when the library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmagik_amd.so")

# e_magik_result_types (include/magik_amd.h)
MAGIK_SUCCESS = 0
MAGIK_ERROR_INVALID_POINTER = 2
MAGIK_UNKNOWN_ENUM_TYPE = 3
MAGIK_INVALID_CUDA_DEVICE = 4
MAGIK_ERROR_HOST_TRACE_ALLOCATED_STEPS_AND_CAPACITY_MISMATCH = 102
MAGIK_ERROR_DEVICE_ALLOCATION_FAILED = 201
MAGIK_ERROR_UNKNOWN_METRIC = 707
MAGIK_ERROR_UNKNOWN = 10000
MGK_ERROR_INVALID_ARGUMENT = 20000
MGK_ERROR_CAPACITY_EXCEEDED = 20001
MGK_ERROR_HIP = 20002

MINKOWSKI, SCHWARZSCHILD, KERR = 0, 1, 2
MASSIVE, MASSLESS = 0, 1

PREC_FP32_REFERENCE, PREC_MIXED, PREC_FP64 = 0, 1, 2

(HIT_NONE, HIT_HORIZON, HIT_ESCAPE, HIT_DISK, HIT_SPHERE, HIT_STEP_LIMIT, HIT_BOX_DIFFUSE, HIT_BOX_EMITTER,
 HIT_MESH_DIFFUSE, HIT_MESH_EMITTER) = range(10)
HIT_NAMES = {0: "none", 1: "horizon", 2: "escape", 3: "disk", 4: "sphere", 5: "step_limit", 6: "box_diffuse",
             7: "box_emitter", 8: "mesh_diffuse", 9: "mesh_emitter"}

SCENE_DISK, SCENE_SPHERE, SCENE_STARS, SCENE_BACKGROUND, SCENE_BOXES, SCENE_MESHES = 1, 2, 4, 8, 16, 32
SCENE_DISK_ON_GEODESIC = 64
CHORD_AFFINE, CHORD_HERMITE = 0, 1
MAX_MESH_TRIANGLES = MAX_INSTANCES = 16777215
MATTER_DIFFUSE, MATTER_ILLUMINANT = 1, 2


class MetricParameters(C.Structure):  # magik_trace_geodesic_metric_parameters_t
    _fields_ = [("metric_type", C.c_int32), ("time_reversed", C.c_bool), ("event_horizon_scalar", C.c_float),
                ("schwarzschild_mass", C.c_float), ("kerr_mass", C.c_float), ("kerr_spin", C.c_float)]


class IntegratorParameters(C.Structure):  # magik_trace_geodesic_integrator_parameters_t
    _fields_ = [("n_integrator_max_substep", C.c_uint32), ("integrator_absolute_error_tolerance", C.c_float),
                ("integrator_relative_error_tolerance", C.c_float)]


class InitialConditions(C.Structure):  # magik_trace_geodesic_inital_conditions_t
    _fields_ = [("test_particle_type", C.c_int32)] + [(n, C.c_float) for n in (
        "spatial_position_x", "spatial_position_y", "spatial_position_z", "x_rotation", "y_rotation", "z_rotation",
        "local_proper_velocity_x", "local_proper_velocity_y", "local_proper_velocity_z", "massive_particle_speed")]


class Step(C.Structure):  # magik_trace_geodesic_step_t
    _fields_ = [("four_position", C.c_float * 4), ("four_momentum", C.c_float * 4), ("moment_derivative", C.c_float * 4),
                ("four_vector_contraction", C.c_float), ("affine_parameter", C.c_float),
                ("delta_affine_parameter", C.c_float)]


class Container(C.Structure):  # magik_trace_geodesic_container_t
    _fields_ = [("capacity", C.c_uint64), ("n_allocated_step", C.c_uint32), ("n_written", C.c_uint32),
                ("steps", C.POINTER(Step))]


class Integrator(C.Structure):  # mgk_integrator_t
    _fields_ = [("metric_type", C.c_int32), ("precision", C.c_int32), ("time_reversed", C.c_uint32),
                ("n_max_substep", C.c_uint32), ("mass_kg", C.c_float), ("spin", C.c_float),
                ("event_horizon_scalar", C.c_float), ("h_max", C.c_float), ("abs_tol", C.c_float),
                ("rel_tol", C.c_float), ("step_limit", C.c_uint32), ("reserved", C.c_uint32)]


class Camera(C.Structure):  # mgk_camera_t
    _fields_ = [("position", C.c_float * 3), ("orientation", C.c_float * 9), ("focal_length_mm", C.c_float),
                ("sensor_width_mm", C.c_float), ("x_resolution", C.c_uint32), ("y_resolution", C.c_uint32),
                ("observer_velocity", C.c_float * 3), ("camera_id", C.c_uint32)]


MAX_BOX = 4


class Box(C.Structure):  # mgk_box_t
    _fields_ = [("min", C.c_float * 3), ("max", C.c_float * 3), ("matter", C.c_uint32), ("reserved", C.c_uint32)]


class Scene(C.Structure):  # mgk_scene_t
    _fields_ = [("celestial_radius_m", C.c_float), ("flags", C.c_uint32), ("disk_r_in", C.c_float),
                ("disk_r_out", C.c_float), ("disk_T_in", C.c_float), ("sphere_center", C.c_float * 3),
                ("sphere_radius", C.c_float), ("sphere_T", C.c_float), ("sphere_checker", C.c_uint32),
                ("background_scale", C.c_float), ("background_T", C.c_float), ("star_scale", C.c_float),
                ("star_T", C.c_float), ("star_radius_rad", C.c_float), ("n_star", C.c_uint32),
                ("star_seed", C.c_uint64), ("n_box", C.c_uint32), ("max_reflections", C.c_uint32),
                ("led_scale", C.c_float), ("chord", C.c_uint32), ("box", Box * MAX_BOX)]


class Mesh(C.Structure):  # mgk_mesh_t
    _fields_ = [("vertices", C.POINTER(C.c_float)), ("n_vertex", C.c_uint32), ("n_triangle", C.c_uint32),
                ("indices", C.POINTER(C.c_uint32))]


class Instance(C.Structure):  # mgk_instance_t
    _fields_ = [("mesh", C.c_uint32), ("matter", C.c_uint32), ("object_to_world", C.c_float * 12)]


class MeshHierarchyInfo(C.Structure):  # mgk_mesh_hierarchy_info_t
    _fields_ = [("n_nodes", C.c_uint64), ("n_triangles", C.c_uint64), ("max_depth", C.c_uint32), ("max_leaf", C.c_uint32),
                ("world_min", C.c_float * 3), ("world_max", C.c_float * 3)]


LAMBDA_UNIFORM = 0   # mgk_lambda_sampling_t
LAMBDA_CMF = 1


class Spectral(C.Structure):  # mgk_spectral_t
    _fields_ = [("n_hero", C.c_uint32), ("hero_min_nm", C.c_float), ("hero_max_nm", C.c_float),
                ("n_bin", C.c_uint32), ("bin_min_nm", C.c_float), ("bin_max_nm", C.c_float), ("sigma_nm", C.c_float),
                ("iso", C.c_float), ("fps", C.c_float), ("shutter_angle", C.c_float), ("white_point_K", C.c_float),
                ("lambda_sampling", C.c_uint32)]


class Display(C.Structure):  # mgk_display_t
    _fields_ = [("scene_colorspace", C.c_uint32), ("view_transform", C.c_uint32), ("display_colorspace", C.c_uint32),
                ("output_format", C.c_uint32), ("white_point_K", C.c_float), ("tint", C.c_float),
                ("reinhard_white_luminance", C.c_float), ("khronos_start_compression", C.c_float),
                ("khronos_desaturation", C.c_float), ("gt7_mode", C.c_uint32), ("gt7_sdr_paper_white", C.c_float),
                ("gt7_reference_luminance", C.c_float), ("gt7_jzazbz_exponent_scale", C.c_float), ("agx_look", C.c_uint32)]


class Hit(C.Structure):  # mgk_hit_t
    _fields_ = [("cls", C.c_uint32), ("g", C.c_float), ("a0", C.c_float), ("a1", C.c_float)]


class BatchCamera(C.Structure):  # mgk_batch_camera_t
    _fields_ = [("camera", Camera), ("spin", C.c_float), ("reserved", C.c_uint32)]


class TraceStats(C.Structure):  # mgk_trace_stats_t
    _fields_ = [("n_rays", C.c_uint64), ("n_steps", C.c_uint64), ("n_trials", C.c_uint64),
                ("n_trials_fp64", C.c_uint64), ("ms", C.c_double), ("n_reflections", C.c_uint64)]


class SurrogateRanges(C.Structure):  # mgk_surrogate_ranges_t
    _fields_ = [(n, C.c_float) for n in ("spin_min", "spin_max", "r_min", "r_max", "theta_min_deg", "theta_max_deg",
                                         "screen_max", "disk_r_max")]


class TrainConfig(C.Structure):  # mgk_train_config_t
    _fields_ = [("batch", C.c_uint32), ("epochs", C.c_uint32), ("lr", C.c_float), ("lr_final", C.c_float),
                ("w_r", C.c_float), ("w_d", C.c_float)]


class Fallback(C.Structure):  # mgk_fallback_t
    _fields_ = [("min_confidence", C.c_float), ("critical_margin", C.c_float), ("direction_margin", C.c_float),
                ("disk_edge_margin", C.c_float), ("judge_stride", C.c_uint32), ("trace_disk", C.c_uint32),
                ("analytic", C.c_uint32)]


class SurrogateStats(C.Structure):  # mgk_surrogate_stats_t
    _fields_ = [(n, C.c_uint64) for n in ("n_paths", "n_fallback", "n_fallback_confidence", "n_fallback_critical",
                                          "n_fallback_disk", "n_fallback_star", "n_fallback_domain", "n_judge",
                                          "judge_class_agree", "judge_n_disk", "judge_n_escape")] + \
               [(n, C.c_double) for n in ("judge_mean_rel_r", "judge_mean_rel_g", "judge_mean_dir", "judge_max_dir")] + \
               [("n_trials", C.c_uint64)] + \
               [(n, C.c_double) for n in ("ms_raygen", "ms_mlp", "ms_trace", "ms_shade", "ms_total")] + \
               [("n_fallback_analytic", C.c_uint64)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


# Every symbol include/magik_amd.h declares, with its ctypes signature.
class ManagerDescriptor(C.Structure):  # magik_manager_descriptor_t (include/magik_manager.h)
    _fields_ = [("display_type", C.c_int32), ("user_device_id", C.c_uint32), ("user_stream", C.c_void_p),
                ("cqs_n_reserved_chunk", C.c_uint32), ("cqs_drop_overflow", C.c_bool)]


QUEUE_HISTORIES = ("advance_A", "advance_B", "intersect", "bxdf", "redistribute", "splat")


class RendererStatistics(C.Structure):  # magik_renderer_statistics_t
    _fields_ = [("x_sample", C.c_size_t), ("total_render_time_seconds", C.c_double),
                ("latest_render_loop_iteration_time_seconds", C.c_double),
                ("total_wavefront_sample_time_seconds", C.c_double), ("latest_wavefront_sample_time_seconds", C.c_double),
                ("latest_wavefront_history_length", C.c_size_t)] + \
               [(f"h_latest_{q}_queue_count_history", C.c_void_p) for q in QUEUE_HISTORIES] + \
               [("total_framebuffer_swapchain_size", C.c_size_t), ("total_wavefront_buffer_size", C.c_size_t)]


class HostContainer(C.Structure):  # magik_framebuffer_container_config_host_t
    _fields_ = [("x_resolution", C.c_uint32), ("y_resolution", C.c_uint32), ("channels", C.c_uint32),
                ("size_of_data", C.c_size_t), ("h_data", C.POINTER(C.c_float))]


class AovCommand(C.Structure):  # magik_command_add_aov_t (remove: without channels)
    _fields_ = [("embedded_type", C.c_uint32), ("length_of_name", C.c_uint32), ("name", C.c_char * 512),
                ("channels", C.c_uint32)]


MAGIK_DISPLAY_SWAPCHAIN, MAGIK_DISPLAY_HEADLESS = 0, 1
FB_CONFIG_HOST, FB_CONFIG_CUDA, FB_CONFIG_OPENGL, FB_CONFIG_VULKAN = 0, 1, 2, 3
CMD = {"render": 1000, "pause": 1001, "clear": 1002, "add_aov": 1003, "remove_aov": 1004, "printf": 2000,
       "minkowski": 4000, "schwarzschild": 4001, "kerr": 4002, "abs_tol": 4100, "rel_tol": 4101, "max_substeps": 4102,
       "velocity": 4200, "celestial_radius": 4300, "schwarzschild_scalar": 4400, "schwarzschild_mass": 4401,
       "kerr_scalar": 4402, "kerr_mass": 4403, "kerr_spin": 4404, "active_camera": 5000, "camera_position": 5001,
       "camera_orientation": 5002, "camera_resolution": 5003, "n_bin": 6000, "wl_max": 6001, "wl_min": 6002,
       "sigma": 6003, "cmf": 6100, "scene_colorspace": 6101, "view_transform": 6102, "display_colorspace": 6103,
       "output_format": 6104, "white_point": 6200, "tint": 6201, "reinhard_white": 6300, "khronos_start": 6301,
       "khronos_desaturation": 6302, "gt7_mode": 6304, "gt7_paper_white": 6305, "gt7_reference_luminance": 6306,
       "gt7_jzazbz_scale": 6407, "agx_look": 6308, "step_limit": 7000, "reflected_limit": 7001}

_P = C.c_void_p
_R = C.c_int32
SIGNATURES = {
    "magik_trace_geodesic": (_R, [C.POINTER(Container), InitialConditions, MetricParameters, IntegratorParameters]),
    "magik_get_version": (_R, [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                               C.POINTER(C.c_char_p)]),
    "mgk_context_create": (_R, [C.c_int, C.POINTER(_P)]),
    "mgk_context_destroy": (_R, [_P]),
    "mgk_set_scene": (_R, [_P, C.POINTER(Scene)]),
    "mgk_set_meshes": (_R, [_P, C.POINTER(Mesh), C.c_uint32, C.POINTER(Instance), C.c_uint32]),
    "mgk_mesh_hierarchy_info": (_R, [C.POINTER(Mesh), C.c_uint32, C.POINTER(Instance), C.c_uint32,
                                     C.POINTER(MeshHierarchyInfo)]),
    "mgk_generate_rays": (_R, [_P, C.POINTER(Integrator), C.POINTER(Camera), C.c_uint64, C.c_uint32, C.c_uint32, _P,
                               C.c_uint64, _P]),
    "mgk_trace_rays": (_R, [_P, C.POINTER(Integrator), C.POINTER(Scene), C.POINTER(C.c_float), _P, C.c_uint64,
                            C.c_uint64, _P, _P, _P, _P]),
    "mgk_trace_rays_noalloc": (_R, [_P, C.POINTER(Integrator), C.POINTER(Scene), C.POINTER(C.c_float), _P, C.c_uint64,
                                    C.c_uint64, _P, _P, _P, _P]),
    "mgk_trace_rays_final": (_R, [_P, C.POINTER(Integrator), C.POINTER(Scene), C.POINTER(C.c_float), _P, C.c_uint64,
                                  C.c_uint64, _P, _P, _P, _P]),
    "mgk_trace_paths": (_R, [_P, C.POINTER(Integrator), C.POINTER(Scene), C.POINTER(Camera), C.c_uint64, C.c_uint32,
                             C.c_uint32, _P, C.c_uint64, _P, _P, _P, _P]),
    "mgk_generate_rays_batch": (_R, [_P, C.POINTER(Integrator), C.POINTER(BatchCamera), C.c_uint32, C.c_uint64,
                                     C.c_uint32, _P, C.c_uint64, C.POINTER(C.c_uint64), _P]),
    "mgk_trace_rays_batch": (_R, [_P, C.POINTER(Integrator), C.POINTER(Scene), _P, C.c_uint64, C.c_uint64, _P, _P, _P,
                                  _P]),
    "mgk_last_trace_stats": (_R, [_P, C.POINTER(TraceStats)]),
    "mgk_shade": (_R, [_P, C.POINTER(Scene), C.POINTER(Spectral), C.POINTER(Camera), C.c_uint64, C.c_uint32,
                       C.c_uint32, _P, _P, _P]),
    "mgk_render_frame": (_R, [_P, C.POINTER(Integrator), C.POINTER(Scene), C.POINTER(Spectral), C.POINTER(Camera),
                              C.c_uint64, C.c_uint32, _P, _P]),
    "mgk_color_weights": (_R, [C.POINTER(Spectral), C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "mgk_cmf_inverse_cdf": (_R, [C.POINTER(Spectral), C.POINTER(C.c_float)]),
    "mgk_display_transform": (_R, [_P, C.POINTER(Display), _P, C.c_uint64, _P, _P]),
    "mgk_camera_boyer_lindquist": (_R, [C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, C.c_uint32,
                                        C.c_uint32, C.POINTER(Camera)]),
    "mgk_surrogate_create": (_R, [C.c_int, C.c_uint64, C.POINTER(SurrogateRanges), C.POINTER(_P)]),
    "mgk_surrogate_destroy": (_R, [_P]),
    "mgk_surrogate_n_params": (_R, [C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "mgk_surrogate_get_weights": (_R, [_P, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "mgk_surrogate_set_weights": (_R, [_P, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "mgk_surrogate_forward": (_R, [_P, _P, C.c_uint64, _P, _P]),
    "mgk_surrogate_gradients": (_R, [_P, _P, _P, C.c_uint32, _P, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "mgk_surrogate_train_step": (_R, [_P, _P, _P, C.c_uint32, C.c_float, C.c_float, C.c_float,
                                      C.POINTER(C.c_float)]),
    "mgk_surrogate_train": (_R, [_P, _P, _P, C.c_uint64, C.POINTER(TrainConfig), C.POINTER(C.c_float)]),
    "mgk_surrogate_generate_data": (_R, [_P, _P, C.POINTER(Integrator), C.POINTER(Scene), C.c_uint64, C.c_uint64,
                                         _P, _P, _P, _P, _P]),
    "mgk_render_frame_surrogate": (_R, [_P, _P, C.POINTER(Integrator), C.POINTER(Scene), C.POINTER(Spectral),
                                        C.POINTER(Camera), C.c_uint64, C.c_uint32, C.POINTER(Fallback), _P, _P,
                                        C.POINTER(SurrogateStats), _P]),
    "mgk_mino_disk": (_R, [C.POINTER(Integrator), C.POINTER(Scene), _P, C.c_uint64, C.c_uint64, _P, _P, _P]),
    "magik_render_manager_create": (_R, [C.POINTER(_P), ManagerDescriptor]),
    "magik_render_manager_destroy": (_R, [_P]),
    "magik_render_manager_fetch_error": (_R, [_P]),
    "magik_cqs_push_command": (_R, [_P, _P]),
    "magik_cqs_dispatch_command_buffer": (_R, [_P, C.POINTER(C.c_bool)]),
    "magik_fetch_renderer_statistics": (_R, [_P, C.POINTER(C.c_bool), C.POINTER(RendererStatistics)]),
    "magik_destroy_renderer_statistics": (_R, [C.POINTER(RendererStatistics)]),
    "magik_configure_framebuffer_collection": (_R, [C.POINTER(_P), C.c_int32]),
    "magik_framebuffer_collection_fetch": (_R, [_P, C.POINTER(C.c_bool), _P]),
    "magik_extract_config_host": (_R, [_P, C.POINTER(HostContainer), _P, C.c_char_p]),
    "magik_destroy_framebuffer_collection": (_R, [_P]),
}

_lib = None


def load():
    """Load libmagik_amd.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
        # One HIP runtime per process: when torch (which carries its own HIP
        # and HSA runtime) is present, load it first so that the library binds
        # to that runtime instead of starting a second one beside it.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            if not hasattr(lib, name):
                raise RuntimeError(f"{LIB_PATH} does not export {name}: rebuild it")
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(code: int, what: str = "") -> None:
    if code != MAGIK_SUCCESS:
        raise RuntimeError(f"magik_amd call {what} failed with result code {code}")
