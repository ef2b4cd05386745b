"""magik_amd: MI355X-native wavefront spectral geodesic tracer.

The product is the C-ABI library ``libmagik_amd.so`` (include/magik_amd.h);
this module is the thin Python host mirror used by the tests and the
benchmark: it fills the ABI structs, hands torch-allocated device buffers to
the library and reads results back.  Nothing here computes on the CPU in place
of the HIP kernels; every call fails loudly when the library is missing.
"""
from __future__ import annotations

import ctypes as C
import math

from . import abi
from . import mesh
from .abi import (LAMBDA_CMF, LAMBDA_UNIFORM, CHORD_AFFINE, CHORD_HERMITE, HIT_BOX_DIFFUSE, HIT_BOX_EMITTER, HIT_DISK, HIT_ESCAPE, HIT_HORIZON,
                  HIT_MESH_DIFFUSE, HIT_MESH_EMITTER, HIT_NONE, HIT_SPHERE, HIT_STEP_LIMIT, KERR, MATTER_DIFFUSE,
                  MATTER_ILLUMINANT, MINKOWSKI, PREC_FP32_REFERENCE, PREC_FP64, PREC_MIXED, SCHWARZSCHILD,
                  SCENE_BACKGROUND, SCENE_BOXES, SCENE_DISK, SCENE_DISK_ON_GEODESIC, SCENE_MESHES, SCENE_SPHERE,
                  SCENE_STARS)

DEFAULT_MASS_KG = 6.7329546075e26       # relativity_settings.h:21,24
G, C_LIGHT = 6.67430e-11, 299792458.0


def length_scale(mass_kg: float = DEFAULT_MASS_KG) -> float:
    """G M / c^2 in metres (metric_kerr.cuh:154-157), in float32 like the library."""
    import numpy as np
    f = np.float32
    return float(f(G) * f(mass_kg) / (f(C_LIGHT) * f(C_LIGHT)))


def r_isco(a: float) -> float:
    """Prograde ISCO radius in units of M (Bardeen, Press & Teukolsky 1972)."""
    a = abs(a)
    z1 = 1 + (1 - a * a) ** (1 / 3) * ((1 + a) ** (1 / 3) + (1 - a) ** (1 / 3))
    z2 = math.sqrt(3 * a * a + z1 * z1)
    return 3 + z2 - math.sqrt((3 - z1) * (3 + z1 + 2 * z2))


def integrator(metric=KERR, precision=PREC_MIXED, spin=0.9, mass_kg=DEFAULT_MASS_KG, abs_tol=1e-7, rel_tol=1e-7,
               n_max=56, step_limit=10000, h_max=1000.0, time_reversed=True, horizon_scalar=1.01) -> abi.Integrator:
    return abi.Integrator(metric_type=metric, precision=precision, time_reversed=int(time_reversed), n_max_substep=n_max,
                          mass_kg=mass_kg, spin=spin, event_horizon_scalar=horizon_scalar, h_max=h_max,
                          abs_tol=abs_tol, rel_tol=rel_tol, step_limit=step_limit)


def spectral(n_hero=4, hero_min=380.0, hero_max=780.0, n_bin=24, bin_min=360.0, bin_max=830.0, sigma=10.0,
             iso=500.0, fps=24.0, shutter=180.0, white_K=5600.0, lambda_sampling=abi.LAMBDA_UNIFORM) -> abi.Spectral:
    """Defaults of color_settings.h:11-14 and optical_system_settings.h:22-27; lambda_sampling
    LAMBDA_CMF draws the hero wavelengths through the reference's CMF inverse CDF."""
    return abi.Spectral(n_hero=n_hero, hero_min_nm=hero_min, hero_max_nm=hero_max, n_bin=n_bin, bin_min_nm=bin_min,
                        bin_max_nm=bin_max, sigma_nm=sigma, iso=iso, fps=fps, shutter_angle=shutter,
                        white_point_K=white_K, lambda_sampling=lambda_sampling)


def cmf_inverse_cdf(spec):
    """The 256 wavelengths (nm) of the inverse CDF that LAMBDA_CMF samples over the hero range."""
    import numpy as np
    out = np.zeros(256, dtype=np.float32)
    abi.check(abi.load().mgk_cmf_inverse_cdf(C.byref(spec), out.ctypes.data_as(C.POINTER(C.c_float))), "mgk_cmf_inverse_cdf")
    return out


def camera_bl(r_obs: float, theta_deg: float, spin: float, fov_deg: float, W: int, H: int,
              mass_kg: float = DEFAULT_MASS_KG, camera_id: int = 0) -> abi.Camera:
    cam = abi.Camera()
    abi.check(abi.load().mgk_camera_boyer_lindquist(r_obs, theta_deg, spin, mass_kg, fov_deg, W, H, C.byref(cam)),
              "mgk_camera_boyer_lindquist")
    cam.camera_id = camera_id
    return cam


def camera_pinhole(position, orientation, fov_deg: float, W: int, H: int, camera_id: int = 0) -> abi.Camera:
    cam = abi.Camera()
    for i in range(3):
        cam.position[i] = position[i]
    for i in range(9):
        cam.orientation[i] = orientation[i]
    cam.sensor_width_mm = 24.0
    cam.focal_length_mm = 12.0 / math.tan(math.radians(fov_deg) / 2)
    cam.x_resolution, cam.y_resolution, cam.camera_id = W, H, camera_id
    return cam


def scene(**kw) -> abi.Scene:
    """A scene; `boxes` is a list of (min xyz, max xyz, matter) in SI metres."""
    s = abi.Scene(celestial_radius_m=1000.0, flags=0, disk_r_in=6.0, disk_r_out=20.0, disk_T_in=1e4,
                  sphere_radius=1.0, sphere_T=6500.0, sphere_checker=16, background_scale=0.0, background_T=6500.0,
                  star_scale=1.0, star_T=6500.0, star_radius_rad=1e-3, n_star=0, star_seed=0, n_box=0,
                  max_reflections=10, led_scale=7e5)
    for k, v in kw.items():
        if k == "sphere_center":
            for i in range(3):
                s.sphere_center[i] = v[i]
        elif k == "boxes":
            if len(v) > abi.MAX_BOX:
                raise ValueError(f"at most {abi.MAX_BOX} boxes")
            s.n_box = len(v)
            for b, (lo, hi, matter) in enumerate(v):
                for i in range(3):
                    s.box[b].min[i], s.box[b].max[i] = lo[i], hi[i]
                s.box[b].matter = matter
        else:
            setattr(s, k, v)
    return s


# ---------------------------------------------------------------------------
# The benchmark workloads.
# ---------------------------------------------------------------------------
def workload(name: str, W: int | None = None, H: int | None = None, spp: int | None = None,
             precision: int | None = None):
    """(integrator, scene, spectral, camera, spp) of a named workload.

    minkowski_sphere  : flat space, 256x256x4, camera at the origin looking at
                        +z (60 deg), emissive 16x16 checkered sphere r=1 at
                        z=5, uniform background sphere r=1000, 1 hero lambda,
                        FP64, tol 1e-8, cap 4096.
    schwarzschild_disk: M=1, 512x512x16, r=30M, theta=85 deg, 40 deg; thin disk
                        6M-20M with T = 1e4 K (r/6)^-3/4; 1e5 seeded stars;
                        tol 1e-7, cap 1e4.  The disk crossing is located on
                        the geodesic (SCENE_DISK_ON_GEODESIC) in every disk
                        workload.
    kerr_disk_1080p   : a=0.9, 1920x1080x32, r=50M, theta=80 deg, 25 deg; disk
                        r_ISCO-30M, same T law; 4 stratified hero lambdas;
                        same stars; tol 1e-7, cap 1e4.
    kerr_disk_4k      : the learned-surrogate frame: a=0.99, 3840x2160x16,
                        otherwise as kerr_disk_1080p (camera r=50M, theta=80
                        deg, 25 deg; disk r_ISCO(0.99)-30M).
    box_led           : the reference's default stage: Kerr a=0.999, M=6.73e26
                        kg, a diffuse 40x20x40 m box with a 20x20 m LED panel
                        at its ceiling, camera at (2,0,9) m, 2195x1860x8, tol
                        1e-3 / 1e-10, step limit 512, 10 reflections.
    box_mesh          : box_led plus triangle meshes (workload_meshes): two
                        diffuse icospheres (20,480 triangles each), a diffuse
                        tessellated plate (8,192) and an illuminant icosphere,
                        traced with Hermite chords.
    """
    if name == "minkowski_sphere":
        W, H, spp = W or 256, H or 256, spp or 4
        integ = integrator(metric=MINKOWSKI, precision=PREC_FP64 if precision is None else precision, spin=0.0,
                           abs_tol=1e-8, rel_tol=1e-8, step_limit=4096)
        # rotation by 180 deg about y: camera -z looks down world +z
        cam = camera_pinhole([0.0, 0.0, 0.0], [-1, 0, 0, 0, 1, 0, 0, 0, -1], 60.0, W, H)
        sc = scene(celestial_radius_m=1000.0, flags=SCENE_SPHERE | SCENE_BACKGROUND, sphere_center=(0.0, 0.0, 5.0),
                   sphere_radius=1.0, sphere_T=6500.0, sphere_checker=16, background_scale=0.1, background_T=6500.0)
        spec = spectral(n_hero=1)
        return integ, sc, spec, cam, spp
    if name == "schwarzschild_disk":
        W, H, spp = W or 512, H or 512, spp or 16
        integ = integrator(metric=SCHWARZSCHILD, precision=PREC_MIXED if precision is None else precision, spin=0.0)
        cam = camera_bl(30.0, 85.0, 0.0, 40.0, W, H)
        sc = scene(celestial_radius_m=1000.0 * length_scale(), flags=SCENE_DISK | SCENE_DISK_ON_GEODESIC | SCENE_STARS, disk_r_in=6.0,
                   disk_r_out=20.0, disk_T_in=1e4, n_star=100000, star_seed=0x5EED, star_scale=1.0,
                   star_radius_rad=1e-3, star_T=6500.0)
        spec = spectral(n_hero=1)
        return integ, sc, spec, cam, spp
    if name in ("kerr_disk_1080p", "kerr_disk_4k"):
        big = name == "kerr_disk_4k"
        W, H, spp = W or (3840 if big else 1920), H or (2160 if big else 1080), spp or (16 if big else 32)
        a = 0.99 if big else 0.9
        integ = integrator(metric=KERR, precision=PREC_MIXED if precision is None else precision, spin=a)
        cam = camera_bl(50.0, 80.0, a, 25.0, W, H)
        sc = scene(celestial_radius_m=1000.0 * length_scale(), flags=SCENE_DISK | SCENE_DISK_ON_GEODESIC | SCENE_STARS,
                   disk_r_in=r_isco(a),
                   disk_r_out=30.0, disk_T_in=1e4, n_star=100000, star_seed=0x5EED, star_scale=1.0,
                   star_radius_rad=1e-3, star_T=6500.0)
        spec = spectral(n_hero=4)
        return integ, sc, spec, cam, spp
    if name in ("box_led", "box_mesh"):
        # the reference's default stage (stage.h:19-22, relativity_settings.h:11-25,
        # optical_system_settings.h:20-34, render_settings.h:16-17): a Kerr hole of
        # 6.73e26 kg (M = 0.5 m), a = 0.999, inside a 40 x 20 x 40 m diffuse box
        # with a 20 x 20 m LED panel at its ceiling; camera at (2, 0, 9) m
        W, H, spp = W or 2195, H or 1860, spp or 8
        integ = integrator(metric=KERR, precision=PREC_MIXED if precision is None else precision, spin=0.999,
                           abs_tol=1e-3, rel_tol=1e-10, step_limit=512)
        cam = abi.Camera()
        cam.position[0], cam.position[1], cam.position[2] = 2.0, 0.0, 9.0
        for i, v in enumerate((1, 0, 0, 0, 1, 0, 0, 0, 1)):
            cam.orientation[i] = v
        cam.focal_length_mm, cam.sensor_width_mm = 12.0, 24.0
        cam.x_resolution, cam.y_resolution = W, H
        sc = scene(celestial_radius_m=1000.0, flags=SCENE_BOXES, max_reflections=10, led_scale=7e5,
                   boxes=[((-20.0, -10.0, -20.0), (20.0, 10.0, 20.0), MATTER_DIFFUSE),
                          ((-10.0, 9.5, -10.0), (10.0, 10.5, 10.0), MATTER_ILLUMINANT)])
        spec = spectral(n_hero=4)
        if name == "box_mesh":
            sc.flags |= SCENE_MESHES
            sc.chord = CHORD_HERMITE
        return integ, sc, spec, cam, spp
    raise ValueError(f"unknown workload {name}")


def workload_meshes(name: str = "box_mesh"):
    """(meshes, instances) of a workload with SCENE_MESHES (Context.set_meshes)."""
    if name != "box_mesh":
        raise ValueError(f"workload {name} has no meshes")
    meshes = [mesh.icosphere(5), mesh.grid_plane(1.0, 64)]
    instances = [(0, MATTER_DIFFUSE, mesh.affine(1.5, (-4.0, -2.0, -3.0))),
                 (0, MATTER_DIFFUSE, mesh.affine(2.0, (4.5, 1.5, -7.0))),
                 (1, MATTER_DIFFUSE, mesh.affine((8.0, 1.0, 8.0), (0.0, -6.0, -5.0), rotate_y_deg=30.0)),
                 (0, MATTER_ILLUMINANT, mesh.affine(0.8, (0.0, 4.0, -4.0)))]
    return meshes, instances


def batch_workload(seed: int = 4, n_cam: int = 64, resolutions=(128, 256, 512, 1024), spp: int = 8,
                   precision: int = PREC_MIXED, fov_deg: float = 30.0):
    """The batched uneven workload: n_cam seeded Boyer-Lindquist cameras with
    spin a ~ U[0, 0.99], inclination ~ U[10, 89] deg, r_obs ~ U[20, 100] M and
    a square resolution drawn from `resolutions`, spp samples each; a thin disk
    from each camera's ISCO to 30 M and the 1e5-star field.  Returns
    (integrator, scene, spectral, batch cameras, spp)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    cams = (abi.BatchCamera * n_cam)()
    for k in range(n_cam):
        a = float(rng.uniform(0.0, 0.99))
        th = float(rng.uniform(10.0, 89.0))
        r = float(rng.uniform(20.0, 100.0))
        res = int(rng.choice(resolutions))
        cams[k].camera = camera_bl(r, th, a, fov_deg, res, res, camera_id=k)
        cams[k].spin = a
    integ = integrator(metric=KERR, precision=precision, spin=0.0)
    sc = scene(celestial_radius_m=1000.0 * length_scale(), flags=SCENE_DISK | SCENE_DISK_ON_GEODESIC | SCENE_STARS, disk_r_in=0.0,
               disk_r_out=30.0,
               disk_T_in=1e4, n_star=100000, star_seed=0x5EED, star_scale=1.0, star_radius_rad=1e-3, star_T=6500.0)
    return integ, sc, spectral(n_hero=4), cams, spp


class Context:
    """One mgk_context on a device; device buffers are torch tensors."""

    def __init__(self, device: int = 0):
        self.lib = abi.load()
        h = C.c_void_p()
        abi.check(self.lib.mgk_context_create(device, C.byref(h)), "mgk_context_create")
        self.h = h
        self.device = device

    def close(self):
        if self.h:
            abi.check(self.lib.mgk_context_destroy(self.h), "mgk_context_destroy")
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _stream():
        import torch
        return C.c_void_p(torch.cuda.current_stream().cuda_stream)

    def set_scene(self, sc):
        abi.check(self.lib.mgk_set_scene(self.h, C.byref(sc)), "mgk_set_scene")

    def set_meshes(self, meshes, instances):
        """Builds and uploads the two-level hierarchy (mgk_set_meshes): meshes
        = [(vertices (n, 3), indices (m, 3))], instances = [(mesh, matter,
        3 x 4 object -> world)].  Empty lists clear it."""
        m, inst, _keep = mesh.to_abi(meshes, instances)
        abi.check(self.lib.mgk_set_meshes(self.h, m, len(meshes), inst, len(instances)), "mgk_set_meshes")

    def generate_rays(self, integ, cam, seed: int, sample_begin: int, spp: int):
        import torch
        n = cam.x_resolution * cam.y_resolution * spp
        dt = torch.float32 if integ.precision == PREC_FP32_REFERENCE else torch.float64
        rays = torch.empty((10, n), dtype=dt, device=f"cuda:{self.device}")
        abi.check(self.lib.mgk_generate_rays(self.h, C.byref(integ), C.byref(cam), seed, sample_begin, spp,
                                             C.c_void_p(rays.data_ptr()), n, self._stream()), "mgk_generate_rays")
        return rays

    def trace(self, integ, sc, cam_pos, rays, n: int | None = None, queue=None, want_steps: bool = True):
        import torch
        stride = rays.shape[1]
        n = stride if n is None else n
        hits = torch.zeros((stride, 4), dtype=torch.int32, device=rays.device)
        steps = torch.zeros(stride, dtype=torch.int32, device=rays.device) if want_steps else None
        cp = (C.c_float * 3)(*cam_pos)
        abi.check(self.lib.mgk_trace_rays(self.h, C.byref(integ), C.byref(sc), cp, C.c_void_p(rays.data_ptr()),
                                          stride, n, C.c_void_p(queue.data_ptr() if queue is not None else None),
                                          C.c_void_p(hits.data_ptr()),
                                          C.c_void_p(steps.data_ptr() if steps is not None else None),
                                          self._stream()), "mgk_trace_rays")
        return hits, steps

    def trace_final(self, integ, sc, cam_pos, rays, n: int | None = None, queue=None):
        """Like trace(), returning (hits, final states (8, stride) float64)."""
        import torch
        stride = rays.shape[1]
        n = stride if n is None else n
        hits = torch.zeros((stride, 4), dtype=torch.int32, device=rays.device)
        final = torch.zeros((8, stride), dtype=torch.float64, device=rays.device)
        cp = (C.c_float * 3)(*cam_pos)
        abi.check(self.lib.mgk_trace_rays_final(self.h, C.byref(integ), C.byref(sc), cp, C.c_void_p(rays.data_ptr()),
                                                stride, n, C.c_void_p(queue.data_ptr() if queue is not None else None),
                                                C.c_void_p(hits.data_ptr()), C.c_void_p(final.data_ptr()),
                                                self._stream()), "mgk_trace_rays_final")
        return hits, final

    def trace_paths(self, integ, sc, cam, seed: int, rays, spp: int, sample_begin: int = 0, want_final: bool = False):
        """Whole paths of a camera's rays with reflections off diffuse boxes
        (mgk_trace_paths); `rays` is overwritten for the reflected paths.
        Returns (hits, steps), and the terminal states (8, stride) with want_final."""
        import torch
        stride = rays.shape[1]
        hits = torch.zeros((stride, 4), dtype=torch.int32, device=rays.device)
        steps = torch.zeros(stride, dtype=torch.int32, device=rays.device)
        final = torch.zeros((8, stride), dtype=torch.float64, device=rays.device) if want_final else None
        abi.check(self.lib.mgk_trace_paths(self.h, C.byref(integ), C.byref(sc), C.byref(cam), seed, sample_begin, spp,
                                           C.c_void_p(rays.data_ptr()), stride, C.c_void_p(hits.data_ptr()),
                                           C.c_void_p(steps.data_ptr()),
                                           C.c_void_p(final.data_ptr() if want_final else None), self._stream()),
                  "mgk_trace_paths")
        return (hits, steps, final) if want_final else (hits, steps)

    def generate_rays_batch(self, integ, cams, seed: int, spp: int):
        """Rays of every batch camera: ((11, n) float64 tensor, offsets (n_cam + 1))."""
        import numpy as np
        import torch
        n = sum(c.camera.x_resolution * c.camera.y_resolution for c in cams) * spp
        rays = torch.empty((11, n), dtype=torch.float64, device=f"cuda:{self.device}")
        off = (C.c_uint64 * (len(cams) + 1))()
        abi.check(self.lib.mgk_generate_rays_batch(self.h, C.byref(integ), cams, len(cams), seed, spp,
                                                   C.c_void_p(rays.data_ptr()), n, off, self._stream()),
                  "mgk_generate_rays_batch")
        return rays, np.array(off[:], dtype=np.uint64)

    def trace_batch(self, integ, sc, rays, n: int | None = None, queue=None, want_steps: bool = True):
        import torch
        stride = rays.shape[1]
        n = stride if n is None else n
        hits = torch.zeros((stride, 4), dtype=torch.int32, device=rays.device)
        steps = torch.zeros(stride, dtype=torch.int32, device=rays.device) if want_steps else None
        abi.check(self.lib.mgk_trace_rays_batch(self.h, C.byref(integ), C.byref(sc), C.c_void_p(rays.data_ptr()), stride, n,
                                                C.c_void_p(queue.data_ptr() if queue is not None else None),
                                                C.c_void_p(hits.data_ptr()),
                                                C.c_void_p(steps.data_ptr() if steps is not None else None),
                                                self._stream()), "mgk_trace_rays_batch")
        return hits, steps

    def shade(self, sc, spec, cam, seed: int, sample_begin: int, spp: int, hits):
        import torch
        xyz = torch.zeros((cam.y_resolution, cam.x_resolution, 3), dtype=torch.float32, device=hits.device)
        abi.check(self.lib.mgk_shade(self.h, C.byref(sc), C.byref(spec), C.byref(cam), seed, sample_begin, spp,
                                     C.c_void_p(hits.data_ptr()), C.c_void_p(xyz.data_ptr()), self._stream()),
                  "mgk_shade")
        return xyz

    def render_frame(self, integ, sc, spec, cam, seed: int, spp: int, out=None):
        import torch
        if out is None:
            out = torch.zeros((cam.y_resolution, cam.x_resolution, 3), dtype=torch.float32,
                              device=f"cuda:{self.device}")
        abi.check(self.lib.mgk_render_frame(self.h, C.byref(integ), C.byref(sc), C.byref(spec), C.byref(cam), seed,
                                            spp, C.c_void_p(out.data_ptr()), self._stream()), "mgk_render_frame")
        return out

    def stats(self) -> abi.TraceStats:
        st = abi.TraceStats()
        abi.check(self.lib.mgk_last_trace_stats(self.h, C.byref(st)), "mgk_last_trace_stats")
        return st


class Manager:
    """A render manager (include/magik_manager.h): the command queue system
    and the worker thread that renders the stage progressively."""

    def __init__(self, device: int = 0, display=abi.MAGIK_DISPLAY_SWAPCHAIN, n_chunk: int = 4096, drop_overflow=False):
        self.lib = abi.load()
        d = abi.ManagerDescriptor(display_type=display, user_device_id=device, user_stream=None,
                                  cqs_n_reserved_chunk=n_chunk, cqs_drop_overflow=drop_overflow)
        h = C.c_void_p()
        abi.check(self.lib.magik_render_manager_create(C.byref(h), d), "magik_render_manager_create")
        self.h = h
        self._last_stats = abi.RendererStatistics()
        self._last_histories = {q: None for q in abi.QUEUE_HISTORIES}
        fb = C.c_void_p()
        abi.check(self.lib.magik_configure_framebuffer_collection(C.byref(fb), abi.FB_CONFIG_HOST),
                  "magik_configure_framebuffer_collection")
        self.fb = fb

    def close(self):
        if self.h:
            self.lib.magik_destroy_framebuffer_collection(self.fb)
            self.lib.magik_render_manager_destroy(self.h)
            self.h = None

    @staticmethod
    def command(name: str, *fields):
        """A command as 4-byte words: its type, then the fields (float or int)."""
        import struct
        words = [abi.CMD[name]]
        for f in fields:
            words.append(struct.unpack("<I", struct.pack("<f", f))[0] if isinstance(f, float) else int(f))
        return (C.c_uint32 * len(words))(*words)

    @staticmethod
    def aov_command(name: str, aov: str, channels: int = 3):
        c = abi.AovCommand(embedded_type=abi.CMD[name], length_of_name=len(aov), name=aov.encode(), channels=channels)
        return c

    def push(self, cmd) -> int:
        return self.lib.magik_cqs_push_command(self.h, C.cast(C.pointer(cmd), C.c_void_p))

    def dispatch(self) -> bool:
        ok = C.c_bool(False)
        abi.check(self.lib.magik_cqs_dispatch_command_buffer(self.h, C.byref(ok)), "magik_cqs_dispatch_command_buffer")
        return ok.value

    def error(self) -> int:
        return self.lib.magik_render_manager_fetch_error(self.h)

    def stats(self):
        """(statistics, whether they are new).  The library fills the struct only
        when the worker has published statistics since the last fetch; the last
        statistics fetched are returned otherwise.  The queue-count histories of
        a fresh fetch are copied into queue_histories() and their host memory is
        released at once, so the returned struct's history pointers are NULL."""
        import numpy as np
        st, new = abi.RendererStatistics(), C.c_bool(False)
        abi.check(self.lib.magik_fetch_renderer_statistics(self.h, C.byref(new), C.byref(st)), "magik_fetch_renderer_statistics")
        if new.value:
            n = st.latest_wavefront_history_length
            hist = {}
            for q in abi.QUEUE_HISTORIES:
                ptr = getattr(st, f"h_latest_{q}_queue_count_history")
                hist[q] = (np.ctypeslib.as_array(C.cast(ptr, C.POINTER(C.c_uint32)), shape=(n,)).copy() if n and ptr
                           else np.zeros(0, np.uint32))
            owned = abi.RendererStatistics.from_buffer_copy(st)
            abi.check(self.lib.magik_destroy_renderer_statistics(C.byref(owned)), "magik_destroy_renderer_statistics")
            for q in abi.QUEUE_HISTORIES:
                setattr(st, f"h_latest_{q}_queue_count_history", None)
            self._last_stats, self._last_histories = st, hist
        return self._last_stats, new.value

    def queue_histories(self) -> dict:
        """The queue counts per wavefront iteration of the latest sample, as of
        the last fresh stats() fetch: {queue name: uint32 array}."""
        return dict(self._last_histories)

    def fetch(self) -> tuple[int, bool]:
        new = C.c_bool(False)
        r = self.lib.magik_framebuffer_collection_fetch(self.h, C.byref(new), self.fb)
        return r, new.value

    def extract(self, aov: str = "test"):
        """(H, W, channels) float32 copy of an AOV of the last fetch, or None."""
        import numpy as np
        c = abi.HostContainer()
        abi.check(self.lib.magik_extract_config_host(self.h, C.byref(c), self.fb, aov.encode()), "magik_extract_config_host")
        if not c.h_data:
            return None
        n = c.x_resolution * c.y_resolution * c.channels
        return np.ctypeslib.as_array(c.h_data, shape=(n,)).copy().reshape(c.y_resolution, c.x_resolution, c.channels)


def hits_to_numpy(hits):
    """(n, 4) int32 device tensor -> structured numpy array cls/g/a0/a1."""
    import numpy as np
    h = hits.cpu().numpy()
    out = np.zeros(h.shape[0], dtype=[("cls", np.uint32), ("g", np.float32), ("a0", np.float32), ("a1", np.float32)])
    out["cls"] = h[:, 0].view(np.uint32)
    out["g"] = h[:, 1].view(np.float32)
    out["a0"] = h[:, 2].view(np.float32)
    out["a1"] = h[:, 3].view(np.float32)
    return out


def color_weights(spec):
    import numpy as np
    W = np.zeros((spec.n_bin, 3), dtype=np.float32)
    M = np.zeros(9, dtype=np.float32)
    abi.check(abi.load().mgk_color_weights(C.byref(spec), W.ctypes.data_as(C.POINTER(C.c_float)),
                                           M.ctypes.data_as(C.POINTER(C.c_float))), "mgk_color_weights")
    return W, M.reshape(3, 3)


# Colour pipeline enums (magik.h:330-383)
SCENE_CIE_XYZ, SCENE_CIE_RGB, SCENE_REC709, SCENE_REC2020, SCENE_ACES_CG, SCENE_ADOBE_RGB = range(6)
DISPLAY_SRGB, DISPLAY_REC709, DISPLAY_PQ, DISPLAY_HLG, DISPLAY_DCI_P3, DISPLAY_P3, DISPLAY_ADOBE_RGB = range(7)
VIEW_NONE, VIEW_REINHARD, VIEW_REINHARD_EXTENDED, VIEW_GT7, VIEW_AGX, VIEW_KHRONOS_PBR_NEUTRAL = range(6)
GT7_ICTCP, GT7_JZAZBZ = 0, 1
AGX_LOOK_NONE, AGX_LOOK_GOLDEN, AGX_LOOK_PUNCHY = 0, 1, 2


def display(**kw) -> abi.Display:
    """Display settings with the reference's colour defaults (color_settings.h:9-30)."""
    d = abi.Display(scene_colorspace=SCENE_REC709, view_transform=VIEW_GT7, display_colorspace=DISPLAY_REC709,
                    output_format=5, white_point_K=5600.0, tint=0.0, reinhard_white_luminance=4.0,
                    khronos_start_compression=0.8, khronos_desaturation=0.15, gt7_mode=GT7_JZAZBZ,
                    gt7_sdr_paper_white=250.0, gt7_reference_luminance=100.0, gt7_jzazbz_exponent_scale=1.7,
                    agx_look=AGX_LOOK_GOLDEN)
    for k, v in kw.items():
        setattr(d, k, v)
    return d


def display_transform(ctx, disp, xyz, out=None):
    """XYZ (..., 3) float32 on the GPU -> display-encoded colour (same shape).
    `out` may be `xyz` itself (in place) or a disjoint contiguous float32 tensor."""
    import torch
    if xyz.dtype != torch.float32 or not xyz.is_cuda or xyz.shape[-1] != 3:
        raise ValueError("display_transform: xyz must be a float32 CUDA tensor of shape (..., 3)")
    if out is xyz:
        if not xyz.is_contiguous():
            raise ValueError("display_transform: in-place output needs a contiguous tensor")
    else:
        xyz = xyz.contiguous()
        if out is None:
            out = torch.empty_like(xyz)
    if out.dtype != torch.float32 or not out.is_contiguous() or out.numel() != xyz.numel():
        raise ValueError("display_transform: out must be a contiguous float32 tensor with as many elements as xyz")
    n = xyz.numel() // 3
    abi.check(ctx.lib.mgk_display_transform(ctx.h, C.byref(disp), C.c_void_p(xyz.data_ptr()), n, C.c_void_p(out.data_ptr()),
                                            Context._stream()), "mgk_display_transform")
    return out


# ---------------------------------------------------------------------------
# Surrogate (learned path)
# ---------------------------------------------------------------------------
LAYER_SHAPES = [(16, 128), (128, 128), (128, 128), (128, 128), (128, 32)]
MINO_UNRESOLVED, MINO_CAPTURED, MINO_ESCAPES, MINO_DISK = range(4)


def mino_disk(integ, sc, rays, r_seed=None):
    """The analytic disk check of fallback.analytic, on the host (mgk_mino_disk):
    rays (n, 10) float64 initial conditions (host) -> (kind (n,) int32, r (n,))."""
    import numpy as np
    rays_soa = np.ascontiguousarray(np.asarray(rays, dtype=np.float64).T)
    n = rays_soa.shape[1]
    kind = np.zeros(n, np.int32)
    r = np.zeros(n, np.float64)
    seed = None if r_seed is None else np.ascontiguousarray(r_seed, dtype=np.float64)
    lib = abi.load()
    abi.check(lib.mgk_mino_disk(C.byref(integ), C.byref(sc), rays_soa.ctypes.data, n, n,
                                seed.ctypes.data if seed is not None else None, kind.ctypes.data, r.ctypes.data),
              "mgk_mino_disk")
    return kind, r


def surrogate_ranges(spin=(0.0, 0.99), r=(20.0, 100.0), theta=(10.0, 89.0), screen=15.0, disk_r_max=30.0):
    return abi.SurrogateRanges(spin_min=spin[0], spin_max=spin[1], r_min=r[0], r_max=r[1], theta_min_deg=theta[0],
                               theta_max_deg=theta[1], screen_max=screen, disk_r_max=disk_r_max)


class Surrogate:
    """The bf16 MLP surrogate on one device."""

    def __init__(self, device: int = 0, seed: int = 1, ranges=None):
        self.lib = abi.load()
        self.ranges = ranges or surrogate_ranges()
        h = C.c_void_p()
        abi.check(self.lib.mgk_surrogate_create(device, seed, C.byref(self.ranges), C.byref(h)), "mgk_surrogate_create")
        self.h = h

    def close(self):
        if self.h:
            abi.check(self.lib.mgk_surrogate_destroy(self.h), "mgk_surrogate_destroy")
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_weights(self):
        import numpy as np
        nw, nb = C.c_uint64(), C.c_uint64()
        self.lib.mgk_surrogate_n_params(C.byref(nw), C.byref(nb))
        w = np.zeros(nw.value, dtype=np.float32)
        b = np.zeros(nb.value, dtype=np.float32)
        abi.check(self.lib.mgk_surrogate_get_weights(self.h, w.ctypes.data_as(C.POINTER(C.c_float)),
                                                     b.ctypes.data_as(C.POINTER(C.c_float))), "get_weights")
        return split_params(w, b)

    def set_weights(self, W, b):
        import numpy as np
        w = np.concatenate([np.ascontiguousarray(x, dtype=np.float32).ravel() for x in W])
        bb = np.concatenate([np.ascontiguousarray(x, dtype=np.float32).ravel() for x in b])
        abi.check(self.lib.mgk_surrogate_set_weights(self.h, w.ctypes.data_as(C.POINTER(C.c_float)),
                                                     bb.ctypes.data_as(C.POINTER(C.c_float))), "set_weights")

    def forward(self, x):
        import torch
        x = x.contiguous()
        y = torch.zeros((x.shape[0], 8), dtype=torch.float32, device=x.device)
        abi.check(self.lib.mgk_surrogate_forward(self.h, C.c_void_p(x.data_ptr()), x.shape[0], C.c_void_p(y.data_ptr()),
                                                 Context._stream()), "mgk_surrogate_forward")
        return y

    def gradients(self, x, dy):
        import numpy as np
        import torch
        n = x.shape[0]
        y = torch.zeros((n, 32), dtype=torch.float32, device=x.device)
        nw, nb = C.c_uint64(), C.c_uint64()
        self.lib.mgk_surrogate_n_params(C.byref(nw), C.byref(nb))
        gw = np.zeros(nw.value, dtype=np.float32)
        gb = np.zeros(nb.value, dtype=np.float32)
        torch.cuda.synchronize()
        abi.check(self.lib.mgk_surrogate_gradients(self.h, C.c_void_p(x.contiguous().data_ptr()),
                                                   C.c_void_p(dy.contiguous().data_ptr()), n, C.c_void_p(y.data_ptr()),
                                                   gw.ctypes.data_as(C.POINTER(C.c_float)),
                                                   gb.ctypes.data_as(C.POINTER(C.c_float))), "mgk_surrogate_gradients")
        return y, split_params(gw, gb)

    def train_step(self, x, t, lr, w_r=10.0, w_d=1.0):
        import numpy as np
        import torch
        torch.cuda.synchronize()
        losses = np.zeros(4, dtype=np.float32)
        abi.check(self.lib.mgk_surrogate_train_step(self.h, C.c_void_p(x.data_ptr()), C.c_void_p(t.data_ptr()), x.shape[0],
                                                    lr, w_r, w_d, losses.ctypes.data_as(C.POINTER(C.c_float))),
                  "mgk_surrogate_train_step")
        return losses

    def train(self, x, t, batch=8192, epochs=4, lr=2e-3, lr_final=2e-5, w_r=10.0, w_d=1.0):
        import numpy as np
        import torch
        torch.cuda.synchronize()
        cfg = abi.TrainConfig(batch=batch, epochs=epochs, lr=lr, lr_final=lr_final, w_r=w_r, w_d=w_d)
        out = np.zeros((epochs, 4), dtype=np.float32)
        abi.check(self.lib.mgk_surrogate_train(self.h, C.c_void_p(x.data_ptr()), C.c_void_p(t.data_ptr()), x.shape[0],
                                               C.byref(cfg), out.ctypes.data_as(C.POINTER(C.c_float))),
                  "mgk_surrogate_train")
        return out

    def generate_data(self, ctx, integ, scene, seed: int, n: int):
        import torch
        rays = torch.zeros((11, n), dtype=torch.float64, device=f"cuda:{ctx.device}")
        x = torch.zeros((n, 16), dtype=torch.float32, device=rays.device)
        t = torch.zeros((n, 8), dtype=torch.float32, device=rays.device)
        hits = torch.zeros((n, 4), dtype=torch.int32, device=rays.device)
        abi.check(self.lib.mgk_surrogate_generate_data(ctx.h, self.h, C.byref(integ), C.byref(scene), seed, n,
                                                       C.c_void_p(rays.data_ptr()), C.c_void_p(x.data_ptr()),
                                                       C.c_void_p(t.data_ptr()), C.c_void_p(hits.data_ptr()),
                                                       Context._stream()), "mgk_surrogate_generate_data")
        return rays, x, t, hits


def split_params(w, b):
    W, B, ow, ob = [], [], 0, 0
    for k, n in LAYER_SHAPES:
        W.append(w[ow:ow + k * n].reshape(k, n))
        B.append(b[ob:ob + n])
        ow += k * n
        ob += n
    return W, B


def fallback(min_confidence=0.9, critical_margin=0.05, direction_margin=2e-3, disk_edge_margin=0.02, judge_stride=1024,
             trace_disk=False, analytic=False) -> abi.Fallback:
    return abi.Fallback(min_confidence=min_confidence, critical_margin=critical_margin,
                        direction_margin=direction_margin, disk_edge_margin=disk_edge_margin,
                        judge_stride=judge_stride, trace_disk=int(trace_disk), analytic=int(analytic))


def render_frame_surrogate(ctx: Context, sur: Surrogate, integ, sc, spec, cam, seed: int, spp: int, fb=None, out=None,
                           hits=None, stats: bool = True):
    """Whole frame through the surrogate with RKF45 fallback; returns (xyz, stats dict or None).
    `hits` (optional (W*H*spp, 4) int32 device tensor) receives every path's hit record."""
    import torch
    if out is None:
        out = torch.zeros((cam.y_resolution, cam.x_resolution, 3), dtype=torch.float32, device=f"cuda:{ctx.device}")
    st = abi.SurrogateStats()
    fb = fb or fallback()
    abi.check(ctx.lib.mgk_render_frame_surrogate(ctx.h, sur.h, C.byref(integ), C.byref(sc), C.byref(spec), C.byref(cam),
                                                 seed, spp, C.byref(fb), C.c_void_p(out.data_ptr()),
                                                 C.c_void_p(hits.data_ptr() if hits is not None else None),
                                                 C.byref(st) if stats else None, Context._stream()),
              "mgk_render_frame_surrogate")
    return out, (st.as_dict() if stats else None)
