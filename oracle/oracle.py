"""ctypes front of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loads ``oracle/liboracle.so`` (the C++ restatement of the reference's per-ray
arithmetic) and, where it was built, ``oracle/_ref/libmagik_ref.so`` (the
reference's own host-compilable headers).  Imported only by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from magik_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libmagik_ref.so")

F32, F64 = 0, 1

_D = C.POINTER(C.c_double)
_F = C.POINTER(C.c_float)
_U = C.POINTER(C.c_uint32)
_SIG = {
    "orc_philox4x32_10": (None, [_U, _U, _U]),
    "orc_path_uniforms": (None, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _F]),
    "orc_rhs": (None, [C.c_int, C.c_int, C.c_double, _D, _D]),
    "orc_metric": (None, [C.c_int, C.POINTER(abi.Integrator), _D, C.c_int, _D]),
    "orc_kerr_schild_r": (C.c_double, [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]),
    "orc_solve_step": (None, [C.c_int, C.POINTER(abi.Integrator), _D, _D, _D, _D, _U]),
    "orc_rkf45_oscillator": (None, [C.c_int, _D, _D, _D, C.c_double, C.c_uint32, C.c_double, C.c_double, _D, _U]),
    "orc_generate_rays": (C.c_int, [C.c_int, C.POINTER(abi.Integrator), C.POINTER(abi.Camera), C.c_uint64, _U, _U,
                                    C.c_uint64, _D]),
    "orc_trace_rays": (C.c_int, [C.c_int, C.POINTER(abi.Integrator), C.POINTER(abi.Scene), _F, _D, C.c_uint64,
                                 C.c_void_p, _U, _U, C.c_int]),
    "orc_bin_weights": (None, [C.c_uint32, C.c_float, C.c_float, _F, C.c_uint32, C.c_float, C.c_float, _F]),
    "orc_cmf_inverse_cdf": (None, [C.c_float, C.c_float, _F, C.c_uint32, C.c_float, C.c_float, _F]),
    "orc_planck": (C.c_double, [C.c_int, C.c_double, C.c_double]),
    "orc_gaussian_cdf": (C.c_double, [C.c_int, C.c_double, C.c_double, C.c_double]),
    "orc_shade": (C.c_int, [C.POINTER(abi.Scene), C.POINTER(abi.Spectral), C.POINTER(abi.Camera), C.c_uint64,
                            C.c_uint32, C.c_uint32, C.c_uint32, _U, C.c_void_p, _F, _F, _D]),
    "orc_host_trace_geodesic": (C.c_int, [C.POINTER(abi.Container), abi.InitialConditions, abi.MetricParameters,
                                          abi.IntegratorParameters]),
    "orc_star_catalogue": (None, [C.c_uint64, C.c_uint32, C.c_float, _F]),
    "orc_rotation": (None, [C.c_float, C.c_float, C.c_float, _F]),
    "orc_sample_cmf": (None, [_F, C.c_uint32, C.c_float, _F]),
    "orc_layout": (None, [C.POINTER(C.c_size_t)]),
    "orc_trace_paths": (C.c_int, [C.c_int, C.POINTER(abi.Integrator), C.POINTER(abi.Scene), C.POINTER(abi.Camera), C.c_uint64,
                                  C.c_uint32, _D, C.c_uint64, C.c_void_p, _U, _U, _D, C.c_int]),
    "orc_aabb_hit": (C.c_int, [C.c_int, _D, _D, C.c_double, _D, _D, _D, C.POINTER(C.c_int)]),
    "orc_offset_geodesic": (None, [_D, C.c_int, _D]),
    "orc_reflect_diffuse": (C.c_int, [C.c_int, C.POINTER(abi.Integrator), _D, C.c_double, C.c_int, C.c_double, C.c_double,
                                      _D, _D]),
    "orc_white_led": (C.c_double, [C.c_int, C.c_double]),
    "orc_manager_layout": (C.c_size_t, [C.POINTER(C.c_size_t)]),
    "orc_set_meshes": (C.c_int, [C.POINTER(abi.Mesh), C.c_uint32, C.POINTER(abi.Instance), C.c_uint32]),
    "orc_seg_tri": (C.c_int, [C.c_int, _D, _D, _D, _D, _D, C.c_double, _D]),
    "orc_mesh_chord": (C.c_int, [C.c_int, _D, _D, _D, _D, C.c_int, C.c_double, _D, _U]),
    "orc_rms_error": (C.c_double, [C.c_int, _D, _D, _D, C.c_double, C.c_double]),
    "orc_sensor_ray": (None, [C.c_int, C.c_double, C.c_double, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_double,
                              C.c_double, _D]),
    "orc_radiance_to_irradiance": (C.c_double, [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                                C.c_double, C.c_double, C.c_uint32, C.c_uint32]),
}

_REF_SIG = {
    "ref_rkf45_oscillator": (None, [_D, _F, _F, C.c_float, C.c_uint, C.c_float, C.c_float, _D]),
    "ref_rkf45_kerr": (None, [_D, C.c_float, _F, _F, C.c_float, C.c_uint, C.c_float, C.c_float, _D]),
    "ref_planck_f32": (C.c_float, [C.c_float, C.c_float]),
    "ref_planck_f64": (C.c_double, [C.c_double, C.c_double]),
    "ref_gaussian_cdf_f32": (C.c_float, [C.c_float, C.c_float, C.c_float]),
    "ref_white_led": (C.c_float, [C.c_float]),
    "ref_rotate": (None, [C.c_float, C.c_float, C.c_float, _F]),
    "ref_sample_cmf": (None, [C.c_float, _F]),
    "ref_cmf_table": (C.c_uint, [_F]),
    "ref_layout": (None, [C.POINTER(C.c_size_t)]),
    "ref_manager_layout": (C.c_size_t, [C.POINTER(C.c_size_t)]),
    "ref_scene_colorspace": (C.c_int, [C.c_uint, _F, _F, _F]),
    "ref_display_colorspace": (C.c_int, [C.c_uint, _F, _F, _F, _F, _F, C.c_uint, _F]),
}

_lib = None
_ref = None


def build() -> None:
    """Compile liboracle.so (and _ref when the reference tree is present)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if os.path.isdir("/root/reference/src"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        for n, (r, a) in _SIG.items():
            f = getattr(L, n)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def ref():
    """The reference-header harness, or None where it could not be built."""
    global _ref
    if _ref is None and os.path.exists(REF_LIB):
        L = C.CDLL(REF_LIB)
        for n, (r, a) in _REF_SIG.items():
            f = getattr(L, n)
            f.restype = r
            f.argtypes = a
        _ref = L
    return _ref


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_D)


def fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_F)


def uptr(a: np.ndarray):
    assert a.dtype == np.uint32 and a.flags.c_contiguous
    return a.ctypes.data_as(_U)


def rhs(prec: int, metric: int, arg: float, y) -> np.ndarray:
    y = np.ascontiguousarray(y, dtype=np.float64)
    out = np.zeros(8)
    lib().orc_rhs(prec, metric, arg, dptr(y), dptr(out))
    return out


def metric_tensor(prec: int, integ, q, inverse: bool) -> np.ndarray:
    """Full 4x4 matrix in index order (x, y, z, t)."""
    q = np.ascontiguousarray(q, dtype=np.float64)
    o = np.zeros(10)
    lib().orc_metric(prec, C.byref(integ), dptr(q), int(inverse), dptr(o))
    tt, tx, ty, tz, xx, xy, xz, yy, yz, zz = o
    return np.array([[xx, xy, xz, tx], [xy, yy, yz, ty], [xz, yz, zz, tz], [tx, ty, tz, tt]])


def solve_step(prec: int, integ, y, h: float):
    y = np.ascontiguousarray(y, dtype=np.float64)
    y1 = np.zeros(8)
    hh, ht = C.c_double(h), C.c_double(0)
    tr = C.c_uint32(0)
    lib().orc_solve_step(prec, C.byref(integ), dptr(y), C.byref(hh), C.byref(ht), dptr(y1), C.byref(tr))
    return y1, hh.value, ht.value, tr.value


def rms_error(prec: int, e, y0, y1, abs_tol: float, rel_tol: float) -> float:
    """The controller's error norm; states as (x, y, z, t, p_x, p_y, p_z, p_t)."""
    a = [np.ascontiguousarray(v, dtype=np.float64) for v in (e, y0, y1)]
    return lib().orc_rms_error(prec, dptr(a[0]), dptr(a[1]), dptr(a[2]), abs_tol, rel_tol)


def sensor_ray(prec: int, sensor_width_mm: float, focal_mm: float, W: int, H: int, px: int, py: int, ju: float, jv: float):
    """(sensor point in metres, unit pinhole direction) of pixel (px, py)."""
    o = np.zeros(6)
    lib().orc_sensor_ray(prec, sensor_width_mm, focal_mm, W, H, px, py, ju, jv, dptr(o))
    return o[:3], o[3:]


def radiance_to_irradiance(prec: int, L, throughput, sensor_w_mm, sensor_h_mm, iso, fps, shutter, W: int, H: int) -> float:
    return lib().orc_radiance_to_irradiance(prec, L, throughput, sensor_w_mm, sensor_h_mm, iso, fps, shutter, W, H)


def generate_rays(prec: int, integ, cam, seed: int, pixels, samples) -> np.ndarray:
    pixels = np.ascontiguousarray(pixels, dtype=np.uint32)
    samples = np.ascontiguousarray(samples, dtype=np.uint32)
    out = np.zeros((len(pixels), 10))
    abi.check(lib().orc_generate_rays(prec, C.byref(integ), C.byref(cam), seed, uptr(pixels), uptr(samples),
                                      len(pixels), dptr(out)), "orc_generate_rays")
    return out


HIT_DTYPE = np.dtype([("cls", np.uint32), ("g", np.float32), ("a0", np.float32), ("a1", np.float32)])


def trace_rays(prec: int, integ, scene, cam_pos, rays: np.ndarray, threads: int = 0):
    rays = np.ascontiguousarray(rays, dtype=np.float64)
    n = rays.shape[0]
    hits = np.zeros(n, dtype=HIT_DTYPE)
    steps = np.zeros(n, dtype=np.uint32)
    trials = np.zeros(n, dtype=np.uint32)
    cp = np.ascontiguousarray(cam_pos, dtype=np.float32)
    abi.check(lib().orc_trace_rays(prec, C.byref(integ), C.byref(scene), fptr(cp), dptr(rays), n,
                                   hits.ctypes.data_as(C.c_void_p), uptr(steps), uptr(trials), threads),
              "orc_trace_rays")
    return hits, steps, trials


def trace_paths(prec: int, integ, scene, cam, seed: int, sample_begin: int, rays: np.ndarray, threads: int = 0,
                want_at: bool = False):
    """Whole paths with reflections off diffuse boxes; returns (hits, steps,
    reflections) and, with want_at, the (n, 8) hit point and momentum of the
    paths that ended on a box."""
    rays = np.ascontiguousarray(rays, dtype=np.float64)
    n = rays.shape[0]
    hits = np.zeros(n, dtype=HIT_DTYPE)
    steps = np.zeros(n, dtype=np.uint32)
    refl = np.zeros(n, dtype=np.uint32)
    at = np.zeros((n, 8)) if want_at else None
    abi.check(lib().orc_trace_paths(prec, C.byref(integ), C.byref(scene), C.byref(cam), seed, sample_begin, dptr(rays), n,
                                    hits.ctypes.data_as(C.c_void_p), uptr(steps), uptr(refl),
                                    dptr(at) if want_at else None, threads), "orc_trace_paths")
    return (hits, steps, refl, at) if want_at else (hits, steps, refl)


def aabb_hit(prec: int, origin, direction, inv_chord: float, bmin, bmax):
    """(hit, t, normal code) of the slab test."""
    o, d = np.ascontiguousarray(origin, dtype=np.float64), np.ascontiguousarray(direction, dtype=np.float64)
    lo, hi = np.ascontiguousarray(bmin, dtype=np.float64), np.ascontiguousarray(bmax, dtype=np.float64)
    t, nc = C.c_double(0), C.c_int(0)
    ok = lib().orc_aabb_hit(prec, dptr(o), dptr(d), inv_chord, dptr(lo), dptr(hi), C.byref(t), C.byref(nc))
    return bool(ok), t.value, nc.value


def offset_geodesic(q, normal_code: int) -> np.ndarray:
    q = np.ascontiguousarray(q, dtype=np.float64)
    out = np.zeros(4)
    lib().orc_offset_geodesic(dptr(q), normal_code, dptr(out))
    return out


def reflect_diffuse(prec: int, integ, at, contraction: float, normal_code: int, u0: float, u1: float):
    """(new state (8), new contraction) of one reflection off diffuse matter."""
    at = np.ascontiguousarray(at, dtype=np.float64)
    out = np.zeros(8)
    c = C.c_double(0)
    abi.check(lib().orc_reflect_diffuse(prec, C.byref(integ), dptr(at), contraction, normal_code, u0, u1, dptr(out),
                                        C.byref(c)), "orc_reflect_diffuse")
    return out, c.value


def set_meshes(meshes, instances) -> None:
    """The oracle's meshes for scenes with SCENE_MESHES: the same inputs as
    magik_amd.Context.set_meshes (lists of (vertices, indices) and of
    (mesh, matter, 3 x 4 object -> world))."""
    from magik_amd import mesh as mesh_mod
    m, inst, keep = mesh_mod.to_abi(meshes, instances)
    if lib().orc_set_meshes(m, len(meshes), inst, len(instances)) != 0:
        raise ValueError("invalid meshes")
    set_meshes._keep = keep


def seg_tri(prec: int, o, d, v0, e1, e2, tmax: float = np.inf):
    """Moller-Trumbore hit parameter of o + t d on a triangle, or None."""
    arr = [np.ascontiguousarray(x, dtype=np.float64) for x in (o, d, v0, e1, e2)]
    t = np.zeros(1)
    ok = lib().orc_seg_tri(prec, *[dptr(x) for x in arr], float(tmax), dptr(t))
    return float(t[0]) if ok else None


def mesh_chord(prec: int, p0, p1, D0=(0, 0, 0), D1=(0, 0, 0), hermite: bool = False, s_max: float = 1.0):
    """The oracle's chord test against the meshes of set_meshes: (s, hit point,
    instance, triangle) or None."""
    arr = [np.ascontiguousarray(x, dtype=np.float64) for x in (p0, p1, D0, D1)]
    out = np.zeros(4)
    it = np.zeros(2, dtype=np.uint32)
    r = lib().orc_mesh_chord(prec, *[dptr(x) for x in arr], int(hermite), float(s_max), dptr(out), uptr(it))
    if r < 0:
        raise RuntimeError("no meshes set")
    return (out[0], out[1:].copy(), int(it[0]), int(it[1])) if r else None


def white_led(prec: int, lam_nm: float) -> float:
    return lib().orc_white_led(prec, lam_nm)


def shade(scene, spec, cam, seed: int, sample_begin: int, spp: int, pixel_ids, hits: np.ndarray,
          weights: np.ndarray) -> np.ndarray:
    pixel_ids = np.ascontiguousarray(pixel_ids, dtype=np.uint32)
    hits = np.ascontiguousarray(hits)
    assert hits.dtype == HIT_DTYPE and hits.size == spp * len(pixel_ids)
    w = np.ascontiguousarray(weights, dtype=np.float32)
    out = np.zeros((len(pixel_ids), 3))
    inv = None
    if spec.lambda_sampling == abi.LAMBDA_CMF:
        inv = cmf_inverse_cdf(spec.hero_min_nm, spec.hero_max_nm, cie_table())
    abi.check(lib().orc_shade(C.byref(scene), C.byref(spec), C.byref(cam), seed, sample_begin, spp, len(pixel_ids),
                              uptr(pixel_ids), hits.ctypes.data_as(C.c_void_p), fptr(w),
                              fptr(inv) if inv is not None else None, dptr(out)), "orc_shade")
    return out


def cmf_inverse_cdf(lo: float, hi: float, table: np.ndarray) -> np.ndarray:
    """The 256-entry inverse CDF of x+y+z (5% floor) over [lo, hi] nm, from a CIE table (n x 3, 360..830 nm)."""
    tab = np.ascontiguousarray(table, dtype=np.float32)
    out = np.zeros(256, dtype=np.float32)
    lib().orc_cmf_inverse_cdf(lo, hi, fptr(tab), tab.shape[0], 360.0, 830.0, fptr(out))
    return out


def bin_weights(n_bin: int, bmin: float, bmax: float, table: np.ndarray) -> np.ndarray:
    tab = np.ascontiguousarray(table, dtype=np.float32)
    W = np.zeros((n_bin, 3), dtype=np.float32)
    lib().orc_bin_weights(n_bin, bmin, bmax, fptr(tab), tab.shape[0], 360.0, 830.0, fptr(W))
    return W


def star_catalogue(seed: int, n: int, scale: float) -> np.ndarray:
    out = np.zeros((n, 4), dtype=np.float32)
    lib().orc_star_catalogue(seed, n, scale, fptr(out))
    return out


def cie_table() -> np.ndarray:
    """The CIE 1931 2-degree table the product ships (magik_amd/data)."""
    path = os.path.join(os.path.dirname(HERE), "magik_amd", "data", "cie1931_2deg_1nm.csv")
    rows = [l.strip().split(",") for l in open(path) if l.strip() and not l.startswith("#")]
    return np.array([[float(x) for x in r[1:]] for r in rows], dtype=np.float32)


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(uptr(c), uptr(k), uptr(o))
    return o


def path_uniforms(seed: int, cam: int, pixel: int, sample: int, block: int = 0) -> np.ndarray:
    u = np.zeros(4, dtype=np.float32)
    lib().orc_path_uniforms(seed, cam, pixel, sample, block, fptr(u))
    return u
