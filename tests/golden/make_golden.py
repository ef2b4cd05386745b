"""Writes tests/golden/ref_vectors.json from the reference's own headers.

Runs in the development container only, where /root/reference exists: it
builds oracle/_ref/libmagik_ref.so (internal_math.h, magik.h,
color_matching_functions.cuh and colorspaces.cuh compiled as they are, see
oracle/Makefile) and
records input/output vectors of the reference arithmetic that the oracle
restates.  The JSON holds data only (inputs, expected outputs).

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import ctypes as C
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402


def f32bits(x):
    return int(np.float32(x).view(np.uint32))


def main():
    O.build()
    R = O.ref()
    assert R is not None, "reference harness not built (needs /root/reference)"
    rng = np.random.default_rng(20261203)
    out = {"source": "reference headers via oracle/_ref (internal_math.h, magik.h, color_matching_functions.cuh, "
                     "colorspaces.cuh; magik.h's manager declarations)"}

    # RKF45 driver (internal_math.h:2160-2205) on the harmonic oscillator
    osc = []
    for h0, tol in [(0.1, 1e-6), (0.5, 1e-6), (3.0, 1e-6), (2.0, 1e-3), (0.01, 1e-9), (50.0, 1e-5)]:
        y = np.array([1.0, 0, 0, 0.0, 0.25, 0, 0, 0])
        hh, ht = C.c_float(h0), C.c_float(0)
        y1 = np.zeros(8)
        R.ref_rkf45_oscillator(O.dptr(y), C.byref(hh), C.byref(ht), 1000.0, 56, tol, tol, O.dptr(y1))
        osc.append({"y0": y.tolist(), "h0": h0, "tol": tol, "y1_bits": [f32bits(v) for v in y1],
                    "h_bits": f32bits(hh.value), "h_taken_bits": f32bits(ht.value)})
    out["rkf45_oscillator"] = osc

    # RKF45 driver on the Kerr equations (driver from the reference, RHS from the oracle)
    kerr = []
    for i in range(24):
        r = rng.uniform(3.0, 60.0)
        th = rng.uniform(0.2, 2.9)
        ph = rng.uniform(-np.pi, np.pi)
        a = float(rng.uniform(-0.99, 0.99))
        x, yv, z = r * np.sin(th) * np.cos(ph), r * np.cos(th), r * np.sin(th) * np.sin(ph)
        p = rng.normal(size=3)
        p = p / np.linalg.norm(p) * 2.0
        pt = -np.linalg.norm(p) * rng.uniform(0.9, 1.1)
        y = np.array([x, yv, z, 6.0e8, p[0], p[1], p[2], pt])
        y = y.astype(np.float32).astype(np.float64)
        h0 = float(rng.choice([0.5, 1.0, 4.0, 20.0]))
        tol = float(rng.choice([1e-3, 1e-5, 1e-6]))
        hh, ht = C.c_float(h0), C.c_float(0)
        y1 = np.zeros(8)
        R.ref_rkf45_kerr(O.dptr(y), a, C.byref(hh), C.byref(ht), 1000.0, 56, tol, 1e-10 if tol == 1e-3 else tol, O.dptr(y1))
        kerr.append({"y0": y.tolist(), "a": a, "h0": h0, "abs_tol": tol, "rel_tol": 1e-10 if tol == 1e-3 else tol,
                     "y1_bits": [f32bits(v) for v in y1], "h_bits": f32bits(hh.value), "h_taken_bits": f32bits(ht.value)})
    out["rkf45_kerr"] = kerr

    # Planck (internal_math.h:2511-2526), Gaussian CDF (:2019-2024)
    lam = np.concatenate([np.linspace(300, 900, 25), rng.uniform(360, 830, 15)])
    T = rng.choice([1000.0, 2600.0, 5600.0, 6500.0, 1e4, 3e4], size=lam.size)
    out["planck"] = [{"lam": float(l), "T": float(t), "f32_bits": f32bits(R.ref_planck_f32(l, t)),
                      "f64": R.ref_planck_f64(l, t)} for l, t in zip(lam, T)]
    cdf = []
    for _ in range(30):
        m, x, s = float(rng.uniform(360, 830)), float(rng.uniform(340, 850)), float(rng.uniform(2, 40))
        cdf.append({"mean": m, "x": x, "sigma": s, "f32_bits": f32bits(R.ref_gaussian_cdf_f32(m, x, s))})
    out["gaussian_cdf"] = cdf

    # mat3 rotate (internal_math.h:1784-1810)
    rot = []
    for _ in range(10):
        a, b, g = (float(v) for v in rng.uniform(-3.2, 3.2, size=3))
        m = (C.c_float * 9)()
        R.ref_rotate(a, b, g, m)
        rot.append({"angles": [a, b, g], "m_bits": [f32bits(v) for v in m]})
    out["rotate"] = rot

    # CIE table interpolation (color_matching_functions.cuh:35-44)
    cmf = []
    for l in np.concatenate([[355.0, 360.0, 555.0, 555.5, 830.0, 835.0], rng.uniform(360, 830, 30)]):
        v = (C.c_float * 3)()
        R.ref_sample_cmf(float(l), v)
        cmf.append({"lam": float(l), "xyz_bits": [f32bits(x) for x in v]})
    out["sample_cmf"] = cmf
    tab = np.zeros((471, 3), dtype=np.float32)
    R.ref_cmf_table(O.fptr(tab))
    out["cmf_table_sum_bits"] = [f32bits(v) for v in tab.sum(axis=0, dtype=np.float64).astype(np.float32)]

    # white LED spectrum of the illuminant matter (internal_math.h:2501-2509)
    lam_led = np.concatenate([np.linspace(300.0, 900.0, 61), np.random.default_rng(20261205).uniform(360, 830, 20)])
    out["white_led"] = [{"lam": float(l), "f32_bits": f32bits(R.ref_white_led(float(l)))} for l in lam_led]

    # colour spaces and display encodes (colorspaces.cuh:24-338)
    scene = []
    for sid in range(6):
        to, frm, wp = (C.c_float * 9)(), (C.c_float * 9)(), (C.c_float * 2)()
        assert R.ref_scene_colorspace(sid, to, frm, wp) == 1
        scene.append({"id": sid, "to_xyz_bits": [f32bits(v) for v in to], "from_xyz_bits": [f32bits(v) for v in frm],
                      "white_xy_bits": [f32bits(v) for v in wp]})
    out["scene_colorspaces"] = scene
    rng_c = np.random.default_rng(20261204)
    probe = np.array([-0.1, 0.0, 1e-7, 1e-4, 1e-3, 0.0031308, 0.004, 0.01, 0.05, 1 / 12, 0.1, 0.18, 0.5, 0.9, 1.0, 1.5, 10.0,
                      100.0], dtype=np.float32)
    rgb = np.concatenate([np.repeat(probe[:, None], 3, axis=1), rng_c.choice(probe, size=(30, 3)),
                          rng_c.uniform(0, 1.2, size=(30, 3))]).astype(np.float32)
    display = []
    for did in range(7):
        to, frm, wp, peak = (C.c_float * 9)(), (C.c_float * 9)(), (C.c_float * 2)(), C.c_float(0)
        enc = np.zeros_like(rgb)
        assert R.ref_display_colorspace(did, to, frm, wp, C.byref(peak), O.fptr(rgb), len(rgb), O.fptr(enc)) == 1
        display.append({"id": did, "to_xyz_bits": [f32bits(v) for v in to], "from_xyz_bits": [f32bits(v) for v in frm],
                        "white_xy_bits": [f32bits(v) for v in wp], "peak_nits": peak.value,
                        "encode_in": rgb.tolist(), "encode_out_bits": [[f32bits(v) for v in r] for r in enc]})
    out["display_colorspaces"] = display

    # the render-manager API: sizes, offsets and enum values (oracle/manager_layout.inc)
    man = (C.c_size_t * 256)()
    k = R.ref_manager_layout(man)
    out["manager_layout"] = list(man[:k])

    # struct layouts and result codes of magik.h
    lay = (C.c_size_t * 14)()
    R.ref_layout(lay)
    out["layout"] = list(lay)

    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_vectors.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
