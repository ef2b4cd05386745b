"""Closed-form pins of three oracle pieces that the end-to-end comparisons only
see composed with others:

* the sensor point and pinhole ray of a pixel (device_sensor_utilities.cuh:35-80,
  thin_lens.cuh:60-76): square pixels of pitch D / max(W, H) (D, the sensor
  "diagonal", spans the longer side), no half-pixel offset, jitter square of
  side D / sqrt(W^2 + H^2), and a horizontal field of view 2 atan(D / 2f) for
  landscape frames;
* the radiance -> sensor irradiance scale (device_color_utilities.cuh:22-29):
  E = L T (A / N) ISO 2 pi shutter / (360 fps), checked against a value worked
  out by hand;
* the step controller's error norm (relativity_utilities.cuh:15-29): the RMS of
  e_i / (atol + rtol max(|y0_i|, |y1_i|)) over the 8 components (Hairer,
  Norsett & Wanner, Solving ODEs I, II.4), on a vector whose norm is sqrt(3)
  exactly and which gives another value for each plausible slip (|y0| alone,
  a divisor of 7, a dropped component).
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

PRECS = [(O.F64, 1e-12), (O.F32, 2e-6)]


def pitch_and_side(D_mm, W, H):
    pitch = D_mm / max(W, H) / 1000.0
    side = D_mm / math.sqrt(W * W + H * H) / 1000.0
    return pitch, side


@pytest.mark.parametrize("prec,tol", PRECS)
@pytest.mark.parametrize("W,H", [(1920, 1080), (512, 512), (1080, 1920), (640, 1)])
def test_sensor_point_closed_form(prec, tol, W, H):
    D, f = 24.0, 20.0
    pitch, side = pitch_and_side(D, W, H)
    rng = np.random.default_rng(W * 7 + H)
    for _ in range(50):
        px, py = int(rng.integers(0, W)), int(rng.integers(0, H))
        ju, jv = float(rng.uniform()), float(rng.uniform())
        s, d = O.sensor_ray(prec, D, f, W, H, px, py, ju, jv)
        want = np.array([(px - W / 2) * pitch + (ju - 0.5) * side, (py - H / 2) * pitch + (jv - 0.5) * side, 0.0])
        assert np.allclose(s, want, rtol=0, atol=tol * D / 1000.0), (px, py, s, want)
        # the ray leaves the sensor point through the lens point (0, 0, -f)
        v = np.array([0.0, 0.0, -f / 1000.0]) - want
        assert np.allclose(d, v / np.linalg.norm(v), atol=4 * tol)


@pytest.mark.parametrize("prec,tol", PRECS)
def test_sensor_centre_axis_edges_and_field_of_view(prec, tol):
    D, f = 24.0, 20.0
    # centre pixel of an even frame with centred jitter: on the optical axis
    s, d = O.sensor_ray(prec, D, f, 1920, 1080, 960, 540, 0.5, 0.5)
    assert np.all(s == 0.0) and np.array_equal(d, [0.0, 0.0, -1.0])
    # pixel 0 (centred jitter): half the sensor width off axis, at the half FOV
    s, d = O.sensor_ray(prec, D, f, 1920, 1080, 0, 540, 0.5, 0.5)
    assert abs(s[0] + D / 2000.0) <= tol * D / 1000.0
    assert abs(math.atan2(d[0], -d[2]) - math.atan(D / 2 / f)) <= 4 * tol
    # portrait: the height spans D, the width W / H of it
    s, _ = O.sensor_ray(prec, D, f, 1080, 1920, 540, 0, 0.5, 0.5)
    assert abs(s[1] + D / 2000.0) <= tol * D / 1000.0
    s, _ = O.sensor_ray(prec, D, f, 1080, 1920, 0, 960, 0.5, 0.5)
    assert abs(s[0] + D * 1080 / 1920 / 2000.0) <= tol * D / 1000.0


@pytest.mark.parametrize("prec,tol", PRECS)
def test_sensor_jitter_square(prec, tol):
    D, W, H = 24.0, 1920, 1080
    pitch, side = pitch_and_side(D, W, H)
    a, _ = O.sensor_ray(prec, D, 20.0, W, H, 100, 200, 0.0, 0.0)
    b, _ = O.sensor_ray(prec, D, 20.0, W, H, 100, 200, 1.0, 1.0)
    assert np.allclose(b[:2] - a[:2], [side, side], rtol=4 * tol)
    # 1080p: the jitter covers 1920 / sqrt(1920^2 + 1080^2) = 87.2 % of the pitch
    assert abs(side / pitch - 0.87157553) < 1e-7


@pytest.mark.parametrize("prec,tol", PRECS)
def test_irradiance_scale_hand_value(prec, tol):
    # L = 2, throughput 3, 24 x 13.5 mm over 1920 x 1080 pixels (1.5625e-10 m^2
    # each), ISO 100, 24 fps, 180 degree shutter:
    # E = 6 * 1.5625e-10 * 100 * 2 pi * 180 / (360 * 24) = 3.90625e-9 pi
    # (pi as the reference holds it, a float constant: internal_math.h:10-16)
    E = O.radiance_to_irradiance(prec, 2.0, 3.0, 24.0, 13.5, 100.0, 24.0, 180.0, 1920, 1080)
    assert abs(E / (3.90625e-9 * float(np.float32(math.pi))) - 1.0) <= tol
    # linear in L, throughput, ISO and the shutter angle; inverse in fps and pixel count
    # (L x 2, ISO x 2, fps x 2, shutter / 2: a factor 1 / 2 overall)
    E2 = O.radiance_to_irradiance(prec, 4.0, 3.0, 24.0, 13.5, 200.0, 48.0, 45.0, 1920, 1080)
    assert abs(E2 / E - 0.5) <= 4 * tol
    E3 = O.radiance_to_irradiance(prec, 2.0, 3.0, 24.0, 13.5, 100.0, 24.0, 180.0, 960, 540)
    assert abs(E3 / E - 4.0) <= 4 * tol


@pytest.mark.parametrize("prec,tol", PRECS)
def test_rms_error_norm_hand_value(prec, tol):
    # atol = rtol = 1e-6; max(|y0|, |y1|) = (1, 1, 3, 3, 1, 1, 0, 0) with y0 the
    # larger on components 0, 2, 4 and y1 on 1, 3, 5, so the scales are
    # 1e-6 (2, 2, 4, 4, 2, 2, 1, 1); the ratios (1, -2, 1, 2, 3, 0, -1, 2) have
    # squares summing to 24, and sqrt(24 / 8) = sqrt(3).
    y0 = [1, 0, -3, 0, 1, 0, 0, 0]
    y1 = [0, -1, 0, 3, 0, 1, 0, 0]
    e = [2e-6, -4e-6, 4e-6, 8e-6, 6e-6, 0.0, -1e-6, 2e-6]
    r = O.rms_error(prec, e, y0, y1, 1e-6, 1e-6)
    assert abs(r - math.sqrt(3.0)) <= tol * math.sqrt(3.0), r
    # the slips this vector tells apart
    assert abs(math.sqrt(24 / 7) - math.sqrt(3)) > 0.1
    only_y0 = np.sqrt(np.mean((np.array(e) / (1e-6 * (1 + np.abs(y0)))) ** 2))
    assert abs(only_y0 - math.sqrt(3)) > 0.1
    # scale invariance: errors and tolerances scaled together leave the norm
    r2 = O.rms_error(prec, [10 * v for v in e], y0, y1, 1e-5, 1e-5)
    assert abs(r2 - r) <= 4 * tol * r
    # zero error, zero norm
    assert O.rms_error(prec, [0.0] * 8, y0, y1, 1e-6, 1e-6) == 0.0
