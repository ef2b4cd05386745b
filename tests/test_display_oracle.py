"""Pins of oracle/display_oracle.py (the display side of the splat) against the
standards and closed forms it restates, so that a transcription slip in a table
or a curve fails here before it is used to check the GPU kernel."""
import numpy as np
import pytest

from oracle import display_oracle as D


def rgb_to_xyz_from_primaries(r, g, b, w):
    """SMPTE RP 177: the RGB -> XYZ matrix of primaries r, g, b and white w (xy)."""
    P = np.array([[p[0] / p[1], 1.0, (1 - p[0] - p[1]) / p[1]] for p in (r, g, b)]).T
    S = np.linalg.solve(P, D.xy_to_XYZ(*w))
    return P * S


PRIMARIES = {   # (table, R, G, B, white, tolerance)
    "rec709 scene": (D.SCENE[2][0], (0.64, 0.33), (0.30, 0.60), (0.15, 0.06), D.WP_D65, 3e-4),
    "rec2020 scene": (D.SCENE[3][0], (0.708, 0.292), (0.170, 0.797), (0.131, 0.046), D.WP_D65, 3e-4),
    "acescg": (D.SCENE[4][0], (0.713, 0.293), (0.165, 0.830), (0.128, 0.044), D.WP_ACES, 3e-4),
    "adobe scene": (D.SCENE[5][0], (0.64, 0.33), (0.21, 0.71), (0.15, 0.06), D.WP_D65, 3e-4),
    "srgb display": (D.DISPLAY[0][0], (0.64, 0.33), (0.30, 0.60), (0.15, 0.06), D.WP_D65, 3e-4),
    "pq display": (D.DISPLAY[2][0], (0.708, 0.292), (0.170, 0.797), (0.131, 0.046), D.WP_D65, 3e-4),
    "hlg display": (D.DISPLAY[3][0], (0.708, 0.292), (0.170, 0.797), (0.131, 0.046), D.WP_D65, 3e-4),
    "dci-p3 display": (D.DISPLAY[4][0], (0.680, 0.320), (0.265, 0.690), (0.150, 0.060), D.WP_DCI, 3e-4),
    "display p3": (D.DISPLAY[5][0], (0.680, 0.320), (0.265, 0.690), (0.150, 0.060), D.WP_D65, 3e-4),
    "adobe display": (D.DISPLAY[6][0], (0.64, 0.33), (0.21, 0.71), (0.15, 0.06), D.WP_D65, 3e-4),
}


@pytest.mark.parametrize("name", list(PRIMARIES))
def test_colour_space_tables_follow_from_their_primaries(name):
    table, r, g, b, w, tol = PRIMARIES[name]
    np.testing.assert_allclose(table, rgb_to_xyz_from_primaries(r, g, b, w), atol=tol)


@pytest.mark.parametrize("space", [*[("scene", k) for k in D.SCENE], *[("display", k) for k in D.DISPLAY]])
def test_from_xyz_is_the_inverse_of_to_xyz(space):
    kind, k = space
    to, frm, _ = (D.SCENE if kind == "scene" else D.DISPLAY)[k]
    np.testing.assert_allclose(frm @ to, np.eye(3), atol=3e-4)   # tables published to 4-9 digits


def test_cie_rgb_is_the_1931_matrix():
    # CIE 1931 RGB -> XYZ with equal-energy white: rows sum to 1 and Y uses (0.17697, 0.81240, 0.01063)
    to = D.SCENE[1][0]
    np.testing.assert_allclose(to.sum(axis=1), [1.0, 1.0, 1.0], atol=1e-3)
    np.testing.assert_allclose(to[1], [0.17697, 0.81240, 0.01063], atol=1e-3)


def test_bradford_maps_white_to_white():
    src, dst = D.xy_to_XYZ(0.35, 0.36), D.xy_to_XYZ(*D.WP_D65)
    np.testing.assert_allclose(D.bradford(src, dst) @ src, dst, rtol=1e-12)
    np.testing.assert_allclose(D.bradford(dst, dst), np.eye(3), atol=1e-12)


@pytest.mark.parametrize("T", [4500.0, 5600.0, 6504.0, 9000.0])
def test_blackbody_white_point_on_the_planckian_locus(T):
    # Kim et al. (2002) cubic approximation of the Planckian locus, 4000-25000 K
    xc = -3.0258469e9 / T ** 3 + 2.1070379e6 / T ** 2 + 0.2226347e3 / T + 0.240390
    yc = 3.0817580 * xc ** 3 - 5.87338670 * xc ** 2 + 3.75112997 * xc - 0.37001483
    v = D.white_point_xyz(T, 0.0)
    s = v.sum()
    assert abs(v[0] / s - xc) < 1.5e-3 and abs(v[1] / s - yc) < 1.5e-3


def test_tint_moves_perpendicular_to_the_locus_by_its_magnitude():
    def uv(v):
        d = v[0] + 15 * v[1] + 3 * v[2]
        return np.array([4 * v[0] / d, 6 * v[1] / d])
    p0, p1 = uv(D.white_point_xyz(5600.0, 0.0)), uv(D.white_point_xyz(5600.0, 0.01))
    tangent = uv(D.blackbody_xyz(5601.0)) - uv(D.blackbody_xyz(5600.0))
    assert np.linalg.norm(p1 - p0) == pytest.approx(0.01, rel=1e-6)
    assert abs(np.dot(p1 - p0, tangent)) < 1e-9 * np.linalg.norm(tangent)


def test_srgb_curve_is_continuous_and_normalised():
    lo = 12.92 * 0.0031308
    hi = 1.055 * 0.0031308 ** (1 / 2.4) - 0.055
    assert abs(lo - hi) < 1e-6
    for d in (0, 5):
        assert D.encode(d, 1.0) == pytest.approx(1.0, abs=1e-12)


def test_pq_known_values_and_round_trip():
    # SMPTE ST 2084: 100 cd/m^2 (0.01 of 10000) -> 0.50808; 10000 -> 1; 0 -> c1^m2 ~ 7.3e-7
    assert D.encode(2, 0.01) == pytest.approx(0.508078, abs=2e-6)
    assert D.encode(2, 1.0) == pytest.approx(1.0, abs=1e-12)
    x = np.geomspace(1e-6, 1.0, 50)
    np.testing.assert_allclose(D.decode_pq(D.encode(2, x)), x, rtol=1e-9)


def test_hlg_known_values():
    # ITU-R BT.2100 HLG OETF: E = 1/12 -> 0.5 (both branches), E = 1 -> 1
    assert D.encode(3, 1 / 12) == pytest.approx(0.5, abs=1e-12)
    a, b, c = 0.17883277, 0.28466892, 0.55991073
    assert a * np.log(12 / 12 - b) + c == pytest.approx(0.5, abs=1e-7)
    assert D.encode(3, 1.0) == pytest.approx(1.0, abs=1e-6)


def settings(**kw):
    s = dict(scene_colorspace=2, view_transform=0, display_colorspace=1, output_format=5, white_point_K=6504.0, tint=0.0,
             reinhard_white_luminance=4.0, khronos_start_compression=0.8, khronos_desaturation=0.15, gt7_mode=1,
             gt7_sdr_paper_white=250.0, gt7_reference_luminance=100.0, gt7_jzazbz_exponent_scale=1.7, agx_look=1)
    s.update(kw)
    return s


def test_reinhard_closed_forms():
    c = np.array([[1.0, 1.0, 1.0], [3.0, 0.5, 0.0]])
    s = settings(view_transform=1, display_colorspace=1)
    # with the scene white = display white and a 1/2.4 power display: (x/(1+x))^(1/2.4) per channel
    xyz = c @ D.SCENE[2][0].T
    wb = D.SCENE[2][1] @ D.bradford(D.white_point_xyz(6504.0, 0.0), D.xy_to_XYZ(*D.WP_D65))
    scene = xyz @ wb.T
    expect = np.power(np.maximum(scene / (1 + scene), 0), 1 / 2.4)
    np.testing.assert_allclose(D.display_transform(xyz, s), expect, rtol=1e-7, atol=1e-12)


def test_reinhard_extended_maps_white_luminance_to_one():
    W = 4.0
    L = np.array([W])
    k = (1 + L / (W * W)) / (1 + L)
    assert (L * k)[0] == pytest.approx(1.0)


def test_khronos_neutral_identity_below_the_knee_and_continuous_at_it():
    v = np.array([[0.2, 0.3, 0.5]])            # min >= 0.08: offset 0.04, peak after offset 0.46 < 0.76
    np.testing.assert_allclose(D.khronos_apply(v, 0.8, 0.15), v - 0.04, atol=1e-12)
    start = 0.8 - 0.04
    below = np.array([[0.5, 0.5, start + 0.04 - 1e-9]])
    above = np.array([[0.5, 0.5, start + 0.04 + 1e-9]])
    np.testing.assert_allclose(D.khronos_apply(below, 0.8, 0.15), D.khronos_apply(above, 0.8, 0.15), atol=1e-7)
    big = D.khronos_apply(np.array([[1e6, 1e6, 1e6]]), 0.8, 0.15)
    assert big.max() <= 1.0


@pytest.mark.parametrize("mode", [0, 1])
def test_gt7_perceptual_spaces_round_trip(mode):
    g = D.gt7_setup(mode, 250.0, 100.0, 1.7)
    rgb = np.random.default_rng(mode).uniform(0.01, 2.0, size=(64, 3))
    np.testing.assert_allclose(D.from_ucs(g, D.to_ucs(g, rgb)), rgb, rtol=2e-4, atol=1e-6)


def test_gt7_curve_joins_its_shoulder_smoothly():
    g = D.gt7_setup(1, 250.0, 100.0, 1.7)
    x0 = g["lin"] * g["peak"]
    sh = lambda x: g["kA"] + g["kB"] * np.exp(x * g["kC"])
    assert sh(x0) == pytest.approx(x0, rel=1e-12)                             # value (the linear part is x there)
    assert (sh(x0 + 1e-6) - sh(x0 - 1e-6)) / 2e-6 == pytest.approx(1.0, rel=1e-6)   # slope
    k = (g["lin"] - 1) / (0.25 - 1)
    assert sh(1e6) == pytest.approx(g["peak"] * (g["lin"] + k), rel=1e-9)    # asymptote kA, above the target
    assert g["target"] == pytest.approx(2.5)                                  # 250 / 100: min() clamps there
    out = D.gt7_apply(g, np.full((1, 3), 1e4))
    np.testing.assert_allclose(out, 1.0, rtol=1e-6)                           # x sdr scale 1 / target


def test_ictcp_matrices_are_inverse_pairs():
    np.testing.assert_allclose(D.ICTCP_INV @ D.ICTCP_OUT, np.eye(3), atol=2e-5)
    np.testing.assert_allclose(D.ICTCP_LMS_INV @ D.ICTCP_LMS, np.eye(3), atol=2e-5)
    np.testing.assert_allclose(D.JZ_LMS_INV @ D.JZ_LMS, np.eye(3), atol=2e-5)


def test_agx_contrast_polynomial_end_points():
    assert D.agx_contrast(0.0) == pytest.approx(0.002857)
    assert D.agx_contrast(1.0) == pytest.approx(0.982057, abs=1e-6)   # the sum of the coefficients
    x = np.linspace(0.05, 1, 96)
    assert np.all(np.diff(D.agx_contrast(x)) > 0)   # monotonic above its toe (the fit dips slightly below x ~ 0.03)


# ---------------------------------------------------------------------------
# Against the reference's own colorspaces.cuh (golden vectors written by
# tests/golden/make_golden.py from oracle/_ref)
# ---------------------------------------------------------------------------
import json  # noqa: E402
import os  # noqa: E402

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.json")))


def f32(bits):
    return np.asarray(bits, dtype=np.uint32).view(np.float32).astype(np.float64)


@pytest.mark.parametrize("entry", GOLD["scene_colorspaces"], ids=lambda e: f"scene{e['id']}")
def test_scene_tables_equal_the_reference(entry):
    to, frm, wp = D.SCENE[entry["id"]]
    np.testing.assert_array_equal(to.astype(np.float32).ravel(), f32(entry["to_xyz_bits"]).astype(np.float32))
    np.testing.assert_array_equal(frm.astype(np.float32).ravel(), f32(entry["from_xyz_bits"]).astype(np.float32))
    np.testing.assert_array_equal(np.float32(wp), f32(entry["white_xy_bits"]).astype(np.float32))


@pytest.mark.parametrize("entry", GOLD["display_colorspaces"], ids=lambda e: f"display{e['id']}")
def test_display_tables_and_encodes_equal_the_reference(entry):
    to, frm, wp = D.DISPLAY[entry["id"]][:3]
    np.testing.assert_array_equal(to.astype(np.float32).ravel(), f32(entry["to_xyz_bits"]).astype(np.float32))
    np.testing.assert_array_equal(frm.astype(np.float32).ravel(), f32(entry["from_xyz_bits"]).astype(np.float32))
    np.testing.assert_array_equal(np.float32(wp), f32(entry["white_xy_bits"]).astype(np.float32))
    x = np.asarray(entry["encode_in"], dtype=np.float32).astype(np.float64)
    ref = f32(entry["encode_out_bits"])
    got = D.encode(entry["id"], x)
    # The reference encodes in FP32 (powf / logf), the oracle in FP64.  PQ raises
    # the ratio (c1 + c2 y) / (1 + c3 y) to m2 = 78.84, which multiplies its FP32
    # rounding (6e-8) by ~80: 3e-5.  A constant wrong in its 4th digit moves PQ by 1e-3.
    rtol = 3e-5 if entry["id"] == 2 else 2e-6
    np.testing.assert_allclose(got, ref, rtol=rtol, atol=2e-7)
