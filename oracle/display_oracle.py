"""NumPy (float64) restatement of the display side of the splat (TEST
INFRASTRUCTURE ONLY: tests/ use it to check mgk_display_transform; the product
never imports it).

Per pixel: linear XYZ -> white-balanced scene colour space -> view transform
-> display primaries -> display transfer function, following
  color_pipeline.cuh:21-48          (white balance, scene -> display matrix),
  host_color_utilities.cuh:29-73    (blackbody + tint white point, Bradford,
                                     convert_scene_colorspace),
  colorspaces.cuh:8-340             (matrices, white points, encodes),
  view_transform.cuh:10-577         (NONE, Reinhard, Khronos PBR Neutral,
                                     GT7, AgX),
  device_color_utilities.cuh:56-72  (the order of the steps).
The colour-space tables and curve constants are published data (see the
device header); test_display_oracle.py checks them against their defining
standards and closed forms.
"""
from __future__ import annotations

import os

import numpy as np

# ---- colour spaces: (to_XYZ, from_XYZ, white point xy) ------------------------
WP_E = (1.0 / 3.0, 1.0 / 3.0)
WP_DCI = (0.314, 0.351)
WP_D65 = (0.31272, 0.32903)
WP_ACES = (0.32168, 0.33767)


def _m(*v):
    return np.array(v, dtype=np.float64).reshape(3, 3)


SCENE = {
    0: (np.eye(3), np.eye(3), WP_E),
    1: (_m(0.49, 0.31, 0.2, 0.1769, 0.8124, 0.0107, 0.0, 0.0099, 0.9901),
        _m(2.36449012, -0.89655263, -0.46793749, -0.51493525, 1.42633279, 0.08860245, 0.00514883, -0.01426189, 1.00911305), WP_E),
    2: (_m(0.41238656, 0.35759149, 0.18045049, 0.21263682, 0.71518298, 0.0721802, 0.01933062, 0.11919716, 0.95037259),
        _m(3.24100326, -1.53739899, -0.49861587, -0.96922426, 1.87592999, 0.04155422, 0.05563942, -0.2040112, 1.05714897), WP_D65),
    3: (_m(6.36953507e-01, 1.44619185e-01, 1.68855854e-01, 2.62698339e-01, 6.78008766e-01, 5.92928953e-02, 4.99407097e-17,
           2.80731358e-02, 1.06082723e+00),
        _m(1.71666343, -0.35567332, -0.25336809, -0.66667384, 1.61645574, 0.0157683, 0.01764248, -0.04277698, 0.94224328), WP_D65),
    4: (_m(0.66245418, 0.13400421, 0.15618769, 0.27222872, 0.67408177, 0.05368952, -0.00557465, 0.00406073, 1.0103391),
        _m(1.64102338, -0.32480329, -0.2364247, -0.66366286, 1.61533159, 0.01675635, 0.01172189, -0.00828444, 0.98839486), WP_ACES),
    5: (_m(0.57667, 0.18556, 0.18823, 0.29734, 0.62736, 0.07529, 0.02703, 0.07069, 0.99134),
        _m(2.04159, -0.56501, -0.34473, -0.96924, 1.87597, 0.04156, 0.01344, -0.11836, 1.01517), WP_D65),
}

DISPLAY = {
    0: (_m(0.4124564, 0.3575761, 0.1804375, 0.2126729, 0.7151522, 0.0721750, 0.0193339, 0.1191920, 0.9503041),
        _m(3.2406255, -1.537208, -0.4986286, -0.9689307, 1.8757561, 0.0415175, 0.0557101, -0.2040211, 1.0569959), WP_D65),
    1: SCENE[2],
    2: (_m(0.63695804830129132, 0.14461690358620838, 0.16888097516417205, 0.26270021201126703, 0.67799807151887104,
           0.059301716469861945, 0.0, 0.028072693049087508, 1.0609850577107909),
        _m(1.7166511879712676, -0.35567078377639239, -0.2533662813736598, -0.66668435183248898, 1.616481236634939,
           0.015768545813911131, 0.017639857445310915, -0.042770613257808655, 0.94210312123547402), WP_D65),
    3: (_m(0.636958063, 0.144616902, 0.168880969, 0.262700200, 0.677998066, 0.059301715, 0.0, 0.028072693, 1.060985090),
        _m(1.716651200, -0.355670780, -0.253366292, -0.666684330, 1.616481180, 0.015768547, 0.017639857, -0.042770613, 0.942103148),
        WP_D65),
    4: (_m(0.44516982, 0.27713441, 0.17228267, 0.20949168, 0.72159525, 0.06891307, 0.0, 0.04706056, 0.90735539),
        _m(2.72539403, -1.01800301, -0.44016320, -0.79516803, 1.68973205, 0.02264719, 0.04124189, -0.08763902, 1.10092938), WP_DCI),
    5: (_m(0.486570954, 0.265667707, 0.198217288, 0.228974566, 0.691738546, 0.0792869106, 0.0, 0.045113381, 1.043944360),
        _m(2.493496890, -0.9313836100, -0.402710795, -0.829488993, 1.7626640800, 0.0236246865, 0.035845831, -0.0761723891,
           0.956884503), WP_D65),
    6: (_m(0.57666904, 0.18555824, 0.18822865, 0.29734498, 0.62736357, 0.07529146, 0.02703136, 0.07068885, 0.99133754),
        _m(2.04158790, -0.56500697, -0.34473135, -0.96924364, 1.87596750, 0.04155506, 0.01344428, -0.11836239, 1.01517499), WP_D65),
}


def xy_to_XYZ(x, y):
    return np.array([x / y, 1.0, (1.0 - x - y) / y])


BRADFORD = _m(0.8951, 0.2664, -0.1614, -0.7502, 1.7135, 0.0367, 0.0389, -0.0685, 1.0296)


def bradford(src, dst):
    """von Kries adaptation in the Bradford cone space (host_color_utilities.cuh:59-68)."""
    ls, ld = BRADFORD @ src, BRADFORD @ dst
    return np.linalg.inv(BRADFORD) @ np.diag(ld / ls) @ BRADFORD


def convert(frm, to):
    """convert_scene_colorspace (host_color_utilities.cuh:70-73)."""
    return to[1] @ bradford(xy_to_XYZ(*frm[2]), xy_to_XYZ(*to[2])) @ frm[0]


# ---- white point: blackbody + tint (host_color_utilities.cuh:29-57) ------------
_CIE = None


def cie_table():
    global _CIE
    if _CIE is None:
        path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "magik_amd", "data",
                            "cie1931_2deg_1nm.csv")
        _CIE = np.loadtxt(path, delimiter=",", comments="#")
    return _CIE


def blackbody_xyz(T):
    t = cie_table()
    lam = t[:, 0] * 1e-9
    h, c, k = 6.62607015e-34, 299792458.0, 1.380649e-23
    B = (2 * h * c * c / lam ** 5) / np.expm1(h * c / (lam * k * T))
    v = (t[:, 1:] * B[:, None]).sum(axis=0)
    return v / v[1]


def white_point_xyz(T, tint):
    a, b = blackbody_xyz(T), blackbody_xyz(T + 1.0)

    def uv(v):
        d = v[0] + 15 * v[1] + 3 * v[2]
        return np.array([4 * v[0] / d, 6 * v[1] / d])

    u0, u1 = uv(a), uv(b)
    n = (u1 - u0) / np.linalg.norm(u1 - u0)
    t = np.array([-n[1], n[0]])
    if t[1] < 0:
        t = -t
    up = u0 + t * tint
    den = 2 * up[0] - 8 * up[1] + 4
    return xy_to_XYZ(3 * up[0] / den, 2 * up[1] / den)


# ---- display transfer functions (colorspaces.cuh) ------------------------------
def encode(disp_id, L):
    L = np.asarray(L, dtype=np.float64)
    if disp_id in (0, 5):
        return np.where(L <= 0.0031308, 12.92 * L, 1.055 * np.power(np.maximum(L, 0.0031308), 1 / 2.4) - 0.055)
    if disp_id == 1:
        return np.power(np.fmax(L, 0.0), 1 / 2.4)
    if disp_id == 2:
        m1, m2 = 2610 / 16384, 2523 / 4096 * 128
        c1, c2, c3 = 3424 / 4096, 2413 / 4096 * 32, 2392 / 4096 * 32
        y = np.power(np.clip(L, 0.0, 1.0), m1)
        return np.power((c1 + c2 * y) / (1 + c3 * y), m2)
    if disp_id == 3:
        a, b, c = 0.17883277, 0.28466892, 0.55991073
        e = np.clip(L, 0.0, 1.0)
        return np.where(e <= 1 / 12, np.sqrt(3 * e), a * np.log(np.maximum(12 * e - b, 1e-300)) + c)
    if disp_id == 4:
        return np.power(np.fmax(L, 0.0), 1 / 2.6)
    return np.power(np.clip(L, 0.0, 1.0), 256 / 563)


def decode(disp_id, V):
    """Inverse of encode on its range (for comparisons in linear display units)."""
    V = np.asarray(V, dtype=np.float64)
    if disp_id in (0, 5):
        return np.where(V <= 0.04045, V / 12.92, np.power(np.maximum((V + 0.055) / 1.055, 0.0), 2.4))
    if disp_id == 1:
        return np.power(np.maximum(V, 0.0), 2.4)
    if disp_id == 2:
        return decode_pq(V)
    if disp_id == 3:
        a, b, c = 0.17883277, 0.28466892, 0.55991073
        return np.where(V <= 0.5, V * V / 3.0, (np.exp((V - c) / a) + b) / 12.0)
    if disp_id == 4:
        return np.power(np.maximum(V, 0.0), 2.6)
    return np.power(np.maximum(V, 0.0), 563 / 256)


def xyz_to_scene_matrix(s):
    """The white-balanced XYZ -> scene matrix of settings s."""
    sc = SCENE[s["scene_colorspace"]]
    return sc[1] @ bradford(white_point_xyz(s["white_point_K"], s["tint"]), xy_to_XYZ(*sc[2]))


def decode_pq(N):
    m1, m2 = 2610 / 16384, 2523 / 4096 * 128
    c1, c2, c3 = 3424 / 4096, 2413 / 4096 * 32, 2392 / 4096 * 32
    p = np.power(np.clip(N, 0, 1), 1 / m2)
    return np.power(np.maximum(p - c1, 0) / (c2 - c3 * p), 1 / m1)


# ---- GT7 (view_transform.cuh:115-461) ------------------------------------------
def smooth_step(x, e0, e1):
    t = (x - e0) / (e1 - e0)
    return np.where(x < e0, 0.0, np.where(x > e1, 1.0, t * t * (3 - 2 * t)))


def pq_to_fb(n, ref, s):
    n = np.clip(n, 0.0, 1.0)
    m1, m2, c1, c2, c3 = 0.1593017578125, 78.84375 * s, 0.8359375, 18.8515625, 18.6875
    p = np.power(n, 1 / m2)
    return np.power(np.maximum(p - c1, 0) / (c2 - c3 * p), 1 / m1) * 10000.0 / ref


def fb_to_pq(v, ref, s):
    m1, m2, c1, c2, c3 = 0.1593017578125, 78.84375 * s, 0.8359375, 18.8515625, 18.6875
    ym = np.power(v * ref / 10000.0, m1)
    return np.power(2.0, m2 * (np.log2(c1 + c2 * ym) - np.log2(1 + c3 * ym)))


ICTCP_LMS = np.array([[1688, 2146, 262], [683, 2951, 462], [99, 309, 3688]]) / 4096.0
ICTCP_OUT = np.array([[2048, 2048, 0], [6610, -13613, 7003], [17933, -17390, -543]]) / 4096.0
ICTCP_INV = np.array([[1, 0.00860904, 0.11103], [1, -0.00860904, -0.11103], [1, 0.560031, -0.320627]])
ICTCP_LMS_INV = np.array([[3.43661, -2.50645, 0.0698454], [-0.79133, 1.9836, -0.192271], [-0.0259499, -0.0989137, 1.12486]])
JZ_LMS = np.array([[0.530004, 0.355704, 0.086090], [0.289388, 0.525395, 0.157481], [0.091098, 0.147588, 0.734234]])
JZ_AB = np.array([[3.524000, -4.066708, 0.542708], [0.199076, 1.096799, -1.295875]])
JZ_INV = np.array([[1.386050432715393e-1, 5.804731615611869e-2], [-1.386050432715393e-1, -5.804731615611869e-2],
                   [-9.601924202631895e-2, -8.118918960560390e-1]])
JZ_LMS_INV = np.array([[2.990669, -2.049742, 0.088977], [-1.634525, 3.145627, -0.483037], [-0.042505, -0.377983, 1.448019]])
JZ_D0 = 1.6295499532821566e-11


def to_ucs(g, rgb):
    rgb = np.asarray(rgb, dtype=np.float64)
    if g["mode"] == 0:
        lms = fb_to_pq(rgb @ ICTCP_LMS.T, g["ref"], 1.0)
        return lms @ ICTCP_OUT.T
    lms = fb_to_pq(rgb @ JZ_LMS.T, g["ref"], g["jz"])
    iz = 0.5 * lms[..., 0] + 0.5 * lms[..., 1]
    jz = 0.44 * iz / (1 - 0.56 * iz) - JZ_D0
    ab = lms @ JZ_AB.T
    return np.concatenate([jz[..., None], ab], axis=-1)


def from_ucs(g, u):
    u = np.asarray(u, dtype=np.float64)
    if g["mode"] == 0:
        lms = pq_to_fb(u @ ICTCP_INV.T, g["ref"], 1.0)
        return np.fmax(lms @ ICTCP_LMS_INV.T, 0.0)
    jz = u[..., 0] + JZ_D0
    iz = jz / (0.44 + 0.56 * jz)
    lmsp = iz[..., None] + u[..., 1:] @ JZ_INV.T
    lms = pq_to_fb(lmsp, g["ref"], g["jz"])
    return lms @ JZ_LMS_INV.T


def gt7_setup(mode, paper, ref, jz):
    target = paper / ref
    alpha, mid, lin, toe = 0.25, 0.538, 0.444, 1.280
    k = (lin - 1) / (alpha - 1)
    g = dict(mode=mode, ref=ref, jz=jz, target=target, sdr=1 / target, peak=target, mid=mid, lin=lin, toe=toe,
             kA=target * lin + target * k, kB=-target * k * np.exp(lin / k), kC=-1 / (k * target),
             blend=0.6, f0=0.98, f1=1.16)
    g["target_ucs"] = to_ucs(g, np.full(3, target))[0]
    return g


def gt7_curve(g, x):
    w = smooth_step(x, 0.0, g["mid"])
    toe = g["mid"] * np.power(np.maximum(x, 0) / g["mid"], g["toe"])
    out = np.where(x >= g["lin"] * g["peak"], g["kA"] + g["kB"] * np.exp(x * g["kC"]), (1 - w) * toe + w * x)
    return np.where(x < 0, 0.0, out)


def gt7_apply(g, rgb):
    u = to_ucs(g, rgb)
    skew = gt7_curve(g, rgb)
    su = to_ucs(g, skew)
    chroma = 1 - smooth_step(u[..., 0] / g["target_ucs"], g["f0"], g["f1"])
    scaled = np.stack([su[..., 0], u[..., 1] * chroma, u[..., 2] * chroma], axis=-1)
    sc = from_ucs(g, scaled)
    return g["sdr"] * np.fmin((1 - g["blend"]) * skew + g["blend"] * sc, g["target"])   # fminf: drops a NaN


# ---- AgX (view_transform.cuh:463-565) ------------------------------------------
AGX_INSET = np.array([[0.856627153315983, 0.137318972929847, 0.11189821299995],
                      [0.0951212405381588, 0.761241990602591, 0.0767994186031903],
                      [0.0482516061458583, 0.101439036467562, 0.811302368396859]]).T
AGX_OUTSET = np.array([[1.1271005818144, -0.14132976349844, -0.14132976349844],
                       [-0.11060664309660, 1.1578237022163, -0.11060664309660],
                       [-0.016493938717835, -0.016493938717834, 1.2519364065950]]).T
AGX_EV = (-12.47393, 4.026069)


def agx_contrast(x):
    return np.polyval([-17.86, 78.01, -126.7, 92.06, -28.72, 4.361, -0.1718, 0.002857], x)


def agx_apply(v, look):
    v = np.fmax(v, 0.0) @ AGX_INSET.T
    t = (np.log2(np.fmax(v, 1e-10)) - AGX_EV[0]) / (AGX_EV[1] - AGX_EV[0])
    e = agx_contrast(np.clip(t, 0.0, 1.0))
    if look != 0:
        luma = e @ np.array([0.2126, 0.7152, 0.0722])
        slope = np.array([1.0, 0.9, 0.5]) if look == 1 else np.ones(3)
        power, sat = (0.8, 1.3) if look == 1 else (1.35, 1.4)
        e = luma[..., None] + sat * (np.power(e * slope, power) - luma[..., None])
    o = e @ AGX_OUTSET.T
    return np.power(np.fmax(o, 0.0), 2.2)


# ---- Khronos PBR Neutral (view_transform.cuh:66-113) ---------------------------
def khronos_apply(v, start_compression, desat):
    start = start_compression - 0.04
    x = np.fmin.reduce(v, axis=-1)
    off = np.where(x < 0.08, x - 6.25 * x * x, 0.04)
    v = v - off[..., None]
    peak = np.fmax.reduce(v, axis=-1)
    d = 1 - start
    newp = 1 - d * d / (peak + d - start)
    scaled = v * (newp / np.where(peak == 0, 1, peak))[..., None]
    g = 1 - 1 / (desat * (peak - newp) + 1)
    comp = scaled * (1 - g[..., None]) + newp[..., None] * g[..., None]
    return np.where((peak < start)[..., None], v, comp)


# ---- the pipeline ----------------------------------------------------------------
def display_transform(xyz, s):
    """xyz: (..., 3); s: dict with the fields of mgk_display_t."""
    sc, dc = SCENE[s["scene_colorspace"]], DISPLAY[s["display_colorspace"]]
    c = np.asarray(xyz, dtype=np.float64) @ xyz_to_scene_matrix(s).T
    v = s["view_transform"]
    if v == 1:
        c = c / (c + 1)
    elif v == 2:
        L = c @ sc[0][1]
        W = s["reinhard_white_luminance"]
        k = np.where(L > 0, L * (1 + L / (W * W)) / (1 + L) / np.where(L > 0, L, 1.0), 0.0)
        c = c * k[..., None]
    elif v == 5:
        c = khronos_apply(c @ convert(sc, SCENE[2]).T, s["khronos_start_compression"], s["khronos_desaturation"]) \
            @ convert(SCENE[2], sc).T
    elif v == 3:
        g = gt7_setup(s["gt7_mode"], s["gt7_sdr_paper_white"], s["gt7_reference_luminance"], s["gt7_jzazbz_exponent_scale"])
        c = gt7_apply(g, c @ convert(sc, SCENE[3]).T) @ convert(SCENE[3], sc).T
    elif v == 4:
        c = agx_apply(c @ convert(sc, SCENE[3]).T, s["agx_look"]) @ convert(SCENE[3], sc).T
    d = c @ convert(sc, dc).T
    return encode(s["display_colorspace"], d)
