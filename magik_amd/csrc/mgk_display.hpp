// Display side of the splat: scene-referred colour -> view transform ->
// display primaries and transfer function (kernel_splat_to_renderbuffer.cu:37-57,
// device_color_utilities.cuh:56-72, color_pipeline.cuh:21-48).  Shared by the host
// parameter set-up and the per-pixel kernel (mgk_display.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>

#include "magik_amd.h"

#define MGK_HD __host__ __device__ __forceinline__

namespace mgk {
namespace disp {

struct M3 {
    float m[9];   // row-major
};
struct C3 {
    float x, y, z;
};

MGK_HD C3 mul(const M3& a, C3 v) {
    return {a.m[0] * v.x + a.m[1] * v.y + a.m[2] * v.z, a.m[3] * v.x + a.m[4] * v.y + a.m[5] * v.z,
            a.m[6] * v.x + a.m[7] * v.y + a.m[8] * v.z};
}
inline M3 mul(const M3& a, const M3& b) {
    M3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[3 * i + j] = a.m[3 * i] * b.m[j] + a.m[3 * i + 1] * b.m[3 + j] + a.m[3 * i + 2] * b.m[6 + j];
    return r;
}
inline M3 transpose(const M3& a) {
    return {{a.m[0], a.m[3], a.m[6], a.m[1], a.m[4], a.m[7], a.m[2], a.m[5], a.m[8]}};
}

// ---------------------------------------------------------------------------
// Colour spaces (colorspaces.cuh:8-340): RGB <-> XYZ matrices, white points
// (CIE xy) and, for displays, the transfer function and nominal peak.  The
// matrices are the published ones (IEC 61966-2-1, ITU-R BT.709 / BT.2020 /
// BT.2100, SMPTE ST 2065-1 AP1, Adobe RGB (1998), SMPTE RP 431-2, Display P3);
// tests/test_display_oracle.py derives them again from the primaries.
// ---------------------------------------------------------------------------
struct Space {
    M3 to_xyz, from_xyz;
    float wx, wy;
};

constexpr float kWpE[2] = {1.0f / 3.0f, 1.0f / 3.0f};
constexpr float kWpDci[2] = {0.314f, 0.351f};
constexpr float kWpD65[2] = {0.31272f, 0.32903f};
constexpr float kWpAces[2] = {0.32168f, 0.33767f};

inline bool scene_space(uint32_t id, Space& s) {
    switch (id) {
        case MAGIK_SCENE_COLORSPACE_CIE_XYZ:
            s = {{{1, 0, 0, 0, 1, 0, 0, 0, 1}}, {{1, 0, 0, 0, 1, 0, 0, 0, 1}}, kWpE[0], kWpE[1]};
            return true;
        case MAGIK_SCENE_COLORSPACE_CIE_RGB:
            s = {{{0.49f, 0.31f, 0.2f, 0.1769f, 0.8124f, 0.0107f, 0.0f, 0.0099f, 0.9901f}},
                 {{2.36449012f, -0.89655263f, -0.46793749f, -0.51493525f, 1.42633279f, 0.08860245f, 0.00514883f, -0.01426189f, 1.00911305f}},
                 kWpE[0], kWpE[1]};
            return true;
        case MAGIK_SCENE_COLORSPACE_LINEAR_REC709:
            s = {{{0.41238656f, 0.35759149f, 0.18045049f, 0.21263682f, 0.71518298f, 0.0721802f, 0.01933062f, 0.11919716f, 0.95037259f}},
                 {{3.24100326f, -1.53739899f, -0.49861587f, -0.96922426f, 1.87592999f, 0.04155422f, 0.05563942f, -0.2040112f, 1.05714897f}},
                 kWpD65[0], kWpD65[1]};
            return true;
        case MAGIK_SCENE_COLORSPACE_LINEAR_REC2020:
            s = {{{6.36953507e-01f, 1.44619185e-01f, 1.68855854e-01f, 2.62698339e-01f, 6.78008766e-01f, 5.92928953e-02f, 4.99407097e-17f,
                   2.80731358e-02f, 1.06082723e+00f}},
                 {{1.71666343f, -0.35567332f, -0.25336809f, -0.66667384f, 1.61645574f, 0.0157683f, 0.01764248f, -0.04277698f, 0.94224328f}},
                 kWpD65[0], kWpD65[1]};
            return true;
        case MAGIK_SCENE_COLORSPACE_ACES_CG:
            s = {{{0.66245418f, 0.13400421f, 0.15618769f, 0.27222872f, 0.67408177f, 0.05368952f, -0.00557465f, 0.00406073f, 1.0103391f}},
                 {{1.64102338f, -0.32480329f, -0.2364247f, -0.66366286f, 1.61533159f, 0.01675635f, 0.01172189f, -0.00828444f, 0.98839486f}},
                 kWpAces[0], kWpAces[1]};
            return true;
        case MAGIK_SCENE_COLORSPACE_ADOBE_RGB_1998:
            s = {{{0.57667f, 0.18556f, 0.18823f, 0.29734f, 0.62736f, 0.07529f, 0.02703f, 0.07069f, 0.99134f}},
                 {{2.04159f, -0.56501f, -0.34473f, -0.96924f, 1.87597f, 0.04156f, 0.01344f, -0.11836f, 1.01517f}},
                 kWpD65[0], kWpD65[1]};
            return true;
        default:
            return false;
    }
}

inline bool display_space(uint32_t id, Space& s) {
    switch (id) {
        case MAGIK_DISPLAY_COLORSPACE_SRGB:
            s = {{{0.4124564f, 0.3575761f, 0.1804375f, 0.2126729f, 0.7151522f, 0.0721750f, 0.0193339f, 0.1191920f, 0.9503041f}},
                 {{3.2406255f, -1.537208f, -0.4986286f, -0.9689307f, 1.8757561f, 0.0415175f, 0.0557101f, -0.2040211f, 1.0569959f}},
                 kWpD65[0], kWpD65[1]};
            return true;
        case MAGIK_DISPLAY_COLORSPACE_REC709:
            return scene_space(MAGIK_SCENE_COLORSPACE_LINEAR_REC709, s);
        case MAGIK_DISPLAY_COLORSPACE_REC2100_PQ:
            s = {{{0.63695804830129132f, 0.14461690358620838f, 0.16888097516417205f, 0.26270021201126703f, 0.67799807151887104f,
                   0.059301716469861945f, 0.0f, 0.028072693049087508f, 1.0609850577107909f}},
                 {{1.7166511879712676f, -0.35567078377639239f, -0.2533662813736598f, -0.66668435183248898f, 1.616481236634939f,
                   0.015768545813911131f, 0.017639857445310915f, -0.042770613257808655f, 0.94210312123547402f}},
                 kWpD65[0], kWpD65[1]};
            return true;
        case MAGIK_DISPLAY_COLORSPACE_REC2100_HLG:
            s = {{{0.636958063f, 0.144616902f, 0.168880969f, 0.262700200f, 0.677998066f, 0.059301715f, 0.0f, 0.028072693f, 1.060985090f}},
                 {{1.716651200f, -0.355670780f, -0.253366292f, -0.666684330f, 1.616481180f, 0.015768547f, 0.017639857f, -0.042770613f, 0.942103148f}},
                 kWpD65[0], kWpD65[1]};
            return true;
        case MAGIK_DISPLAY_COLORSPACE_DCI_P3_THEATRICAL_WHITE_POINT:
            s = {{{0.44516982f, 0.27713441f, 0.17228267f, 0.20949168f, 0.72159525f, 0.06891307f, 0.0f, 0.04706056f, 0.90735539f}},
                 {{2.72539403f, -1.01800301f, -0.44016320f, -0.79516803f, 1.68973205f, 0.02264719f, 0.04124189f, -0.08763902f, 1.10092938f}},
                 kWpDci[0], kWpDci[1]};
            return true;
        case MAGIK_DISPLAY_COLORSPACE_DISPLAY_P3:
            s = {{{0.486570954f, 0.265667707f, 0.198217288f, 0.228974566f, 0.691738546f, 0.0792869106f, 0.0f, 0.045113381f, 1.043944360f}},
                 {{2.493496890f, -0.9313836100f, -0.402710795f, -0.829488993f, 1.7626640800f, 0.0236246865f, 0.035845831f, -0.0761723891f,
                   0.956884503f}},
                 kWpD65[0], kWpD65[1]};
            return true;
        case MAGIK_DISPLAY_COLORSPACE_ADOBE_RGB:
            s = {{{0.57666904f, 0.18555824f, 0.18822865f, 0.29734498f, 0.62736357f, 0.07529146f, 0.02703136f, 0.07068885f, 0.99133754f}},
                 {{2.04158790f, -0.56500697f, -0.34473135f, -0.96924364f, 1.87596750f, 0.04155506f, 0.01344428f, -0.11836239f, 1.01517499f}},
                 kWpD65[0], kWpD65[1]};
            return true;
        default:
            return false;
    }
}

// The reference's clamp (internal_math.h:2064-2069): comparisons, so a NaN
// passes through, unlike fminf / fmaxf, which drop it.
MGK_HD float clampc(float x, float lo, float hi) { return x < lo ? lo : (x > hi ? hi : x); }

// x^y for the display curves, as exp2(y log2 x) on the hardware's v_log_f32 and
// v_exp_f32: the rounding of y log2 x gives ~1e-6 relative error for the
// exponents used here, at a few instructions against powf's accurate sequence.
// x < 0 or NaN gives NaN, x = 0 gives 0 (y > 0), as powf does for these
// non-integer exponents.  The host side (parameter set-up) uses powf.
MGK_HD float fpow(float x, float y) {
#if defined(__HIP_DEVICE_COMPILE__)
    const bool tiny = x < 1.17549435e-38f;   // scale denormals into v_log_f32's range
    const float l = __builtin_amdgcn_logf(tiny ? x * 4294967296.0f : x) - (tiny ? 32.0f : 0.0f);
    return __builtin_amdgcn_exp2f(y * l);
#else
    return powf(x, y);
#endif
}

// ---------------------------------------------------------------------------
// Display transfer functions (encode), per channel.  sRGB and Display P3: the
// IEC 61966-2-1 piecewise curve; Rec.709: a pure 1/2.4 power (the reference's
// choice, BT.1886 inverse); BT.2100 PQ on [0, 1] = [0, 10000] cd/m^2 and HLG
// OETF on [0, 1]; DCI-P3: 1/2.6 power; Adobe RGB: 256/563 power.
// ---------------------------------------------------------------------------
MGK_HD float encode_srgb(float L) { return L <= 0.0031308f ? 12.92f * L : 1.055f * fpow(L, 1.0f / 2.4f) - 0.055f; }
MGK_HD float encode_pq(float L) {
    const float m1 = 2610.0f / 16384.0f, m2 = 2523.0f / 4096.0f * 128.0f;
    const float c1 = 3424.0f / 4096.0f, c2 = 2413.0f / 4096.0f * 32.0f, c3 = 2392.0f / 4096.0f * 32.0f;
    const float y = fpow(clampc(L, 0.0f, 1.0f), m1);
    return fpow((c1 + c2 * y) / (1.0f + c3 * y), m2);
}
MGK_HD float encode_hlg(float L) {
    const float a = 0.17883277f, b = 0.28466892f, c = 0.55991073f;
    const float e = clampc(L, 0.0f, 1.0f);
    return e <= 1.0f / 12.0f ? sqrtf(3.0f * e) : a * logf(12.0f * e - b) + c;
}
MGK_HD float encode(uint32_t id, float L) {
    switch (id) {
        case MAGIK_DISPLAY_COLORSPACE_SRGB:
        case MAGIK_DISPLAY_COLORSPACE_DISPLAY_P3: return encode_srgb(L);
        case MAGIK_DISPLAY_COLORSPACE_REC709: return fpow(fmaxf(L, 0.0f), 1.0f / 2.4f);
        case MAGIK_DISPLAY_COLORSPACE_REC2100_PQ: return encode_pq(L);
        case MAGIK_DISPLAY_COLORSPACE_REC2100_HLG: return encode_hlg(L);
        case MAGIK_DISPLAY_COLORSPACE_DCI_P3_THEATRICAL_WHITE_POINT: return fpow(fmaxf(L, 0.0f), 1.0f / 2.6f);
        default: return fpow(clampc(L, 0.0f, 1.0f), 256.0f / 563.0f);   // Adobe RGB
    }
}

// ---------------------------------------------------------------------------
// GT7 tone mapping (view_transform.cuh:115-461), after Polyphony Digital's
// published reference (SIGGRAPH 2025 shading course, "gt7_tone_mapping"):
// a toe / linear / exponential-shoulder curve per channel, blended with a copy
// whose chroma is faded near the target luminance in a perceptual space
// (ICtCp, BT.2100, or Jzazbz, Safdar et al. 2017, with a scaled PQ exponent).
// ---------------------------------------------------------------------------
struct Gt7 {
    uint32_t mode;                 // e_magik_gran_turismo_7_mode_types
    float ref_lum, jz_exp;         // reference luminance (cd/m^2 per framebuffer unit), Jzazbz PQ exponent scale
    float peak, mid, lin, toe;     // curve: peak intensity, mid point, linear section, toe strength
    float kA, kB, kC;              // shoulder kA + kB exp(kC x)
    float sdr_scale, target, target_ucs;
    float blend, fade0, fade1;
};

MGK_HD float smooth_step(float x, float e0, float e1) {
    if (x < e0) return 0.0f;
    if (x > e1) return 1.0f;
    const float t = (x - e0) / (e1 - e0);
    return t * t * (3.0f - 2.0f * t);
}
// ST 2084 in framebuffer units (x ref_lum cd/m^2), PQ exponent m2 scaled by s
MGK_HD float pq_to_fb(float n, float ref_lum, float s) {
    n = clampc(n, 0.0f, 1.0f);
    const float m1 = 0.1593017578125f, m2 = 78.84375f * s, c1 = 0.8359375f, c2 = 18.8515625f, c3 = 18.6875f;
    const float np = fpow(n, 1.0f / m2);
    float l = np - c1;
    if (l < 0.0f) l = 0.0f;
    l = fpow(l / (c2 - c3 * np), 1.0f / m1);
    return l * 10000.0f / ref_lum;
}
MGK_HD float fb_to_pq(float v, float ref_lum, float s) {
    const float m1 = 0.1593017578125f, m2 = 78.84375f * s, c1 = 0.8359375f, c2 = 18.8515625f, c3 = 18.6875f;
    const float ym = fpow(v * ref_lum / 10000.0f, m1);
    return exp2f(m2 * (log2f(c1 + c2 * ym) - log2f(1.0f + c3 * ym)));
}
MGK_HD void to_ucs(const Gt7& g, const float* rgb, float* u) {
    if (g.mode == MAGIK_GRAN_TURISMO_7_MODE_ICTCP) {
        const float l = fb_to_pq((rgb[0] * 1688.0f + rgb[1] * 2146.0f + rgb[2] * 262.0f) / 4096.0f, g.ref_lum, 1.0f);
        const float m = fb_to_pq((rgb[0] * 683.0f + rgb[1] * 2951.0f + rgb[2] * 462.0f) / 4096.0f, g.ref_lum, 1.0f);
        const float s = fb_to_pq((rgb[0] * 99.0f + rgb[1] * 309.0f + rgb[2] * 3688.0f) / 4096.0f, g.ref_lum, 1.0f);
        u[0] = (2048.0f * l + 2048.0f * m) / 4096.0f;
        u[1] = (6610.0f * l - 13613.0f * m + 7003.0f * s) / 4096.0f;
        u[2] = (17933.0f * l - 17390.0f * m - 543.0f * s) / 4096.0f;
    } else {
        const float l = fb_to_pq(rgb[0] * 0.530004f + rgb[1] * 0.355704f + rgb[2] * 0.086090f, g.ref_lum, g.jz_exp);
        const float m = fb_to_pq(rgb[0] * 0.289388f + rgb[1] * 0.525395f + rgb[2] * 0.157481f, g.ref_lum, g.jz_exp);
        const float s = fb_to_pq(rgb[0] * 0.091098f + rgb[1] * 0.147588f + rgb[2] * 0.734234f, g.ref_lum, g.jz_exp);
        const float iz = 0.5f * l + 0.5f * m;
        u[0] = (0.44f * iz) / (1.0f - 0.56f * iz) - 1.6295499532821566e-11f;
        u[1] = 3.524000f * l - 4.066708f * m + 0.542708f * s;
        u[2] = 0.199076f * l + 1.096799f * m - 1.295875f * s;
    }
}
MGK_HD void from_ucs(const Gt7& g, const float* u, float* rgb) {
    if (g.mode == MAGIK_GRAN_TURISMO_7_MODE_ICTCP) {
        const float l = pq_to_fb(u[0] + 0.00860904f * u[1] + 0.11103f * u[2], g.ref_lum, 1.0f);
        const float m = pq_to_fb(u[0] - 0.00860904f * u[1] - 0.11103f * u[2], g.ref_lum, 1.0f);
        const float s = pq_to_fb(u[0] + 0.560031f * u[1] - 0.320627f * u[2], g.ref_lum, 1.0f);
        rgb[0] = fmaxf(3.43661f * l - 2.50645f * m + 0.0698454f * s, 0.0f);
        rgb[1] = fmaxf(-0.79133f * l + 1.9836f * m - 0.192271f * s, 0.0f);
        rgb[2] = fmaxf(-0.0259499f * l - 0.0989137f * m + 1.12486f * s, 0.0f);
    } else {
        const float jz = u[0] + 1.6295499532821566e-11f;
        const float iz = jz / (0.44f + 0.56f * jz);
        const float l = pq_to_fb(iz + u[1] * 1.386050432715393e-1f + u[2] * 5.804731615611869e-2f, g.ref_lum, g.jz_exp);
        const float m = pq_to_fb(iz - u[1] * 1.386050432715393e-1f - u[2] * 5.804731615611869e-2f, g.ref_lum, g.jz_exp);
        const float s = pq_to_fb(iz - u[1] * 9.601924202631895e-2f - u[2] * 8.118918960560390e-1f, g.ref_lum, g.jz_exp);
        rgb[0] = l * 2.990669f - m * 2.049742f + s * 0.088977f;
        rgb[1] = -l * 1.634525f + m * 3.145627f - s * 0.483037f;
        rgb[2] = -l * 0.042505f - m * 0.377983f + s * 1.448019f;
    }
}
MGK_HD float gt7_curve(const Gt7& g, float x) {
    if (x < 0.0f) return 0.0f;
    if (x >= g.lin * g.peak) return g.kA + g.kB * expf(x * g.kC);
    const float w = smooth_step(x, 0.0f, g.mid);
    return (1.0f - w) * (g.mid * fpow(x / g.mid, g.toe)) + w * x;
}
// SDR set-up (view_transform.cuh:328-413): paper white sets the target, the
// curve is (alpha 0.25, mid 0.538, linear 0.444, toe 1.28), fade 0.98-1.16, blend 0.6
inline Gt7 gt7_setup(uint32_t mode, float paper_white, float ref_lum, float jz_exp) {
    Gt7 g{};
    g.mode = mode;
    g.ref_lum = ref_lum;
    g.jz_exp = jz_exp;
    g.target = paper_white / ref_lum;
    g.sdr_scale = 1.0f / g.target;
    const float alpha = 0.25f;
    g.peak = g.target;
    g.mid = 0.538f;
    g.lin = 0.444f;
    g.toe = 1.280f;
    const float k = (g.lin - 1.0f) / (alpha - 1.0f);
    g.kA = g.peak * g.lin + g.peak * k;
    g.kB = -g.peak * k * std::exp(g.lin / k);
    g.kC = -1.0f / (k * g.peak);
    g.blend = 0.6f;
    g.fade0 = 0.98f;
    g.fade1 = 1.16f;
    const float w[3] = {g.target, g.target, g.target};
    float u[3];
    to_ucs(g, w, u);
    g.target_ucs = u[0];
    return g;
}
MGK_HD void gt7_apply(const Gt7& g, const float* rgb, float* out) {
    float u[3], skew[3], su[3], sc[3];
    to_ucs(g, rgb, u);
    for (int i = 0; i < 3; ++i) skew[i] = gt7_curve(g, rgb[i]);
    to_ucs(g, skew, su);
    const float chroma = 1.0f - smooth_step(u[0] / g.target_ucs, g.fade0, g.fade1);
    const float scaled[3] = {su[0], u[1] * chroma, u[2] * chroma};
    from_ucs(g, scaled, sc);
    for (int i = 0; i < 3; ++i) out[i] = g.sdr_scale * fminf((1.0f - g.blend) * skew[i] + g.blend * sc[i], g.target);
}

// ---------------------------------------------------------------------------
// AgX (view_transform.cuh:463-565), after Filament's ToneMapper: inset matrix,
// log2 encoding over [-12.47393, 4.026069] EV, the 7th-order contrast
// polynomial, an optional ASC-CDL look, outset matrix, 2.2 power.
// ---------------------------------------------------------------------------
MGK_HD float agx_contrast(float x) {
    const float x2 = x * x, x4 = x2 * x2, x6 = x4 * x2;
    return -17.86f * x6 * x + 78.01f * x6 - 126.7f * x4 * x + 92.06f * x4 - 28.72f * x2 * x + 4.361f * x2 - 0.1718f * x + 0.002857f;
}
constexpr float kAgxInset[9] = {0.856627153315983f, 0.0951212405381588f, 0.0482516061458583f, 0.137318972929847f, 0.761241990602591f,
                                0.101439036467562f,  0.11189821299995f,   0.0767994186031903f, 0.811302368396859f};
constexpr float kAgxOutset[9] = {1.1271005818144f,   -0.11060664309660f, -0.016493938717835f, -0.14132976349844f, 1.1578237022163f,
                                 -0.016493938717834f, -0.14132976349844f, -0.11060664309660f,  1.2519364065950f};

// ---------------------------------------------------------------------------
// The whole per-pixel pipeline.
// ---------------------------------------------------------------------------
struct Params {
    uint32_t view, display;
    M3 xyz_to_scene;       // white-balanced XYZ -> scene colour space
    M3 scene_to_display;   // scene -> display primaries, with the white-point adaptation
    M3 to_work, from_work; // scene <-> the view transform's working space (Rec.709 or Rec.2020)
    C3 lum;                // scene-space luminance weights (Reinhard extended)
    float reinhard_white;
    float khr_start, khr_desat;
    uint32_t agx_look;
    Gt7 gt7;
};

MGK_HD C3 view_transform(const Params& P, C3 c) {
    switch (P.view) {
        case MAGIK_VIEW_TRANSFORM_REINHARD_SIMPLE:
            return {c.x / (c.x + 1.0f), c.y / (c.y + 1.0f), c.z / (c.z + 1.0f)};
        case MAGIK_VIEW_TRANSFORM_REINHARD_EXTENDED: {
            const float L = P.lum.x * c.x + P.lum.y * c.y + P.lum.z * c.z;
            if (!(L > 0.0f)) return {0.0f, 0.0f, 0.0f};
            const float L0 = L * (1.0f + L / (P.reinhard_white * P.reinhard_white)) / (1.0f + L);
            const float s = L0 / L;
            return {c.x * s, c.y * s, c.z * s};
        }
        case MAGIK_VIEW_TRANSFORM_KHRONOS_PBR_NEUTRAL: {
            C3 v = mul(P.to_work, c);
            const float start = P.khr_start - 0.04f;
            const float x = fminf(v.x, fminf(v.y, v.z));
            const float off = x < 0.08f ? x - 6.25f * x * x : 0.04f;
            v = {v.x - off, v.y - off, v.z - off};
            const float peak = fmaxf(v.x, fmaxf(v.y, v.z));
            if (peak < start) return mul(P.from_work, v);
            const float d = 1.0f - start;
            const float np = 1.0f - d * d / (peak + d - start);
            const float k = np / peak;
            v = {v.x * k, v.y * k, v.z * k};
            const float g = 1.0f - 1.0f / (P.khr_desat * (peak - np) + 1.0f);
            return mul(P.from_work, C3{v.x * (1.0f - g) + np * g, v.y * (1.0f - g) + np * g, v.z * (1.0f - g) + np * g});
        }
        case MAGIK_VIEW_TRANSFORM_GRAN_TURISMO_7: {
            const C3 v = mul(P.to_work, c);   // Rec.2020; the curve clamps negatives itself
            const float in[3] = {v.x, v.y, v.z};
            float out[3];
            gt7_apply(P.gt7, in, out);
            return mul(P.from_work, C3{out[0], out[1], out[2]});
        }
        case MAGIK_VIEW_TRANSFORM_AGX: {
            C3 v = mul(P.to_work, c);   // Rec.2020
            v = {fmaxf(v.x, 0.0f), fmaxf(v.y, 0.0f), fmaxf(v.z, 0.0f)};
            float e[3] = {kAgxInset[0] * v.x + kAgxInset[1] * v.y + kAgxInset[2] * v.z,
                          kAgxInset[3] * v.x + kAgxInset[4] * v.y + kAgxInset[5] * v.z,
                          kAgxInset[6] * v.x + kAgxInset[7] * v.y + kAgxInset[8] * v.z};
            const float lo = -12.47393f, hi = 4.026069f;
            for (int i = 0; i < 3; ++i) {
                const float t = (log2f(fmaxf(e[i], 1e-10f)) - lo) / (hi - lo);
                e[i] = agx_contrast(clampc(t, 0.0f, 1.0f));
            }
            if (P.agx_look != MAGIK_AGX_LOOK_NONE) {
                const float luma = 0.2126f * e[0] + 0.7152f * e[1] + 0.0722f * e[2];
                const bool golden = P.agx_look == MAGIK_AGX_LOOK_GOLDEN;
                const float slope[3] = {1.0f, golden ? 0.9f : 1.0f, golden ? 0.5f : 1.0f};
                const float power = golden ? 0.8f : 1.35f, sat = golden ? 1.3f : 1.4f;
                for (int i = 0; i < 3; ++i) e[i] = luma + sat * (fpow(e[i] * slope[i], power) - luma);
            }
            C3 o{kAgxOutset[0] * e[0] + kAgxOutset[1] * e[1] + kAgxOutset[2] * e[2],
                 kAgxOutset[3] * e[0] + kAgxOutset[4] * e[1] + kAgxOutset[5] * e[2],
                 kAgxOutset[6] * e[0] + kAgxOutset[7] * e[1] + kAgxOutset[8] * e[2]};
            o = {fpow(fmaxf(o.x, 0.0f), 2.2f), fpow(fmaxf(o.y, 0.0f), 2.2f), fpow(fmaxf(o.z, 0.0f), 2.2f)};
            return mul(P.from_work, o);
        }
        default:
            return c;
    }
}

MGK_HD C3 display_pixel(const Params& P, C3 xyz) {
    const C3 scene = mul(P.xyz_to_scene, xyz);
    const C3 d = mul(P.scene_to_display, view_transform(P, scene));
    return {encode(P.display, d.x), encode(P.display, d.y), encode(P.display, d.z)};
}

}  // namespace disp
}  // namespace mgk
