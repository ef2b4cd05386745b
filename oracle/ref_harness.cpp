// Harness around the reference's own host-compilable headers, built only when
// the reference sources are present (see oracle/Makefile, target `ref`).
// Compiles internal_math.h, magik.h, color_matching_functions.cuh and
// colorspaces.cuh as they lie in the reference tree -- no stand-in headers --
// and exposes their arithmetic through extern "C" so tests can pin the oracle
// against it.
// TEST INFRASTRUCTURE ONLY; output goes to oracle/_ref/.
#include "magik.h"
#include "internal_math.h"
#include "color_matching_functions.cuh"
#include "colorspaces.cuh"
#include "oracle.hpp"

#include <cstddef>

namespace {

// An 8-wide geodesic state with the interface the reference RKF45 driver
// (internal_math.h:2160-2205) expects: +, -, float * state and
// compute_rms_error.  Written for this harness; the norm follows
// relativity_utilities.cuh:15-29.
struct hstate {
    magik::vec4f q{}, p{};
    float compute_rms_error(hstate e, hstate y0, hstate y1, float abs_tol, float rel_tol) {
        auto c = [&](float ev, float a, float b) { return ev / (abs_tol + rel_tol * fmaxf(fabsf(a), fabsf(b))); };
        float s0 = c(e.q.t, y0.q.t, y1.q.t), s1 = c(e.q.x, y0.q.x, y1.q.x), s2 = c(e.q.y, y0.q.y, y1.q.y), s3 = c(e.q.z, y0.q.z, y1.q.z);
        float s4 = c(e.p.t, y0.p.t, y1.p.t), s5 = c(e.p.x, y0.p.x, y1.p.x), s6 = c(e.p.y, y0.p.y, y1.p.y), s7 = c(e.p.z, y0.p.z, y1.p.z);
        float total = (s0 * s0 + s1 * s1 + s2 * s2 + s3 * s3 + s4 * s4 + s5 * s5 + s6 * s6 + s7 * s7) / 8.0f;
        return sqrtf(total);
    }
};
hstate operator+(hstate a, hstate b) { return {a.q + b.q, a.p + b.p}; }
hstate operator-(hstate a, hstate b) { return {a.q - b.q, a.p - b.p}; }
hstate operator*(float s, hstate a) { return {s * a.q, s * a.p}; }

hstate load(const double* y) {
    return {{float(y[0]), float(y[1]), float(y[2]), float(y[3])}, {float(y[4]), float(y[5]), float(y[6]), float(y[7])}};
}
void store(const hstate& s, double* y) {
    y[0] = s.q.x; y[1] = s.q.y; y[2] = s.q.z; y[3] = s.q.t; y[4] = s.p.x; y[5] = s.p.y; y[6] = s.p.z; y[7] = s.p.t;
}

hstate oscillator(hstate s) {
    hstate d{};
    d.q.x = s.p.x;
    d.p.x = -s.q.x;
    d.q.t = 1.0f;
    return d;
}

// The oracle's Kerr right-hand side in the harness' state type.
void put(const magik::mat3f& m, float* out) {
    const float v[9] = {m.m11, m.m12, m.m13, m.m21, m.m22, m.m23, m.m31, m.m32, m.m33};
    for (int i = 0; i < 9; ++i) out[i] = v[i];
}

template <class S>
void scene_space(float* to, float* from, float* wp) {
    put(S::colorspace_to_XYZ, to);
    put(S::XYZ_to_colorspace, from);
    wp[0] = S::white_point_cie_xy.x;
    wp[1] = S::white_point_cie_xy.y;
}

template <class D>
void display_space(float* to, float* from, float* wp, float* peak) {
    scene_space<D>(to, from, wp);
    *peak = D::peak_luminance;
}

template <class D>
void encode_all(const float* in, unsigned n, float* out) {
    for (unsigned i = 0; i < n; ++i) {
        const magik::vec3f v = D::encode(magik::vec3f{in[3 * i], in[3 * i + 1], in[3 * i + 2]});
        out[3 * i] = v.x; out[3 * i + 1] = v.y; out[3 * i + 2] = v.z;
    }
}

hstate kerr(hstate s, float a) {
    oracle::State<float> o{{s.q.x, s.q.y, s.q.z, s.q.t}, {s.p.x, s.p.y, s.p.z, s.p.t}};
    oracle::State<float> r = oracle::kerr_rhs(o, a);
    return {{r.q.x, r.q.y, r.q.z, r.q.t}, {r.p.x, r.p.y, r.p.z, r.p.t}};
}

}  // namespace

extern "C" {

void ref_rkf45_oscillator(const double y[8], float* h, float* h_taken, float h_max, unsigned n_max, float abs_tol, float rel_tol,
                          double y1[8]) {
    hstate r = magik::solve_runge_kutta_fehlberg_45_step(oscillator, load(y), h, h_taken, h_max, n_max, abs_tol, rel_tol);
    store(r, y1);
}

void ref_rkf45_kerr(const double y[8], float a, float* h, float* h_taken, float h_max, unsigned n_max, float abs_tol, float rel_tol,
                    double y1[8]) {
    hstate r = magik::solve_runge_kutta_fehlberg_45_step(kerr, load(y), h, h_taken, h_max, n_max, abs_tol, rel_tol, a);
    store(r, y1);
}

float ref_planck_f32(float lam, float T) { return magik::sample_plancks_law<float>(lam, T); }
double ref_planck_f64(double lam, double T) { return magik::sample_plancks_law<double>(lam, T); }
float ref_gaussian_cdf_f32(float mean, float x, float sigma) { return magik::norm_gaussian_cdf<float>(mean, x, sigma); }
float ref_white_led(float lam) { return magik::evaluate_white_led_spectrum<float>(lam); }

void ref_rotate(float a, float b, float g, float out[9]) {
    magik::mat3f m;
    m = m.make_identity();
    m = m.rotate(a, b, g);
    float v[9] = {m.m11, m.m12, m.m13, m.m21, m.m22, m.m23, m.m31, m.m32, m.m33};
    for (int i = 0; i < 9; ++i) out[i] = v[i];
}

void ref_sample_cmf(float lam, float out[3]) {
    using C = magik::generic_color_matching_function_t<MAGIK_CMF_CIE_1931>;
    magik::vec3f v = magik::sample_equidistant_spectral_data(C::values, C::n_entries, C::min_wavelength_nm, C::max_wavelength_nm, lam);
    out[0] = v.x; out[1] = v.y; out[2] = v.z;
}

unsigned ref_cmf_table(float* out) {
    using C = magik::generic_color_matching_function_t<MAGIK_CMF_CIE_1931>;
    for (unsigned i = 0; i < C::n_entries; ++i) {
        out[3 * i] = C::values[i].x; out[3 * i + 1] = C::values[i].y; out[3 * i + 2] = C::values[i].z;
    }
    return C::n_entries;
}

// Scene colour space `id` (colorspaces.cuh:24-63): RGB -> XYZ and XYZ -> RGB
// (row-major) and the white point's CIE xy.  Returns 0 for an unknown id.
int ref_scene_colorspace(unsigned id, float* to, float* from, float* wp) {
    using namespace magik;
    switch (id) {
        case MAGIK_SCENE_COLORSPACE_CIE_XYZ: scene_space<generic_scene_colorspace_t<MAGIK_SCENE_COLORSPACE_CIE_XYZ>>(to, from, wp); return 1;
        case MAGIK_SCENE_COLORSPACE_CIE_RGB: scene_space<generic_scene_colorspace_t<MAGIK_SCENE_COLORSPACE_CIE_RGB>>(to, from, wp); return 1;
        case MAGIK_SCENE_COLORSPACE_LINEAR_REC709: scene_space<generic_scene_colorspace_t<MAGIK_SCENE_COLORSPACE_LINEAR_REC709>>(to, from, wp); return 1;
        case MAGIK_SCENE_COLORSPACE_LINEAR_REC2020: scene_space<generic_scene_colorspace_t<MAGIK_SCENE_COLORSPACE_LINEAR_REC2020>>(to, from, wp); return 1;
        case MAGIK_SCENE_COLORSPACE_ACES_CG: scene_space<generic_scene_colorspace_t<MAGIK_SCENE_COLORSPACE_ACES_CG>>(to, from, wp); return 1;
        case MAGIK_SCENE_COLORSPACE_ADOBE_RGB_1998: scene_space<generic_scene_colorspace_t<MAGIK_SCENE_COLORSPACE_ADOBE_RGB_1998>>(to, from, wp); return 1;
        default: return 0;
    }
}

// Display colour space `id` (colorspaces.cuh:90-338): matrices, white point,
// peak luminance, and its encode applied to n RGB triples (in -> out).
int ref_display_colorspace(unsigned id, float* to, float* from, float* wp, float* peak, const float* in, unsigned n, float* out) {
    using namespace magik;
    switch (id) {
#define MGK_REF_DISPLAY(ID)                                                   \
    case ID:                                                                  \
        display_space<generic_display_colorspace_t<ID>>(to, from, wp, peak); \
        encode_all<generic_display_colorspace_t<ID>>(in, n, out);           \
        return 1;
        MGK_REF_DISPLAY(MAGIK_DISPLAY_COLORSPACE_SRGB)
        MGK_REF_DISPLAY(MAGIK_DISPLAY_COLORSPACE_REC709)
        MGK_REF_DISPLAY(MAGIK_DISPLAY_COLORSPACE_REC2100_PQ)
        MGK_REF_DISPLAY(MAGIK_DISPLAY_COLORSPACE_REC2100_HLG)
        MGK_REF_DISPLAY(MAGIK_DISPLAY_COLORSPACE_DCI_P3_THEATRICAL_WHITE_POINT)
        MGK_REF_DISPLAY(MAGIK_DISPLAY_COLORSPACE_DISPLAY_P3)
        MGK_REF_DISPLAY(MAGIK_DISPLAY_COLORSPACE_ADOBE_RGB)
#undef MGK_REF_DISPLAY
        default: return 0;
    }
}

// The render-manager API's layout list (manager_layout.inc); returns its length.
size_t ref_manager_layout(size_t* out) {
    size_t i = 0;
#define L(x) out[i++] = (size_t)(x);
#define CMD(T, V) L(sizeof(T)) L(V)
#include "manager_layout.inc"
#undef L
#undef CMD
    return i;
}

// sizeof / offsetof of the magik.h structs the library re-declares.
void ref_layout(size_t* out) {
    size_t i = 0;
    out[i++] = sizeof(magik_trace_geodesic_metric_parameters_t);
    out[i++] = offsetof(magik_trace_geodesic_metric_parameters_t, time_reversed);
    out[i++] = offsetof(magik_trace_geodesic_metric_parameters_t, event_horizon_scalar);
    out[i++] = offsetof(magik_trace_geodesic_metric_parameters_t, kerr_spin);
    out[i++] = sizeof(magik_trace_geodesic_integrator_parameters_t);
    out[i++] = sizeof(magik_trace_geodesic_inital_conditions_t);
    out[i++] = offsetof(magik_trace_geodesic_inital_conditions_t, massive_particle_speed);
    out[i++] = sizeof(magik_trace_geodesic_step_t);
    out[i++] = offsetof(magik_trace_geodesic_step_t, four_vector_contraction);
    out[i++] = sizeof(magik_trace_geodesic_container_t);
    out[i++] = offsetof(magik_trace_geodesic_container_t, n_written);
    out[i++] = offsetof(magik_trace_geodesic_container_t, steps);
    out[i++] = (size_t)MAGIK_ERROR_HOST_TRACE_ALLOCATED_STEPS_AND_CAPACITY_MISMATCH;
    out[i++] = (size_t)MAGIK_ERROR_UNKNOWN_METRIC;
}

}  // extern "C"
