// C entry points of the CPU oracle (test infrastructure only: loaded by tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the
// product).  prec = 0 runs the FP32 reference restatement, prec = 1 the
// all-FP64 oracle.  Structs are the ones of include/magik_amd.h so a test can
// hand the same settings to the oracle and to the HIP library.
#include "oracle.hpp"
#include "magik_amd.h"
#include "magik_manager.h"

#include <omp.h>
#include <cstddef>
#include <cstdio>

using namespace oracle;

namespace {

template <class R> MetricParams<R> params_from(const mgk_integrator_t* in) {
    MetricParams<R> p = make_params<R>(in->metric_type, in->time_reversed != 0, in->mass_kg, in->spin,
                                       in->event_horizon_scalar, in->abs_tol, in->rel_tol, in->n_max_substep);
    p.h_max = R(in->h_max);
    return p;
}

// The meshes of orc_set_meshes, in both precisions.
MeshSet<float> g_mesh_f;
MeshSet<double> g_mesh_d;
bool g_mesh_valid = false;
template <class R> const MeshSet<R>* mesh_set();
template <> const MeshSet<float>* mesh_set<float>() { return g_mesh_valid ? &g_mesh_f : nullptr; }
template <> const MeshSet<double>* mesh_set<double>() { return g_mesh_valid ? &g_mesh_d : nullptr; }

// Instance maps as the library stores them: the inverse of object_to_world
// by cofactors in double, rounded to float.
template <class R> bool fill_meshes(MeshSet<R>& ms, const mgk_mesh_t* meshes, uint32_t n_mesh, const mgk_instance_t* inst, uint32_t n_inst) {
    ms = MeshSet<R>();
    for (uint32_t m = 0; m < n_mesh; ++m) {
        std::vector<typename MeshSet<R>::Tri> tris(meshes[m].n_triangle);
        for (uint32_t t = 0; t < meshes[m].n_triangle; ++t) {
            const float* v[3];
            for (int k = 0; k < 3; ++k) {
                const uint32_t vi = meshes[m].indices[3 * t + k];
                if (vi >= meshes[m].n_vertex) return false;
                v[k] = meshes[m].vertices + 3 * vi;
            }
            auto& T = tris[t];
            T.v0 = {R(v[0][0]), R(v[0][1]), R(v[0][2])};
            // edges as floats, as stored
            T.e1 = {R(v[1][0] - v[0][0]), R(v[1][1] - v[0][1]), R(v[1][2] - v[0][2])};
            T.e2 = {R(v[2][0] - v[0][0]), R(v[2][1] - v[0][1]), R(v[2][2] - v[0][2])};
            const double a[3] = {double(v[1][0]) - v[0][0], double(v[1][1]) - v[0][1], double(v[1][2]) - v[0][2]};
            const double b[3] = {double(v[2][0]) - v[0][0], double(v[2][1]) - v[0][1], double(v[2][2]) - v[0][2]};
            double n[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
            const double l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            if (l > 0) { n[0] /= l; n[1] /= l; n[2] /= l; }
            T.n = {R(float(n[0])), R(float(n[1])), R(float(n[2]))};
        }
        ms.mesh.push_back(std::move(tris));
    }
    for (uint32_t k = 0; k < n_inst; ++k) {
        const float* m = inst[k].object_to_world;
        if (inst[k].mesh >= n_mesh) return false;
        const double a = m[0], b = m[1], c = m[2], d = m[4], e = m[5], f = m[6], g = m[8], h = m[9], i = m[10];
        const double det = a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
        if (!(std::fabs(det) > 0)) return false;
        const double r[9] = {(e * i - f * h) / det, (c * h - b * i) / det, (b * f - c * e) / det,
                             (f * g - d * i) / det, (a * i - c * g) / det, (c * d - a * f) / det,
                             (d * h - e * g) / det, (b * g - a * h) / det, (a * e - b * d) / det};
        typename MeshSet<R>::Inst I;
        I.mesh = inst[k].mesh;
        I.matter = inst[k].matter;
        for (int row = 0; row < 3; ++row) {
            for (int col = 0; col < 3; ++col) I.w2o[4 * row + col] = R(float(r[3 * row + col]));
            I.w2o[4 * row + 3] = R(float(-(r[3 * row] * m[3] + r[3 * row + 1] * m[7] + r[3 * row + 2] * m[11])));
            for (int col = 0; col < 3; ++col) I.nm[3 * row + col] = R(float(r[3 * col + row]));
        }
        I.lo = {R(INFINITY), R(INFINITY), R(INFINITY)};
        I.hi = {R(-INFINITY), R(-INFINITY), R(-INFINITY)};
        for (uint32_t vi = 0; vi < meshes[I.mesh].n_vertex; ++vi) {
            const float* v = meshes[I.mesh].vertices + 3 * vi;
            I.lo = {std::min(I.lo.x, R(v[0])), std::min(I.lo.y, R(v[1])), std::min(I.lo.z, R(v[2]))};
            I.hi = {std::max(I.hi.x, R(v[0])), std::max(I.hi.y, R(v[1])), std::max(I.hi.z, R(v[2]))};
        }
        ms.inst.push_back(I);
    }
    return true;
}

template <class R> Scene<R> scene_from(const mgk_scene_t* s, uint32_t step_limit) {
    Scene<R> sc;
    sc.celestial_radius_m = R(s->celestial_radius_m);
    sc.disk = (s->flags & MGK_SCENE_DISK) != 0;
    sc.disk_on_geodesic = (s->flags & MGK_SCENE_DISK_ON_GEODESIC) != 0;
    sc.disk_r_in = R(s->disk_r_in);
    sc.disk_r_out = R(s->disk_r_out);
    sc.disk_T_in = R(s->disk_T_in);
    sc.sphere = (s->flags & MGK_SCENE_SPHERE) != 0;
    sc.sphere_center = {R(s->sphere_center[0]), R(s->sphere_center[1]), R(s->sphere_center[2])};
    sc.sphere_radius = R(s->sphere_radius);
    sc.sphere_T = R(s->sphere_T);
    sc.sphere_checker = s->sphere_checker;
    sc.background_scale = R(s->background_scale);
    sc.background_T = R(s->background_T);
    sc.step_limit = step_limit;
    if (s->flags & MGK_SCENE_BOXES) {
        sc.n_box = s->n_box < 4 ? s->n_box : 4;
        for (uint32_t b = 0; b < sc.n_box; ++b) {
            sc.box[b].mn = {R(s->box[b].min[0]), R(s->box[b].min[1]), R(s->box[b].min[2])};
            sc.box[b].mx = {R(s->box[b].max[0]), R(s->box[b].max[1]), R(s->box[b].max[2])};
            sc.box[b].matter = s->box[b].matter;
        }
    }
    sc.max_reflections = s->max_reflections;
    sc.led_scale = R(s->led_scale);
    if (s->flags & MGK_SCENE_MESHES) sc.meshes = mesh_set<R>();
    sc.hermite = s->chord == MGK_CHORD_HERMITE;
    return sc;
}

template <class R> Camera<R> camera_from(const mgk_camera_t* c) {
    Camera<R> cam;
    cam.position = {R(c->position[0]), R(c->position[1]), R(c->position[2])};
    for (int i = 0; i < 9; ++i) cam.orientation.m[i] = R(c->orientation[i]);
    cam.focal_mm = R(c->focal_length_mm);
    cam.sensor_width_mm = R(c->sensor_width_mm);
    cam.W = c->x_resolution;
    cam.H = c->y_resolution;
    cam.observer_velocity = {R(c->observer_velocity[0]), R(c->observer_velocity[1]), R(c->observer_velocity[2])};
    return cam;
}

template <class R> State<R> load_state(const double* y) {
    return {{R(y[0]), R(y[1]), R(y[2]), R(y[3])}, {R(y[4]), R(y[5]), R(y[6]), R(y[7])}};
}
template <class R> void store_state(const State<R>& s, double* y) {
    y[0] = double(s.q.x); y[1] = double(s.q.y); y[2] = double(s.q.z); y[3] = double(s.q.t);
    y[4] = double(s.p.x); y[5] = double(s.p.y); y[6] = double(s.p.z); y[7] = double(s.p.t);
}

// Ray generation + IC for one (pixel, sample): kernel_generate_thin_lens_ray.cu:51-77
// followed by kernel_generate_initial_conditions.cu:54-74.
template <class R>
void gen_ray(const MetricParams<R>& p, const Camera<R>& cam, uint64_t seed, uint32_t cam_id, uint32_t pixel, uint32_t sample,
             double* out10) {
    float u[4];
    path_uniforms(seed, cam_id, pixel, sample, 0, u);
    uint32_t px = pixel % cam.W, py = pixel / cam.W;
    V3<R> sensor = sensor_position<R>(cam.sensor_width_mm, cam.W, cam.H, px, py, R(u[0]), R(u[1]));
    V3<R> dir, orig;
    pinhole_ray(sensor, cam.focal_mm, &dir, &orig);
    orig = mul(cam.orientation, orig) + cam.position;
    RayIC<R> ic = initial_conditions(cam, p, dir, orig);
    store_state(ic.y, out10);
    out10[8] = double(ic.h);
    out10[9] = double(ic.contraction);
}

template <class R> void trace_one(const MetricParams<R>& p, const Scene<R>& sc, V3<R> origin, const double* ray, mgk_hit_t* hit,
                                  uint32_t* steps, uint32_t* trials) {
    RayIC<R> ic;
    ic.y = load_state<R>(ray);
    ic.h = R(ray[8]);
    ic.contraction = R(ray[9]);
    Hit<R> h = trace_ray(ic, p, sc, origin);
    hit->cls = h.cls;
    hit->g = float(h.g);
    hit->a0 = float(h.a0);
    hit->a1 = float(h.a1);
    if (steps) *steps = h.steps;
    if (trials) *trials = h.trials;
}

// Emission of a terminated path at observed wavelength lam, before the g^5
// boost (matter_examples.cuh:60-95 for the emitter contract; the disk,
// sphere, background and stars are this build's scene additions).
double emission(const mgk_scene_t* s, const StarIndex* stars, const mgk_hit_t& h, double lam_emit) {
    switch (h.cls) {
        case MGK_HIT_DISK: {
            double T = double(s->disk_T_in) * std::pow(double(h.a0) / double(s->disk_r_in), -0.75);
            return planck<double>(lam_emit, T);
        }
        case MGK_HIT_SPHERE: {
            double th = std::acos(std::fmax(-1.0, std::fmin(1.0, double(h.a0))));
            double ph = double(h.a1) + double(kPi);
            uint32_t n = s->sphere_checker;
            uint32_t ci = std::min(n - 1, uint32_t(th / double(kPi) * n));
            uint32_t cj = std::min(n - 1, uint32_t(ph / (2.0 * double(kPi)) * n));
            return ((ci + cj) & 1u) ? 0.0 : planck<double>(lam_emit, double(s->sphere_T));
        }
        case MGK_HIT_BOX_EMITTER:
        case MGK_HIT_MESH_EMITTER:
            return double(s->led_scale) * white_led<double>(lam_emit);
        case MGK_HIT_ESCAPE: {
            double v = 0;
            if (s->flags & MGK_SCENE_BACKGROUND) v += double(s->background_scale) * planck<double>(lam_emit, double(s->background_T));
            if ((s->flags & MGK_SCENE_STARS) && stars) {
                double c = double(h.a0), sn = std::sqrt(std::fmax(0.0, 1.0 - c * c));
                double b = stars->lookup(sn * std::cos(double(h.a1)), c, sn * std::sin(double(h.a1)));
                if (b > 0) v += b * planck<double>(lam_emit, double(s->star_T));
            }
            return v;
        }
        default:
            return 0.0;
    }
}

}  // namespace

extern "C" {

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) { philox4x32_10(ctr, key, out); }

void orc_path_uniforms(uint64_t seed, uint32_t cam, uint32_t pixel, uint32_t sample, uint32_t block, float u[4]) {
    path_uniforms(seed, cam, pixel, sample, block, u);
}

// Right-hand side of the geodesic equation; arg = spin for Kerr (as used by the
// equations), mass sign for Schwarzschild.
void orc_rhs(int prec, int metric, double arg, const double y[8], double out[8]) {
    if (prec == 0) {
        State<float> s = load_state<float>(y), r;
        r = metric == kKerr ? kerr_rhs(s, float(arg)) : metric == kSchwarzschild ? schwarzschild_rhs(s, float(arg)) : minkowski_rhs(s);
        store_state(r, out);
    } else {
        State<double> s = load_state<double>(y), r;
        r = metric == kKerr ? kerr_rhs(s, arg) : metric == kSchwarzschild ? schwarzschild_rhs(s, arg) : minkowski_rhs(s);
        store_state(r, out);
    }
}

// g_{mu nu} (inverse = 0) or g^{mu nu} (inverse = 1) as {tt,tx,ty,tz,xx,xy,xz,yy,yz,zz}.
void orc_metric(int prec, const mgk_integrator_t* integ, const double q[4], int inverse, double out[10]) {
    auto run = [&](auto zero) {
        using R = decltype(zero);
        MetricParams<R> p = params_from<R>(integ);
        Sym4<R> g = compute_g(V4<R>{R(q[0]), R(q[1]), R(q[2]), R(q[3])}, p, inverse != 0);
        R v[10] = {g.tt, g.tx, g.ty, g.tz, g.xx, g.xy, g.xz, g.yy, g.yz, g.zz};
        for (int i = 0; i < 10; ++i) out[i] = double(v[i]);
    };
    if (prec == 0) run(0.0f); else run(0.0);
}

double orc_kerr_schild_r(int prec, double x, double y, double z, double a) {
    if (prec == 0) return double(kerr_schild_r<float>(float(x), float(y), float(z), float(a)));
    return kerr_schild_r<double>(x, y, z, a);
}

// One adaptive step (metric<T>::solve_step): returns y1, updates h, h_taken.
void orc_solve_step(int prec, const mgk_integrator_t* integ, const double y[8], double* h, double* h_taken, double y1[8],
                    uint32_t* trials) {
    auto run = [&](auto zero) {
        using R = decltype(zero);
        MetricParams<R> p = params_from<R>(integ);
        R hh = R(*h), ht = 0;
        State<R> r = solve_step(load_state<R>(y), p, &hh, &ht, trials);
        store_state(r, y1);
        *h = double(hh);
        *h_taken = double(ht);
    };
    if (prec == 0) run(0.0f); else run(0.0);
}

// RKF45 on the harmonic oscillator y'' = -y embedded in the 8-wide state
// (q.x = y, p.x = y'), for controller checks against the reference driver.
void orc_rkf45_oscillator(int prec, const double y[8], double* h, double* h_taken, double h_max, uint32_t n_max,
                          double abs_tol, double rel_tol, double y1[8], uint32_t* trials) {
    auto run = [&](auto zero) {
        using R = decltype(zero);
        auto f = [](const State<R>& s) {
            State<R> d{};
            d.q.x = s.p.x;
            d.p.x = -s.q.x;
            d.q.t = R(1);
            return d;
        };
        R hh = R(*h), ht = 0;
        State<R> r = rkf45_step<R>(f, load_state<R>(y), &hh, &ht, R(h_max), n_max, R(abs_tol), R(rel_tol), trials);
        store_state(r, y1);
        *h = double(hh);
        *h_taken = double(ht);
    };
    if (prec == 0) run(0.0f); else run(0.0);
}

// Rays for (pixel[i], sample[i]); out is n x 10 (x y z t px py pz pt h u.p).
int orc_generate_rays(int prec, const mgk_integrator_t* integ, const mgk_camera_t* cam, uint64_t seed, const uint32_t* pixel,
                      const uint32_t* sample, uint64_t n, double* out) {
    if (!integ || !cam || !pixel || !sample || !out) return MAGIK_ERROR_INVALID_POINTER;
    if (prec == 0) {
        MetricParams<float> p = params_from<float>(integ);
        Camera<float> c = camera_from<float>(cam);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < int64_t(n); ++i) gen_ray(p, c, seed, cam->camera_id, pixel[i], sample[i], out + 10 * i);
    } else {
        MetricParams<double> p = params_from<double>(integ);
        Camera<double> c = camera_from<double>(cam);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < int64_t(n); ++i) gen_ray(p, c, seed, cam->camera_id, pixel[i], sample[i], out + 10 * i);
    }
    return MAGIK_SUCCESS;
}

// Trace n rays (n x 10 doubles, as produced above) to termination with
// `threads` OpenMP threads (0 = all).
int orc_trace_rays(int prec, const mgk_integrator_t* integ, const mgk_scene_t* scene, const float cam_pos[3], const double* rays,
                   uint64_t n, mgk_hit_t* hits, uint32_t* steps, uint32_t* trials, int threads) {
    if (!integ || !scene || !cam_pos || !rays || !hits) return MAGIK_ERROR_INVALID_POINTER;
    if (threads > 0) omp_set_num_threads(threads);
    auto run = [&](auto zero) {
        using R = decltype(zero);
        MetricParams<R> p = params_from<R>(integ);
        Scene<R> sc = scene_from<R>(scene, integ->step_limit);
        V3<R> o{R(cam_pos[0]), R(cam_pos[1]), R(cam_pos[2])};
#pragma omp parallel for schedule(dynamic, 16)
        for (int64_t i = 0; i < int64_t(n); ++i)
            trace_one(p, sc, o, rays + 10 * i, hits + i, steps ? steps + i : nullptr, trials ? trials + i : nullptr);
    };
    if (prec == 0) run(0.0f); else run(0.0);
    return MAGIK_SUCCESS;
}

// Whole paths with reflections off diffuse boxes (trace_path): rays as above,
// ray i being sample sample_begin + i / (W H) of pixel i % (W H) of `cam`,
// whose seed and camera id key the reflections' uniforms.  reflections[i]
// (optional) receives the number of reflections of path i, at[8 i ..] (optional)
// the hit point and momentum of a path that ended on a box.
int orc_trace_paths(int prec, const mgk_integrator_t* integ, const mgk_scene_t* scene, const mgk_camera_t* cam, uint64_t seed,
                    uint32_t sample_begin, const double* rays, uint64_t n, mgk_hit_t* hits, uint32_t* steps, uint32_t* reflections,
                    double* at, int threads) {
    if (!integ || !scene || !cam || !rays || !hits) return MAGIK_ERROR_INVALID_POINTER;
    if (threads > 0) omp_set_num_threads(threads);
    const uint64_t npix = uint64_t(cam->x_resolution) * cam->y_resolution;
    auto run = [&](auto zero) {
        using R = decltype(zero);
        MetricParams<R> p = params_from<R>(integ);
        Scene<R> sc = scene_from<R>(scene, integ->step_limit);
        V3<R> o{R(cam->position[0]), R(cam->position[1]), R(cam->position[2])};
#pragma omp parallel for schedule(dynamic, 16)
        for (int64_t i = 0; i < int64_t(n); ++i) {
            const double* ray = rays + 10 * i;
            RayIC<R> ic;
            ic.y = load_state<R>(ray);
            ic.h = R(ray[8]);
            ic.contraction = R(ray[9]);
            uint32_t k = 0;
            Hit<R> h = trace_path(ic, p, sc, o, seed, cam->camera_id, uint32_t(uint64_t(i) % npix),
                                  sample_begin + uint32_t(uint64_t(i) / npix), &k);
            hits[i].cls = h.cls;
            hits[i].g = float(h.g);
            hits[i].a0 = float(h.a0);
            hits[i].a1 = float(h.a1);
            if (steps) steps[i] = h.steps;
            if (reflections) reflections[i] = k;
            if (at) store_state(h.at, at + 8 * i);
        }
    };
    if (prec == 0) run(0.0f); else run(0.0);
    return MAGIK_SUCCESS;
}

// The box pieces on their own, for the tests: the slab test of a chord
// (origin, unit direction, 1 / chord length) with its normal code, the
// self-intersection offset, one reflection, and the white LED spectrum.
int orc_aabb_hit(int prec, const double o[3], const double d[3], double inv_chord, const double bmin[3], const double bmax[3],
                 double* t, int* normal_code) {
    auto run = [&](auto zero) {
        using R = decltype(zero);
        R tt = 0;
        int nc = 0;
        bool ok = aabb_hit<R>(&tt, V3<R>{R(d[0]), R(d[1]), R(d[2])}, V3<R>{R(o[0]), R(o[1]), R(o[2])}, R(inv_chord),
                              V3<R>{R(bmin[0]), R(bmin[1]), R(bmin[2])}, V3<R>{R(bmax[0]), R(bmax[1]), R(bmax[2])}, &nc);
        *t = double(tt);
        *normal_code = nc;
        return ok ? 1 : 0;
    };
    return prec == 0 ? run(0.0f) : run(0.0);
}

void orc_offset_geodesic(const double q[4], int normal_code, double out[4]) {
    V4<double> r = offset_geodesic(V4<double>{q[0], q[1], q[2], q[3]}, normal_of<double>(normal_code));
    out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.t;
}

int orc_reflect_diffuse(int prec, const mgk_integrator_t* integ, const double at[8], double contraction, int normal_code, double u0,
                        double u1, double out[8], double* contraction_out) {
    if (!integ || !at || !out || !contraction_out) return MAGIK_ERROR_INVALID_POINTER;
    auto run = [&](auto zero) {
        using R = decltype(zero);
        MetricParams<R> p = params_from<R>(integ);
        State<R> y;
        R c = 0;
        reflect_diffuse<R>(p, load_state<R>(at), R(contraction), normal_code, R(u0), R(u1), &y, &c);
        store_state(y, out);
        *contraction_out = double(c);
    };
    if (prec == 0) run(0.0f); else run(0.0);
    return MAGIK_SUCCESS;
}

// Meshes for the scenes with MGK_SCENE_MESHES (the library's mgk_set_meshes
// inputs); n_mesh = 0 clears them.  Returns 0, or -1 for an invalid input.
int orc_set_meshes(const mgk_mesh_t* meshes, uint32_t n_mesh, const mgk_instance_t* inst, uint32_t n_inst) {
    g_mesh_valid = false;
    if (n_mesh == 0) return 0;
    if (!meshes || !inst || n_inst == 0) return -1;
    if (!fill_meshes(g_mesh_f, meshes, n_mesh, inst, n_inst) || !fill_meshes(g_mesh_d, meshes, n_mesh, inst, n_inst)) return -1;
    g_mesh_valid = true;
    return 0;
}

// One segment against one triangle (test access to seg_tri).
int orc_seg_tri(int prec, const double o[3], const double d[3], const double v0[3], const double e1[3], const double e2[3], double tmax,
                double* t) {
    auto run = [&](auto zero) {
        using R = decltype(zero);
        auto v = [](const double* x) { return V3<R>{R(x[0]), R(x[1]), R(x[2])}; };
        R tt = 0;
        const bool ok = seg_tri<R>(v(o), v(d), v(v0), v(e1), v(e2), R(tmax), &tt);
        *t = double(tt);
        return ok ? 1 : 0;
    };
    return prec == 0 ? run(0.0f) : run(0.0);
}

// One chord against the meshes of orc_set_meshes (test access to mesh_chord):
// out = {s, x, y, z}, inst_tri = {instance, triangle}.  Returns 1 on a hit.
int orc_mesh_chord(int prec, const double p0[3], const double p1[3], const double D0[3], const double D1[3], int hermite, double s_max,
                   double out[4], uint32_t inst_tri[2]) {
    auto run = [&](auto zero) {
        using R = decltype(zero);
        const MeshSet<R>* ms = mesh_set<R>();
        if (!ms) return -1;
        auto v = [](const double* x) { return V3<R>{R(x[0]), R(x[1]), R(x[2])}; };
        R s = 0;
        V3<R> x{};
        uint32_t k = 0, j = 0;
        if (!mesh_chord<R>(*ms, v(p0), v(p1), v(D0), v(D1), hermite != 0, R(s_max), &s, &x, &k, &j)) return 0;
        out[0] = double(s); out[1] = double(x.x); out[2] = double(x.y); out[3] = double(x.z);
        inst_tri[0] = k;
        inst_tri[1] = j;
        return 1;
    };
    return prec == 0 ? run(0.0f) : run(0.0);
}

double orc_white_led(int prec, double lam_nm) {
    if (prec == 0) return double(white_led<float>(float(lam_nm)));
    return white_led<double>(lam_nm);
}

// The wavelength sampler's inverse CDF over [lo, hi] (256 entries).
void orc_cmf_inverse_cdf(float lo, float hi, const float* tab, uint32_t n_tab, float tmin, float tmax, float* out) {
    cmf_inverse_cdf(tab, n_tab, tmin, tmax, lo, hi, out);
}

// Bin CMF weights (host_color_utilities.cuh:145-178) from a CIE table (n x 3).
void orc_bin_weights(uint32_t n_bin, float bmin, float bmax, const float* tab, uint32_t n_tab, float tmin, float tmax, float* W) {
    bin_weights(W, n_bin, bmin, bmax, tab, n_tab, tmin, tmax);
}

double orc_planck(int prec, double lam_nm, double T) {
    if (prec == 0) return double(planck<float>(float(lam_nm), float(T)));
    return planck<double>(lam_nm, T);
}

double orc_gaussian_cdf(int prec, double mean, double x, double sigma) {
    if (prec == 0) return double(norm_gaussian_cdf<float>(float(mean), float(x), float(sigma)));
    return norm_gaussian_cdf<double>(mean, x, sigma);
}

// Shade pixels: hits[s * n_pix + j] is sample s of pixel pixel_ids[j]; weights
// are n_bin x 3 XYZ bin weights.  Per sample the hero wavelengths are
// stratified from the block-0 uniform u[2]; each is redshifted (lambda g), its
// emission boosted by g^5 (bxdf_utilities.cuh:156-163), converted to sensor
// irradiance (device_color_utilities.cuh:22-29) and spread over the bins by
// Gaussian-CDF differences (kernel_redistribute_radiance.cu:64-88); the bins
// are accumulated over samples and folded with the weights / spp
// (device_color_utilities.cuh:40-54).  Computed in double.
// With spec->lambda_sampling = MGK_LAMBDA_CMF the hero wavelengths are the
// stratified primary samples mapped through inv_cdf (256 entries; the draw of
// device_sensor_utilities.cuh:25-33) at throughput 1 / (n_hero pdf)
// (kernel_generate_thin_lens_ray.cu:73-74).
int orc_shade(const mgk_scene_t* scene, const mgk_spectral_t* spec, const mgk_camera_t* cam, uint64_t seed, uint32_t sample_begin,
              uint32_t spp, uint32_t n_pix, const uint32_t* pixel_ids, const mgk_hit_t* hits, const float* weights, const float* inv_cdf,
              double* xyz) {
    if (!scene || !spec || !cam || !pixel_ids || !hits || !weights || !xyz) return MAGIK_ERROR_INVALID_POINTER;
    const bool cmf = spec->lambda_sampling == MGK_LAMBDA_CMF;
    if (cmf && !inv_cdf) return MAGIK_ERROR_INVALID_POINTER;
    StarIndex idx;
    bool have_stars = (scene->flags & MGK_SCENE_STARS) && scene->n_star > 0;
    if (have_stars) idx.build(star_catalogue(scene->star_seed, scene->n_star, scene->star_scale), scene->star_radius_rad);
    double inv_b = double(cam->sensor_width_mm) / std::sqrt(double(float(cam->x_resolution * cam->x_resolution + cam->y_resolution * cam->y_resolution)));
    double sw = double(cam->x_resolution) * inv_b, sh = double(cam->y_resolution) * inv_b;
    double span = double(spec->hero_max_nm) - double(spec->hero_min_nm);
    double bw = (double(spec->bin_max_nm) - double(spec->bin_min_nm)) / double(spec->n_bin);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t j = 0; j < int64_t(n_pix); ++j) {
        double bins[64] = {0};
        for (uint32_t s = 0; s < spp; ++s) {
            const mgk_hit_t& h = hits[uint64_t(s) * n_pix + j];
            float u[4];
            path_uniforms(seed, cam->camera_id, pixel_ids[j], sample_begin + s, 0, u);
            double g = double(h.g);
            for (uint32_t k = 0; k < spec->n_hero; ++k) {
                double f = double(u[2]) + double(k) / double(spec->n_hero);
                f -= std::floor(f);
                double lam = double(spec->hero_min_nm) + span * f;
                double thr = span / double(spec->n_hero);
                if (cmf) {
                    const double x = f * 255.0;
                    const uint32_t kk = uint32_t(std::fmin(x, 254.0));
                    const double fr = x - double(kk), c0 = double(inv_cdf[kk]), c1 = double(inv_cdf[kk + 1]);
                    lam = c0 * (1.0 - fr) + c1 * fr;
                    thr = 255.0 * (c1 - c0) / double(spec->n_hero);
                }
                double L = emission(scene, have_stars ? &idx : nullptr, h, lam * g) * (g * g * g * g * g);
                if (std::isnan(L) || std::isinf(L) || L < 0) L = 0;
                if (!(L > 0)) continue;
                double E = radiance_to_irradiance<double>(L, thr, sw, sh, spec->iso, spec->fps, spec->shutter_angle,
                                                          cam->x_resolution, cam->y_resolution);
                for (uint32_t i = 0; i < spec->n_bin; ++i) {
                    double x1 = double(spec->bin_min_nm) + double(i) * bw, x2 = x1 + bw;
                    bins[i] += E * (norm_gaussian_cdf<double>(lam, x2, spec->sigma_nm) - norm_gaussian_cdf<double>(lam, x1, spec->sigma_nm));
                }
            }
        }
        double c[3] = {0, 0, 0};
        for (uint32_t i = 0; i < spec->n_bin; ++i)
            for (int k = 0; k < 3; ++k) c[k] += bins[i] * double(weights[3 * i + k]);
        for (int k = 0; k < 3; ++k) xyz[3 * j + k] = c[k] / double(spp);
    }
    return MAGIK_SUCCESS;
}

// Restatement of magik_trace_geodesic (src/magik.cpp:196-207 with
// relativity/host_trace_geodesic.cu:13-144) in FP32: the host-API initial
// conditions (rotate(x, y, z) applied as rotation(alpha = x_rot, beta = y_rot,
// gamma = z_rot); rotated seeds for the observer, identity seeds for the
// momentum; t0 = 0), then solve_step / rhs per step until the horizon or the
// container is full, then positions scaled to SI.
int orc_host_trace_geodesic(magik_trace_geodesic_container_t* c, magik_trace_geodesic_inital_conditions_t ic,
                            magik_trace_geodesic_metric_parameters_t mp, magik_trace_geodesic_integrator_parameters_t ip) {
    if (!c || !c->steps) return MAGIK_ERROR_INVALID_POINTER;
    if (c->capacity != sizeof(magik_trace_geodesic_step_t) * c->n_allocated_step)
        return MAGIK_ERROR_HOST_TRACE_ALLOCATED_STEPS_AND_CAPACITY_MISMATCH;
    if (int(mp.metric_type) != kMinkowski && int(mp.metric_type) != kSchwarzschild && int(mp.metric_type) != kKerr) return MAGIK_ERROR_UNKNOWN;
    using R = float;
    float mass = int(mp.metric_type) == kKerr ? mp.kerr_mass : mp.schwarzschild_mass;
    MetricParams<R> p = make_params<R>(mp.metric_type, mp.time_reversed, mass, mp.kerr_spin, mp.event_horizon_scalar,
                                       ip.integrator_absolute_error_tolerance, ip.integrator_relative_error_tolerance,
                                       ip.n_integrator_max_substep);
    M3<R> m = rotation<R>(ic.x_rotation, ic.y_rotation, ic.z_rotation);
    V3<R> dir = mul(m, V3<R>{0, 0, R(-1)});
    V3<R> bx = mul(m, V3<R>{R(1), 0, 0}), by = mul(m, V3<R>{0, R(1), 0}), bz = mul(m, V3<R>{0, 0, R(1)});
    Seeds<R> rs;
    rs.X = {bx.x, bx.y, bx.z, 0};
    rs.Y = {by.x, by.y, by.z, 0};
    rs.Z = {bz.x, bz.y, bz.z, 0};
    R sq = 0;
    V3<R> lpv{ic.local_proper_velocity_x, ic.local_proper_velocity_y, ic.local_proper_velocity_z};
    if (ic.test_particle_type == MASSIVE) {
        sq = R(-1);
        dir = dir * (ic.massive_particle_speed / kSpeedOfLight);
    }
    R L = length_scale(p);
    V4<R> q{ic.spatial_position_x / L, ic.spatial_position_y / L, ic.spatial_position_z / L, R(0) / (L / kSpeedOfLight)};
    V4<R> u = rest_frame_four_velocity(q, rs, lpv / kSpeedOfLight, p);
    V4<R> pm = direction_to_momentum(q, Seeds<R>{}, dir / L, u, sq, p);
    State<R> y0{q, pm};
    State<R> ys = rhs(y0, p);
    c->n_written = 0;
    auto put = [&](magik_trace_geodesic_step_t& st, const State<R>& y, const State<R>& d, R contraction, R affine, R dh) {
        st.four_position[0] = y.q.x; st.four_position[1] = y.q.y; st.four_position[2] = y.q.z; st.four_position[3] = y.q.t;
        st.four_momentum[0] = y.p.x; st.four_momentum[1] = y.p.y; st.four_momentum[2] = y.p.z; st.four_momentum[3] = y.p.t;
        st.moment_derivative[0] = d.p.x; st.moment_derivative[1] = d.p.y; st.moment_derivative[2] = d.p.z; st.moment_derivative[3] = d.p.t;
        st.four_vector_contraction = contraction;
        st.affine_parameter = affine;
        st.delta_affine_parameter = dh;
    };
    if (c->n_allocated_step == 0) return MAGIK_SUCCESS;
    put(c->steps[0], y0, ys, dot4(u, pm), R(0), initial_h(p));
    c->n_written = 1;
    while (c->n_written < c->n_allocated_step) {
        auto& prev = c->steps[c->n_written - 1];
        State<R> a{{prev.four_position[0], prev.four_position[1], prev.four_position[2], prev.four_position[3]},
                   {prev.four_momentum[0], prev.four_momentum[1], prev.four_momentum[2], prev.four_momentum[3]}};
        R h = prev.delta_affine_parameter, ht = 0;
        State<R> b = solve_step(a, p, &h, &ht);
        State<R> d = rhs(b, p);
        put(c->steps[c->n_written], b, d, prev.four_vector_contraction, prev.affine_parameter + ht, h);
        c->n_written++;
        if (in_domain(b.q, p)) break;
    }
    for (uint32_t i = 0; i < c->n_written; i++)
        for (int k = 0; k < 4; ++k) c->steps[i].four_position[k] *= L;
    return MAGIK_SUCCESS;
}

void orc_rotation(float a, float b, float g, float out[9]) {
    M3<float> m = rotation<float>(a, b, g);
    for (int i = 0; i < 9; ++i) out[i] = m.m[i];
}

void orc_sample_cmf(const float* tab, uint32_t n, float lam, float out[3]) { sample_cmf(tab, n, 360.0f, 830.0f, lam, out); }

// sizeof / offsetof of the magik.h structs as include/magik_amd.h declares them.
void orc_layout(size_t* out) {
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

// The render-manager API's layout list (manager_layout.inc); returns its length.
size_t orc_manager_layout(size_t* out) {
    size_t i = 0;
#define L(x) out[i++] = (size_t)(x);
#define CMD(T, V) L(sizeof(T)) L(V)
#include "manager_layout.inc"
#undef L
#undef CMD
    return i;
}

// Star catalogue exactly as the oracle generates it (n x 4: dx dy dz brightness).
void orc_star_catalogue(uint64_t seed, uint32_t n, float scale, float* out) {
    std::vector<Star> s = star_catalogue(seed, n, scale);
    for (uint32_t i = 0; i < n; ++i) {
        out[4 * i] = s[i].dx; out[4 * i + 1] = s[i].dy; out[4 * i + 2] = s[i].dz; out[4 * i + 3] = s[i].brightness;
    }
}

// The three per-ray pieces that the end-to-end comparisons only see composed:
// the error norm of the step controller, the sensor point of a pixel and its
// pinhole ray, and the radiance -> sensor irradiance scale.
double orc_rms_error(int prec, const double e[8], const double y0[8], const double y1[8], double abs_tol, double rel_tol) {
    if (prec == 0)
        return double(rms_error(load_state<float>(e), load_state<float>(y0), load_state<float>(y1), float(abs_tol), float(rel_tol)));
    return rms_error(load_state<double>(e), load_state<double>(y0), load_state<double>(y1), abs_tol, rel_tol);
}

// Sensor point (metres) and the pinhole ray's unit direction of pixel (px, py)
// with jitter uniforms (ju, jv): out = {sx, sy, sz, dx, dy, dz}.
void orc_sensor_ray(int prec, double sensor_width_mm, double focal_mm, uint32_t W, uint32_t H, uint32_t px, uint32_t py, double ju,
                    double jv, double out[6]) {
    auto run = [&](auto zero) {
        using R = decltype(zero);
        V3<R> sp = sensor_position<R>(R(sensor_width_mm), W, H, px, py, R(ju), R(jv));
        V3<R> d, o;
        pinhole_ray<R>(sp, R(focal_mm), &d, &o);
        out[0] = double(sp.x); out[1] = double(sp.y); out[2] = double(sp.z);
        out[3] = double(d.x); out[4] = double(d.y); out[5] = double(d.z);
    };
    if (prec == 0) run(0.0f); else run(0.0);
}

double orc_radiance_to_irradiance(int prec, double L, double throughput, double sensor_w_mm, double sensor_h_mm, double iso,
                                  double fps, double shutter, uint32_t W, uint32_t H) {
    if (prec == 0)
        return double(radiance_to_irradiance<float>(float(L), float(throughput), float(sensor_w_mm), float(sensor_h_mm), float(iso),
                                                    float(fps), float(shutter), W, H));
    return radiance_to_irradiance<double>(L, throughput, sensor_w_mm, sensor_h_mm, iso, fps, shutter, W, H);
}

}  // extern "C"
