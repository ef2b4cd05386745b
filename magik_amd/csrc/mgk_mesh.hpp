// Triangle meshes in a two-level bounding volume hierarchy, and the chord of
// a completed step against them (affine segment or cubic Hermite curve).
//
// The reference reserves these stages (top- and bottom-level, affine and
// Hermite intersection: wavefront.cpp:94-100) and stores the end-point
// derivatives of every step for them (kernel_advance_geodesic.cu:67-72); it
// has no implementation.  Layouts, all in HBM and read-only while tracing:
//
//   nodes      2 float4 per node: {lo.xyz, skip}, {hi.xyz, info}.  Every
//              hierarchy is laid out in depth-first order, so the first child
//              of an interior node is the next node, and `skip` is the node
//              after the subtree (kBvhEnd at the end of the hierarchy): the
//              traversal needs no stack.  info = 0 for an interior node; a
//              leaf holds (count << 24) | first.  Bottom levels (one per mesh)
//              index triangle slots; the top level's leaves hold one instance.
//   triangles  3 float4 per slot, in leaf order: {v0, index in the mesh},
//              {v1 - v0, 0}, {v2 - v0, 0} (object space, SI metres).
//   normals    1 float4 per triangle of a mesh, in the mesh's own order: the
//              unit geometric normal (object space).
//   instances  kInstF4 float4: rows 0-2 the world -> object affine map,
//              rows 3-5 the normal map object -> world (the transposed linear
//              part of rows 0-2) with .w = bottom-level root, first normal,
//              matter.
#pragma once

#include "mgk_device.hpp"
#include "mgk_internal.hpp"

namespace mgk {

template <class T> __device__ __forceinline__ T min_(T a, T b) { return a < b ? a : b; }
template <class T> __device__ __forceinline__ T max_(T a, T b) { return a > b ? a : b; }

// Slab test of o + t d, t in [0, tmax], against [lo, hi] (Kay & Kajiya,
// SIGGRAPH 1986).  inv = 1 / d; an axis with d = 0 gives +-inf bounds, and a
// NaN from 0 * inf drops out of the min / max (fminf semantics), which only
// widens the test.
template <class T>
__device__ __forceinline__ bool slab(const T o[3], const T inv[3], const float4 lo, const float4 hi, T tmax) {
    const T ax = (T(lo.x) - o[0]) * inv[0], bx = (T(hi.x) - o[0]) * inv[0];
    const T ay = (T(lo.y) - o[1]) * inv[1], by = (T(hi.y) - o[1]) * inv[1];
    const T az = (T(lo.z) - o[2]) * inv[2], bz = (T(hi.z) - o[2]) * inv[2];
    const T tn = fmax(fmax(fmin(ax, bx), fmin(ay, by)), fmax(fmin(az, bz), T(0)));
    const T tf = fmin(fmin(fmax(ax, bx), fmax(ay, by)), fmin(fmax(az, bz), tmax));
    return tn <= tf;
}

// Segment o + t d against a two-sided triangle (Moller & Trumbore, JGT 1997),
// d not normalised: accepts 0 <= t < tmax.
template <class T>
__device__ __forceinline__ bool seg_tri(const T o[3], const T d[3], const float4* __restrict__ tri, T tmax, T& t_out) {
#pragma clang fp contract(off)
    const float4 A = tri[0], E1 = tri[1], E2 = tri[2];
    const T e1x = T(E1.x), e1y = T(E1.y), e1z = T(E1.z), e2x = T(E2.x), e2y = T(E2.y), e2z = T(E2.z);
    const T px = d[1] * e2z - d[2] * e2y, py = d[2] * e2x - d[0] * e2z, pz = d[0] * e2y - d[1] * e2x;
    const T det = e1x * px + e1y * py + e1z * pz;
    if (det == T(0)) return false;
    const T inv = T(1) / det;
    const T tx = o[0] - T(A.x), ty = o[1] - T(A.y), tz = o[2] - T(A.z);
    const T u = (tx * px + ty * py + tz * pz) * inv;
    if (u < T(0) || u > T(1)) return false;
    const T qx = ty * e1z - tz * e1y, qy = tz * e1x - tx * e1z, qz = tx * e1y - ty * e1x;
    const T v = (d[0] * qx + d[1] * qy + d[2] * qz) * inv;
    if (v < T(0) || u + v > T(1)) return false;
    const T t = (e2x * qx + e2y * qy + e2z * qz) * inv;
    if (!(t >= T(0) && t < tmax)) return false;
    t_out = t;
    return true;
}

// Closest triangle along o + t d (world space, t in [0, tb)) through the top
// level and the bottom level of every instance it reaches.  The segment is
// taken into each instance's object space, where t keeps its meaning.  On a
// hit tb is lowered to it and the instance and triangle slot are returned.
template <class T>
__device__ __forceinline__ bool mesh_segment(const TraceParams& P, const T o[3], const T d[3], T& tb, uint32_t& hit_inst,
                                          uint32_t& hit_slot) {
#pragma clang fp contract(off)
    const float4* __restrict__ nodes = P.mesh_nodes;
    const T inv[3] = {T(1) / d[0], T(1) / d[1], T(1) / d[2]};
    bool found = false;
    uint32_t n = P.mesh_root;
    while (n != kBvhEnd) {
        const float4 lo = nodes[2u * n], hi = nodes[2u * n + 1u];
        const uint32_t skip = __float_as_uint(lo.w), info = __float_as_uint(hi.w);
        if (!slab(o, inv, lo, hi, tb)) { n = skip; continue; }
        if (info == 0u) { ++n; continue; }
        const uint32_t k = info & kLeafMask;
        const float4* __restrict__ I = P.mesh_inst + uint64_t(kInstF4) * k;
        const float4 r0 = I[0], r1 = I[1], r2 = I[2];
        const T oo[3] = {T(r0.x) * o[0] + T(r0.y) * o[1] + T(r0.z) * o[2] + T(r0.w),
                         T(r1.x) * o[0] + T(r1.y) * o[1] + T(r1.z) * o[2] + T(r1.w),
                         T(r2.x) * o[0] + T(r2.y) * o[1] + T(r2.z) * o[2] + T(r2.w)};
        const T dd[3] = {T(r0.x) * d[0] + T(r0.y) * d[1] + T(r0.z) * d[2], T(r1.x) * d[0] + T(r1.y) * d[1] + T(r1.z) * d[2],
                         T(r2.x) * d[0] + T(r2.y) * d[1] + T(r2.z) * d[2]};
        const T ii[3] = {T(1) / dd[0], T(1) / dd[1], T(1) / dd[2]};
        uint32_t m = __float_as_uint(I[3].w);
        while (m != kBvhEnd) {
            const float4 a = nodes[2u * m], b = nodes[2u * m + 1u];
            const uint32_t sk = __float_as_uint(a.w), in = __float_as_uint(b.w);
            if (!slab(oo, ii, a, b, tb)) { m = sk; continue; }
            if (in == 0u) { ++m; continue; }
            const uint32_t first = in & kLeafMask, end = first + (in >> kLeafShift);
            for (uint32_t j = first; j < end; ++j) {
                T t;
                if (seg_tri(oo, dd, P.mesh_tris + 3u * uint64_t(j), tb, t)) {
                    tb = t;
                    hit_inst = k;
                    hit_slot = j;
                    found = true;
                }
            }
            m = sk;
        }
        n = skip;
    }
    return found;
}

// Point and derivative of the cubic Hermite curve with end points p0, p1 and
// end tangents D0, D1 (already scaled by the curve's parameter length).
template <class T>
__device__ __forceinline__ T hermite(T p0, T D0, T p1, T D1, T s) {
#pragma clang fp contract(off)
    const T s2 = s * s, s3 = s2 * s;
    return (T(2) * s3 - T(3) * s2 + T(1)) * p0 + (s3 - T(2) * s2 + s) * D0 + (T(3) * s2 - T(2) * s3) * p1 + (s3 - s2) * D1;
}
template <class T>
__device__ __forceinline__ T hermite_d(T p0, T D0, T p1, T D1, T s) {
#pragma clang fp contract(off)
    const T s2 = s * s;
    return (T(6) * s2 - T(6) * s) * (p0 - p1) + (T(3) * s2 - T(4) * s + T(1)) * D0 + (T(3) * s2 - T(2) * s) * D1;
}

// Result of the mesh test of one chord.
template <class T> struct MeshHit {
    T s;                 // curve parameter of the hit in [0, 1]
    T x[3];              // hit point, world SI metres
    uint32_t inst, tri;  // instance, triangle index in its mesh
    uint32_t matter;
};

// The first hit along n straight sub-chords through the curve's points at
// s_lo + j (s_hi - s_lo) / n, j = 0..n, accepted below the curve parameter
// s_max.  Returns the sub-chord's index (-1: no hit), the hit's curve
// parameter and its point on the sub-chord.
template <class T>
__device__ __forceinline__ int curve_pass(const TraceParams& P, const T p0[3], const T D0[3], const T p1[3], const T D1[3], T s_lo,
                                          T s_hi, int n, T s_max, T& s_hit, T x[3], uint32_t& inst, uint32_t& slot) {
#pragma clang fp contract(off)
    const T ds = (s_hi - s_lo) / T(n);
    T q[3];
    for (int c = 0; c < 3; ++c) q[c] = s_lo == T(0) ? p0[c] : hermite(p0[c], D0[c], p1[c], D1[c], s_lo);
    for (int j = 0; j < n; ++j) {
        const T tmax = min_((s_max - s_lo) / ds - T(j), T(1));
        if (!(tmax > T(0))) break;
        const T s1 = s_lo + T(j + 1) * ds;
        T e[3];
        for (int c = 0; c < 3; ++c) e[c] = s1 >= T(1) ? p1[c] : hermite(p0[c], D0[c], p1[c], D1[c], s1);
        const T d[3] = {e[0] - q[0], e[1] - q[1], e[2] - q[2]};
        T tb = tmax;
        if (mesh_segment<T>(P, q, d, tb, inst, slot)) {
            s_hit = s_lo + (T(j) + tb) * ds;
            for (int c = 0; c < 3; ++c) x[c] = q[c] + tb * d[c];
            return j;
        }
        q[0] = e[0]; q[1] = e[1]; q[2] = e[2];
    }
    return -1;
}

// The chord of one completed step against the meshes.  p0, p1: end points
// (world SI); for Hermite chords D0, D1 are the end tangents times the step's
// parameter length.  Hits are accepted below s_max (the closest hit of the
// other traceables, in chord units).  A Hermite chord is followed along
// kHermiteSegments straight sub-chords through points of the curve, and the
// first that hits along kHermiteFine finer ones.  Newton's method on the hit
// triangle's plane then moves the hit onto the cubic itself; where the
// crossing found that way leaves the triangle (the curve passes into a
// neighbour the sub-chord missed), the sub-chord's hit is kept, so a hit
// point always lies on the triangle its record names.
template <class T>
__device__ __forceinline__ bool mesh_chord(const TraceParams& P, const T p0[3], const T p1[3], const T D0[3], const T D1[3],
                                           bool hermite_chord, T s_max, MeshHit<T>& h) {
#pragma clang fp contract(off)
    // cull against the world bounds of the top level: the chord's bounds, or the
    // Hermite curve's Bezier control polygon (p0, p0 + D0/3, p1 - D1/3, p1), which holds the curve
    for (int c = 0; c < 3; ++c) {
        T lo = min_(p0[c], p1[c]), hi = max_(p0[c], p1[c]);
        if (hermite_chord) {
            const T b1 = p0[c] + D0[c] / T(3), b2 = p1[c] - D1[c] / T(3);
            lo = min_(lo, min_(b1, b2));
            hi = max_(hi, max_(b1, b2));
        }
        if (hi < T(P.mesh_lo[c]) || lo > T(P.mesh_hi[c])) return false;
    }
    uint32_t inst = 0, slot = 0;
    if (!hermite_chord) {
        const T d[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
        T tb = min_(s_max, T(1));
        if (!mesh_segment<T>(P, p0, d, tb, inst, slot)) return false;
        h.s = tb;
        for (int c = 0; c < 3; ++c) h.x[c] = p0[c] + tb * d[c];
    } else {
        T s_hit, x[3];
        const int k = curve_pass<T>(P, p0, D0, p1, D1, T(0), T(1), kHermiteSegments, s_max, s_hit, x, inst, slot);
        if (k < 0) return false;
        T lo = T(k) / T(kHermiteSegments), hi = T(k + 1) / T(kHermiteSegments);
        {
            T s2, x2[3];
            uint32_t i2 = 0, sl2 = 0;
            const int j = curve_pass<T>(P, p0, D0, p1, D1, lo, hi, kHermiteFine, s_max, s2, x2, i2, sl2);
            if (j >= 0) {
                const T ds = (hi - lo) / T(kHermiteFine);
                lo = lo + T(j) * ds;
                hi = lo + ds;
                s_hit = s2;
                x[0] = x2[0]; x[1] = x2[1]; x[2] = x2[2];
                inst = i2;
                slot = sl2;
            }
        }
        // Newton on the triangle's world plane n_w . x + c = 0, n_w = N (e1 x e2),
        // c = (e1 x e2) . (t - v0) with x_obj = R x + t: the plane in object space
        const float4* __restrict__ I = P.mesh_inst + uint64_t(kInstF4) * inst;
        const float4* __restrict__ tri = P.mesh_tris + 3u * uint64_t(slot);
        const float4 A = tri[0], E1 = tri[1], E2 = tri[2];
        const T e1[3] = {T(E1.x), T(E1.y), T(E1.z)}, e2[3] = {T(E2.x), T(E2.y), T(E2.z)};
        const T nx = e1[1] * e2[2] - e1[2] * e2[1], ny = e1[2] * e2[0] - e1[0] * e2[2], nz = e1[0] * e2[1] - e1[1] * e2[0];
        const float4 r0 = I[0], r1 = I[1], r2 = I[2];
        const T cc = nx * (T(r0.w) - T(A.x)) + ny * (T(r1.w) - T(A.y)) + nz * (T(r2.w) - T(A.z));
        const T wx = T(r0.x) * nx + T(r1.x) * ny + T(r2.x) * nz;
        const T wy = T(r0.y) * nx + T(r1.y) * ny + T(r2.y) * nz;
        const T wz = T(r0.z) * nx + T(r1.z) * ny + T(r2.z) * nz;
        T s = s_hit;
        for (int it = 0; it < 3; ++it) {
            const T f = wx * hermite(p0[0], D0[0], p1[0], D1[0], s) + wy * hermite(p0[1], D0[1], p1[1], D1[1], s) +
                        wz * hermite(p0[2], D0[2], p1[2], D1[2], s) + cc;
            const T fd = wx * hermite_d(p0[0], D0[0], p1[0], D1[0], s) + wy * hermite_d(p0[1], D0[1], p1[1], D1[1], s) +
                         wz * hermite_d(p0[2], D0[2], p1[2], D1[2], s);
            if (fd == T(0)) break;
            s = min_(max_(s - f / fd, lo), hi);
        }
        T xc[3];
        for (int c = 0; c < 3; ++c) xc[c] = hermite(p0[c], D0[c], p1[c], D1[c], s);
        // barycentric coordinates of the curve's crossing in object space
        const T w[3] = {T(r0.x) * xc[0] + T(r0.y) * xc[1] + T(r0.z) * xc[2] + T(r0.w) - T(A.x),
                        T(r1.x) * xc[0] + T(r1.y) * xc[1] + T(r1.z) * xc[2] + T(r1.w) - T(A.y),
                        T(r2.x) * xc[0] + T(r2.y) * xc[1] + T(r2.z) * xc[2] + T(r2.w) - T(A.z)};
        const T d00 = e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2], d01 = e1[0] * e2[0] + e1[1] * e2[1] + e1[2] * e2[2];
        const T d11 = e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2];
        const T d20 = w[0] * e1[0] + w[1] * e1[1] + w[2] * e1[2], d21 = w[0] * e2[0] + w[1] * e2[1] + w[2] * e2[2];
        const T den = d00 * d11 - d01 * d01;
        const T bu = (d11 * d20 - d01 * d21) / den, bv = (d00 * d21 - d01 * d20) / den;
        const T eps = T(kBaryEps);
        if (bu >= -eps && bv >= -eps && bu + bv <= T(1) + eps && s < s_max) {
            h.s = s;
            h.x[0] = xc[0]; h.x[1] = xc[1]; h.x[2] = xc[2];
        } else {
            h.s = s_hit;
            h.x[0] = x[0]; h.x[1] = x[1]; h.x[2] = x[2];
        }
    }
    const float4* __restrict__ I = P.mesh_inst + uint64_t(kInstF4) * inst;
    h.inst = inst;
    h.tri = __float_as_uint(P.mesh_tris[3u * uint64_t(slot)].w);
    h.matter = __float_as_uint(I[5].w);
    return true;
}

}  // namespace mgk
