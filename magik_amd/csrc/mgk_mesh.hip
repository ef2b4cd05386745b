// Host build of the two-level mesh hierarchy traced by mgk_mesh.hpp.
//
// Bottom level: one binned-SAH bounding volume hierarchy per mesh (Wald,
// "On fast construction of SAH-based bounding volume hierarchies", IEEE
// RT 2007: 16 centroid bins on the widest axis, cost = area x count), leaves
// of at most kLeafMax triangles.  Top level: the same build over the
// instances' world bounds, one instance per leaf.  Both are emitted in
// depth-first order with skip links, so the device walks them without a stack.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <vector>

#include "mgk_internal.hpp"

namespace mgk {
namespace {

constexpr uint32_t kLeafMax = 8;   // the traversal is bound by its dependent node loads: on box_mesh, 8 beat 2, 4 and 16
constexpr int kBins = 16;
constexpr int kMaxSahDepth = 48;   // deeper: median splits only, which bound the depth by log2(n)

struct Box3 {
    float lo[3] = {INFINITY, INFINITY, INFINITY};
    float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    void grow(const float p[3]) {
        for (int c = 0; c < 3; ++c) { lo[c] = std::min(lo[c], p[c]); hi[c] = std::max(hi[c], p[c]); }
    }
    void grow(const Box3& b) {
        for (int c = 0; c < 3; ++c) { lo[c] = std::min(lo[c], b.lo[c]); hi[c] = std::max(hi[c], b.hi[c]); }
    }
    double area() const {
        if (!(hi[0] >= lo[0])) return 0.0;
        const double x = double(hi[0]) - lo[0], y = double(hi[1]) - lo[1], z = double(hi[2]) - lo[2];
        return 2.0 * (x * y + y * z + z * x);
    }
};

struct Prim {
    Box3 box;
    float c[3];      // centroid
    uint32_t id;     // triangle index in the mesh, or instance index
};

struct Node {
    Box3 box;
    uint32_t skip = kBvhEnd, info = 0;
};

// Depth-first build of prims[b, e) into nodes; returns nothing, fills skip
// links with the index one past the subtree (patched to kBvhEnd by the caller
// at the end of the hierarchy).  Leaves reference prims by position
// (`by_position`) or hold the prim's id (top level, one per leaf).
void build_node(std::vector<Prim>& prims, uint32_t b, uint32_t e, uint32_t leaf_max, bool by_position, std::vector<Node>& nodes,
                int depth = 0) {
    const uint32_t me = uint32_t(nodes.size());
    nodes.emplace_back();
    Box3 box, cbox;
    for (uint32_t i = b; i < e; ++i) { box.grow(prims[i].box); cbox.grow(prims[i].c); }
    nodes[me].box = box;
    const uint32_t n = e - b;
    auto make_leaf = [&]() {
        nodes[me].info = by_position ? ((n << kLeafShift) | b) : ((1u << kLeafShift) | prims[b].id);
        nodes[me].skip = uint32_t(nodes.size());
    };
    if (n <= leaf_max) { make_leaf(); return; }
    int axis = 0;
    for (int c = 1; c < 3; ++c)
        if (double(cbox.hi[c]) - cbox.lo[c] > double(cbox.hi[axis]) - cbox.lo[axis]) axis = c;
    const double ext = double(cbox.hi[axis]) - cbox.lo[axis];
    uint32_t mid = b + n / 2;
    if (ext > 0.0 && depth < kMaxSahDepth) {
        std::array<Box3, kBins> bb;
        std::array<uint32_t, kBins> bn{};
        auto bin_of = [&](const Prim& p) {
            int k = int(double(kBins) * (double(p.c[axis]) - cbox.lo[axis]) / ext);
            return std::min(std::max(k, 0), kBins - 1);
        };
        for (uint32_t i = b; i < e; ++i) { const int k = bin_of(prims[i]); bb[k].grow(prims[i].box); ++bn[k]; }
        std::array<double, kBins> left_cost{};
        Box3 acc;
        uint32_t cnt = 0;
        for (int k = 0; k < kBins - 1; ++k) { acc.grow(bb[k]); cnt += bn[k]; left_cost[k] = acc.area() * cnt; }
        acc = Box3();
        cnt = 0;
        double best = INFINITY;
        int split = -1;
        for (int k = kBins - 1; k > 0; --k) {
            acc.grow(bb[k]);
            cnt += bn[k];
            const double cost = left_cost[k - 1] + acc.area() * cnt;
            if (cost < best) { best = cost; split = k; }
        }
        if (split > 0) {
            auto it = std::partition(prims.begin() + b, prims.begin() + e, [&](const Prim& p) { return bin_of(p) < split; });
            mid = uint32_t(it - prims.begin());
            if (mid == b || mid == e) mid = b + n / 2;
        }
    }
    if (mid == b + n / 2 && ext > 0.0) {
        // no useful bin split: median on the axis
        std::nth_element(prims.begin() + b, prims.begin() + mid, prims.begin() + e,
                         [&](const Prim& x, const Prim& y) { return x.c[axis] < y.c[axis]; });
    }
    build_node(prims, b, mid, leaf_max, by_position, nodes, depth + 1);
    build_node(prims, mid, e, leaf_max, by_position, nodes, depth + 1);
    nodes[me].skip = uint32_t(nodes.size());
}

// Appends a hierarchy to the global node array with its indices offset by the
// array's size; skip links past the hierarchy become kBvhEnd.  Returns the root.
uint32_t append(std::vector<Node>& local, std::vector<float4>& out) {
    const uint32_t base = uint32_t(out.size() / 2), end = uint32_t(local.size());
    for (const Node& nd : local) {
        const uint32_t skip = nd.skip >= end ? kBvhEnd : nd.skip + base;
        float4 a, b;
        a.x = nd.box.lo[0]; a.y = nd.box.lo[1]; a.z = nd.box.lo[2];
        b.x = nd.box.hi[0]; b.y = nd.box.hi[1]; b.z = nd.box.hi[2];
        std::memcpy(&a.w, &skip, 4);
        std::memcpy(&b.w, &nd.info, 4);
        out.push_back(a);
        out.push_back(b);
    }
    return base;
}

float4 f4(float x, float y, float z, float w) {
    float4 v;
    v.x = x; v.y = y; v.z = z; v.w = w;
    return v;
}
float f_of(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// Inverse of a 3 x 4 affine map in double; false if singular.
bool invert_affine(const float m[12], double out[12]) {
    const double a = m[0], b = m[1], c = m[2], d = m[4], e = m[5], f = m[6], g = m[8], h = m[9], i = m[10];
    const double A = e * i - f * h, B = -(d * i - f * g), C = d * h - e * g;
    const double det = a * A + b * B + c * C;
    if (!(std::fabs(det) > 0.0) || !std::isfinite(det)) return false;
    const double r[9] = {A / det, -(b * i - c * h) / det, (b * f - c * e) / det,
                         B / det, (a * i - c * g) / det, -(a * f - c * d) / det,
                         C / det, -(a * h - b * g) / det, (a * e - b * d) / det};
    for (int row = 0; row < 3; ++row) {
        out[4 * row + 0] = r[3 * row + 0];
        out[4 * row + 1] = r[3 * row + 1];
        out[4 * row + 2] = r[3 * row + 2];
        out[4 * row + 3] = -(r[3 * row + 0] * m[3] + r[3 * row + 1] * m[7] + r[3 * row + 2] * m[11]);
    }
    return true;
}

// Rounds a double bound outward to float.
float down(double x) {
    float f = float(x);
    return double(f) > x ? std::nextafter(f, -INFINITY) : f;
}
float up(double x) {
    float f = float(x);
    return double(f) < x ? std::nextafter(f, INFINITY) : f;
}

}  // namespace

e_magik_result_types build_meshes(const mgk_mesh_t* meshes, uint32_t n_mesh, const mgk_instance_t* inst, uint32_t n_inst,
                                  MeshBuild& out) {
    out = MeshBuild();
    if (n_mesh == 0) return MAGIK_SUCCESS;
    if (!meshes || (n_inst > 0 && !inst)) return MAGIK_ERROR_INVALID_POINTER;
    if (n_inst == 0 || n_inst > MGK_MAX_INSTANCES) return MGK_ERROR_INVALID_ARGUMENT;
    uint64_t total = 0;
    for (uint32_t m = 0; m < n_mesh; ++m) {
        const mgk_mesh_t& M = meshes[m];
        if (M.n_triangle == 0 || M.n_triangle > MGK_MAX_MESH_TRIANGLES) return MGK_ERROR_INVALID_ARGUMENT;
        if (!M.vertices || !M.indices) return MAGIK_ERROR_INVALID_POINTER;
        total += M.n_triangle;
    }
    if (total > MGK_MAX_MESH_TRIANGLES) return MGK_ERROR_CAPACITY_EXCEEDED;
    std::vector<uint32_t> root(n_mesh), normal_base(n_mesh);
    std::vector<Box3> mesh_box(n_mesh);
    for (uint32_t m = 0; m < n_mesh; ++m) {
        const mgk_mesh_t& M = meshes[m];
        std::vector<Prim> prims(M.n_triangle);
        normal_base[m] = uint32_t(out.normals.size());
        for (uint32_t t = 0; t < M.n_triangle; ++t) {
            const float* v[3];
            for (int k = 0; k < 3; ++k) {
                const uint32_t vi = M.indices[3u * t + k];
                if (vi >= M.n_vertex) return MGK_ERROR_INVALID_ARGUMENT;
                v[k] = M.vertices + 3u * vi;
            }
            Prim& p = prims[t];
            for (int k = 0; k < 3; ++k) p.box.grow(v[k]);
            for (int c = 0; c < 3; ++c) p.c[c] = float((double(v[0][c]) + v[1][c] + v[2][c]) / 3.0);
            p.id = t;
            const double e1[3] = {double(v[1][0]) - v[0][0], double(v[1][1]) - v[0][1], double(v[1][2]) - v[0][2]};
            const double e2[3] = {double(v[2][0]) - v[0][0], double(v[2][1]) - v[0][1], double(v[2][2]) - v[0][2]};
            double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
            const double l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            if (l > 0.0) { n[0] /= l; n[1] /= l; n[2] /= l; }
            out.normals.push_back(f4(float(n[0]), float(n[1]), float(n[2]), 0.0f));
        }
        std::vector<Node> local;
        local.reserve(2 * M.n_triangle / kLeafMax + 2);
        build_node(prims, 0, M.n_triangle, kLeafMax, true, local);
        // leaves refer to prims by position: offset them by the mesh's first slot
        const uint32_t slot0 = uint32_t(out.tris.size() / 3);
        for (Node& nd : local)
            if (nd.info) nd.info = (nd.info & ~kLeafMask) | ((nd.info & kLeafMask) + slot0);
        for (const Prim& p : prims) {
            const uint32_t t = p.id;
            const float* v0 = M.vertices + 3u * M.indices[3u * t];
            const float* v1 = M.vertices + 3u * M.indices[3u * t + 1];
            const float* v2 = M.vertices + 3u * M.indices[3u * t + 2];
            out.tris.push_back(f4(v0[0], v0[1], v0[2], f_of(t)));
            out.tris.push_back(f4(v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2], 0.0f));
            out.tris.push_back(f4(v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2], 0.0f));
        }
        mesh_box[m] = local[0].box;
        root[m] = append(local, out.nodes);
    }
    out.mesh_roots = root;
    // instances: records and world bounds
    std::vector<Prim> prims(n_inst);
    Box3 world;
    for (uint32_t k = 0; k < n_inst; ++k) {
        const mgk_instance_t& I = inst[k];
        if (I.mesh >= n_mesh) return MGK_ERROR_INVALID_ARGUMENT;
        if (I.matter != MGK_MATTER_DIFFUSE && I.matter != MGK_MATTER_ILLUMINANT) return MGK_ERROR_INVALID_ARGUMENT;
        for (int j = 0; j < 12; ++j)
            if (!std::isfinite(I.object_to_world[j])) return MGK_ERROR_INVALID_ARGUMENT;
        double w2o[12];
        if (!invert_affine(I.object_to_world, w2o)) return MGK_ERROR_INVALID_ARGUMENT;
        out.any_diffuse = out.any_diffuse || I.matter == MGK_MATTER_DIFFUSE;
        for (int r = 0; r < 3; ++r)
            out.inst.push_back(f4(float(w2o[4 * r]), float(w2o[4 * r + 1]), float(w2o[4 * r + 2]), float(w2o[4 * r + 3])));
        const uint32_t meta[3] = {root[I.mesh], normal_base[I.mesh], I.matter};
        for (int r = 0; r < 3; ++r)
            out.inst.push_back(f4(float(w2o[r]), float(w2o[4 + r]), float(w2o[8 + r]), f_of(meta[r])));
        // world bounds: the 8 corners of the mesh's bounds through the object -> world map
        const Box3& mb = mesh_box[I.mesh];
        double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int corner = 0; corner < 8; ++corner) {
            const double x = (corner & 1) ? mb.hi[0] : mb.lo[0], y = (corner & 2) ? mb.hi[1] : mb.lo[1],
                         z = (corner & 4) ? mb.hi[2] : mb.lo[2];
            for (int c = 0; c < 3; ++c) {
                const float* m = I.object_to_world + 4 * c;
                const double w = double(m[0]) * x + double(m[1]) * y + double(m[2]) * z + double(m[3]);
                lo[c] = std::min(lo[c], w);
                hi[c] = std::max(hi[c], w);
            }
        }
        // widened by 1e-6 relative: the device maps segments into object space in float
        Prim& p = prims[k];
        for (int c = 0; c < 3; ++c) {
            const double pad = 1e-6 * std::max({std::fabs(lo[c]), std::fabs(hi[c]), hi[c] - lo[c]});
            p.box.lo[c] = down(lo[c] - pad);
            p.box.hi[c] = up(hi[c] + pad);
            p.c[c] = float(0.5 * (lo[c] + hi[c]));
        }
        p.id = k;
        world.grow(p.box);
    }
    std::vector<Node> top;
    build_node(prims, 0, n_inst, 1, false, top);
    out.root = append(top, out.nodes);
    for (int c = 0; c < 3; ++c) { out.lo[c] = world.lo[c]; out.hi[c] = world.hi[c]; }
    return MAGIK_SUCCESS;
}

}  // namespace mgk

namespace {

// Walks one hierarchy from its root with the skip links, as the device does,
// checking the layout; counts leaf entries per slot.
bool walk(const mgk::MeshBuild& b, uint32_t root, std::vector<uint32_t>& seen, bool top, uint32_t n_inst, uint32_t& depth,
          uint32_t& max_leaf) {
    using namespace mgk;
    auto u = [](float f) { uint32_t x; std::memcpy(&x, &f, 4); return x; };
    const uint64_t n_nodes = b.nodes.size() / 2;
    std::vector<uint32_t> ends;   // skip targets of the open interior nodes (for the depth)
    uint32_t n = root;
    uint64_t steps = 0;
    while (n != kBvhEnd) {
        if (n >= n_nodes || ++steps > n_nodes) return false;
        while (!ends.empty() && n >= ends.back()) ends.pop_back();
        depth = std::max(depth, uint32_t(ends.size()) + 1);
        const float4 lo = b.nodes[2 * uint64_t(n)], hi = b.nodes[2 * uint64_t(n) + 1];
        const uint32_t skip = u(lo.w), info = u(hi.w);
        if (skip != kBvhEnd && skip <= n) return false;
        if (info == 0) {
            // two children inside the parent's box: n + 1, and the node that the
            // first child's skip link reaches, before the parent's own skip target
            const uint64_t end = skip == kBvhEnd ? n_nodes : skip;
            auto inside = [&](uint64_t c) {
                const float4 clo = b.nodes[2 * c], chi = b.nodes[2 * c + 1];
                return !(clo.x < lo.x || clo.y < lo.y || clo.z < lo.z || chi.x > hi.x || chi.y > hi.y || chi.z > hi.z);
            };
            const uint64_t c1 = uint64_t(n) + 1;
            if (c1 >= end || !inside(c1)) return false;
            const uint32_t c2 = u(b.nodes[2 * c1].w);
            if (c2 == kBvhEnd || c2 <= c1 || c2 >= end || !inside(c2)) return false;
            ends.push_back(uint32_t(end));
            ++n;
            continue;
        }
        const uint32_t first = info & kLeafMask, count = info >> kLeafShift;
        if (count == 0 || (top && count != 1)) return false;
        max_leaf = std::max(max_leaf, count);
        for (uint32_t j = first; j < first + count; ++j) {
            if (j >= seen.size() || (top && j >= n_inst)) return false;
            ++seen[j];
        }
        n = skip;
    }
    return true;
}

}  // namespace

MGK_API e_magik_result_types mgk_mesh_hierarchy_info(const mgk_mesh_t* meshes, uint32_t n_mesh, const mgk_instance_t* instances,
                                                     uint32_t n_instance, mgk_mesh_hierarchy_info_t* info) {
    using namespace mgk;
    if (!info) return MAGIK_ERROR_INVALID_POINTER;
    std::memset(info, 0, sizeof(*info));
    MeshBuild b;
    e_magik_result_types r = build_meshes(meshes, n_mesh, instances, n_instance, b);
    if (r != MAGIK_SUCCESS || n_mesh == 0) return r;
    auto u = [](float f) { uint32_t x; std::memcpy(&x, &f, 4); return x; };
    const uint64_t n_slots = b.tris.size() / 3;
    // bottom levels: every mesh's own root, whether an instance uses it or not;
    // every slot in exactly one leaf; each instance names one of those roots
    std::vector<uint32_t> seen(n_slots, 0);
    uint32_t depth = 0, max_leaf = 0, top_depth = 0, top_leaf = 0;
    for (uint32_t root : b.mesh_roots)
        if (!walk(b, root, seen, false, n_instance, depth, max_leaf)) return MGK_ERROR_INVALID_ARGUMENT;
    for (uint32_t c : seen)
        if (c != 1u) return MGK_ERROR_INVALID_ARGUMENT;
    for (uint32_t k = 0; k < n_instance; ++k) {
        const uint32_t root = u(b.inst[uint64_t(kInstF4) * k + 3].w);
        if (std::find(b.mesh_roots.begin(), b.mesh_roots.end(), root) == b.mesh_roots.end()) return MGK_ERROR_INVALID_ARGUMENT;
    }
    std::vector<uint32_t> seen_top(n_instance, 0);
    if (!walk(b, b.root, seen_top, true, n_instance, top_depth, top_leaf)) return MGK_ERROR_INVALID_ARGUMENT;
    for (uint32_t c : seen_top)
        if (c != 1u) return MGK_ERROR_INVALID_ARGUMENT;
    info->n_nodes = b.nodes.size() / 2;
    info->n_triangles = n_slots;
    info->max_depth = depth;
    info->max_leaf = max_leaf;
    for (int c = 0; c < 3; ++c) { info->world_min[c] = b.lo[c]; info->world_max[c] = b.hi[c]; }
    return MAGIK_SUCCESS;
}
