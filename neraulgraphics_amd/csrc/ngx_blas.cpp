// Host-side bottom-level trees: one binned-SAH BVH per mesh, built once when a context is created, then
// laid out for the tracer.
//
// The tree is the reference's (src/bvh.cpp:46-174 through src/mesh_build.cpp:91-157): 16 bins, at most
// 4 triangles a leaf, a depth limit of 60, the same cost comparisons and the same std::partition, so
// the node array, the triangle order and therefore the order in which a ray meets triangles are the
// reference's. Triangles are packed as v0, e1 = v1 - v0, e2 = v2 - v0 in leaf order. After that the
// tree is re-laid as 64-byte wide nodes (both children's boxes in the parent) for the wavefront walk.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>
#include "ngx_host.h"

namespace ngx::host {
namespace {

constexpr int BINS = 16;
constexpr int MAX_DEPTH = 60;

Box empty_box() {
    const float inf = std::numeric_limits<float>::infinity();
    return Box{{inf, inf, inf}, {-inf, -inf, -inf}};
}

void grow(Box &b, const Box &o) {
    for (int k = 0; k < 3; ++k) {
        b.lo[k] = std::min(b.lo[k], o.lo[k]);
        b.hi[k] = std::max(b.hi[k], o.hi[k]);
    }
}

float area(const Box &b) {
    const float dx = b.hi[0] - b.lo[0], dy = b.hi[1] - b.lo[1], dz = b.hi[2] - b.lo[2];
    if (dx < 0.0f || dy < 0.0f || dz < 0.0f) return 0.0f;
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}

struct Job {
    int node;
    uint32_t begin, end;
    int depth;
};

}  // namespace

int sah_build(const Box *bounds, const int n, const int max_leaf, std::vector<BlasNode> &nodes, std::vector<uint32_t> &order) {
    nodes.clear();
    order.resize(static_cast<size_t>(n));
    if (n <= 0) return 0;
    std::iota(order.begin(), order.end(), 0u);
    std::vector<float> centre(3 * static_cast<size_t>(n));
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) centre[3 * static_cast<size_t>(i) + k] = 0.5f * (bounds[i].lo[k] + bounds[i].hi[k]);
    nodes.reserve(2 * static_cast<size_t>(n));
    nodes.emplace_back();
    std::vector<Job> jobs{{0, 0, static_cast<uint32_t>(n), 0}};
    auto bin_of = [&](const uint32_t prim, const int axis, const float lo, const float scale) {
        return std::min(BINS - 1, static_cast<int>((centre[3 * static_cast<size_t>(prim) + axis] - lo) * scale));
    };
    while (!jobs.empty()) {
        const Job job = jobs.back();
        jobs.pop_back();
        const uint32_t count = job.end - job.begin;
        uint32_t *prims = order.data() + job.begin;
        Box box = empty_box(), cbox = empty_box();
        for (uint32_t i = 0; i < count; ++i) {
            grow(box, bounds[prims[i]]);
            for (int k = 0; k < 3; ++k) {
                cbox.lo[k] = std::min(cbox.lo[k], centre[3 * static_cast<size_t>(prims[i]) + k]);
                cbox.hi[k] = std::max(cbox.hi[k], centre[3 * static_cast<size_t>(prims[i]) + k]);
            }
        }
        BlasNode node{};
        for (int k = 0; k < 3; ++k) {
            node.lo[k] = box.lo[k];
            node.hi[k] = box.hi[k];
        }
        int best_axis = -1, best_bin = 0;
        float best_cost = std::numeric_limits<float>::infinity();
        if (count > 1) {
            for (int axis = 0; axis < 3; ++axis) {
                const float extent = cbox.hi[axis] - cbox.lo[axis];
                if (!(extent > 0.0f)) continue;
                const float scale = static_cast<float>(BINS) / extent;
                int bin_count[BINS] = {};
                Box bin_box[BINS];
                for (Box &b : bin_box) b = empty_box();
                for (uint32_t i = 0; i < count; ++i) {
                    const int b = bin_of(prims[i], axis, cbox.lo[axis], scale);
                    ++bin_count[b];
                    grow(bin_box[b], bounds[prims[i]]);
                }
                float right_area[BINS - 1];
                int right_count[BINS - 1];
                Box acc = empty_box();
                int running = 0;
                for (int i = BINS - 1; i >= 1; --i) {
                    grow(acc, bin_box[i]);
                    running += bin_count[i];
                    right_area[i - 1] = area(acc);
                    right_count[i - 1] = running;
                }
                acc = empty_box();
                running = 0;
                for (int i = 0; i < BINS - 1; ++i) {
                    grow(acc, bin_box[i]);
                    running += bin_count[i];
                    if (running == 0 || right_count[i] == 0) continue;
                    const float cost = static_cast<float>(running) * area(acc) + static_cast<float>(right_count[i]) * right_area[i];
                    if (cost < best_cost) {
                        best_cost = cost;
                        best_axis = axis;
                        best_bin = i;
                    }
                }
            }
        }
        const float leaf_cost = static_cast<float>(count) * area(box);
        const float split_cost = area(box) + best_cost;
        const bool small = count <= static_cast<uint32_t>(max_leaf);
        if (count == 1 || job.depth >= MAX_DEPTH || (small && (best_axis < 0 || leaf_cost <= split_cost))) {
            node.a = static_cast<int32_t>(job.begin);
            node.count = static_cast<int32_t>(count);
            nodes[job.node] = node;
            continue;
        }
        uint32_t mid;
        if (best_axis >= 0) {
            const float extent = cbox.hi[best_axis] - cbox.lo[best_axis];
            const float scale = static_cast<float>(BINS) / extent;
            const uint32_t *p = std::partition(prims, prims + count,
                [&](const uint32_t prim) { return bin_of(prim, best_axis, cbox.lo[best_axis], scale) <= best_bin; });
            mid = job.begin + static_cast<uint32_t>(p - prims);
            if (mid == job.begin || mid == job.end) mid = job.begin + count / 2;
        } else {
            mid = job.begin + count / 2;
        }
        const int left = static_cast<int>(nodes.size());
        nodes.resize(nodes.size() + 2);
        node.a = left;
        node.count = 0;
        nodes[job.node] = node;
        jobs.push_back({left + 1, mid, job.end, job.depth + 1});
        jobs.push_back({left, job.begin, mid, job.depth + 1});
    }
    return static_cast<int>(nodes.size());
}

void build_meshes(const MeshInput &in, BuiltMeshes &out) {
    out = BuiltMeshes{};
    bool any_normals = false;
    for (int m = 0; m < in.mesh_count; ++m) any_normals = any_normals || (in.normals && in.has_normals && in.has_normals[m]);
    for (int m = 0; m < in.mesh_count; ++m) {
        const float *v = in.vertices + 3 * static_cast<size_t>(in.vertex_offsets[m]);
        const uint32_t *ix = in.indices + in.index_offsets[m];
        const int tri_count = (in.index_offsets[m + 1] - in.index_offsets[m]) / 3;
        std::vector<Box> tb(static_cast<size_t>(tri_count));
        for (int t = 0; t < tri_count; ++t) {
            Box b = empty_box();
            for (int k = 0; k < 3; ++k) {
                const float *p = v + 3 * static_cast<size_t>(ix[3 * t + k]);
                for (int a = 0; a < 3; ++a) {
                    b.lo[a] = std::fmin(b.lo[a], p[a]);
                    b.hi[a] = std::fmax(b.hi[a], p[a]);
                }
            }
            tb[t] = b;
        }
        std::vector<BlasNode> tree;
        std::vector<uint32_t> order;
        sah_build(tb.data(), tri_count, 4, tree, order);
        const int node_base = static_cast<int>(out.nodes.size());
        const int tri_base = static_cast<int>(out.triangles.size() / 12);
        out.roots.push_back(node_base);
        out.node_base.push_back(node_base);
        out.node_count.push_back(static_cast<int>(tree.size()));
        out.tri_base.push_back(tri_base);
        out.bounds.push_back(Box{{tree[0].lo[0], tree[0].lo[1], tree[0].lo[2]}, {tree[0].hi[0], tree[0].hi[1], tree[0].hi[2]}});
        for (BlasNode node : tree) {
            node.a += node.count > 0 ? tri_base : node_base;
            out.nodes.push_back(node);
        }
        for (const uint32_t t : order) {
            for (int k = 0; k < 3; ++k) out.tri_verts.push_back(in.vertex_offsets[m] + static_cast<int>(ix[3 * t + k]));
            const float *a = v + 3 * static_cast<size_t>(ix[3 * t]);
            const float *b = v + 3 * static_cast<size_t>(ix[3 * t + 1]);
            const float *c = v + 3 * static_cast<size_t>(ix[3 * t + 2]);
            const float packed[12] = {a[0], a[1], a[2], 0.0f, b[0] - a[0], b[1] - a[1], b[2] - a[2], 0.0f,
                                      c[0] - a[0], c[1] - a[1], c[2] - a[2], 0.0f};
            out.triangles.insert(out.triangles.end(), packed, packed + 12);
            if (any_normals) {
                float rec[16] = {};
                const bool has = in.normals && in.has_normals && in.has_normals[m];
                if (has) {
                    const float *nrm = in.normals + 3 * static_cast<size_t>(in.vertex_offsets[m]);
                    for (int k = 0; k < 3; ++k)
                        for (int a2 = 0; a2 < 3; ++a2) rec[4 * k + a2] = nrm[3 * static_cast<size_t>(ix[3 * t + k]) + a2];
                }
                const int32_t flags = has ? 1 : 0;
                std::memcpy(&rec[15], &flags, sizeof(flags));
                out.attributes.insert(out.attributes.end(), rec, rec + 16);
            }
        }
    }
    // The traversal layout: a wide node for every interior node, in the same order.
    std::vector<int> wide_index(out.nodes.size(), -1);
    int wide_count = 0;
    for (size_t i = 0; i < out.nodes.size(); ++i)
        if (out.nodes[i].count == 0) wide_index[i] = wide_count++;
    out.wide.resize(static_cast<size_t>(wide_count));
    out.node_wide = wide_index;
    out.node_parent.assign(out.nodes.size(), -1);
    for (size_t i = 0; i < out.nodes.size(); ++i)
        if (out.nodes[i].count == 0) out.node_parent[out.nodes[i].a] = out.node_parent[out.nodes[i].a + 1] = static_cast<int>(i);
    out.wide_node.assign(static_cast<size_t>(wide_count), -1);
    for (size_t i = 0; i < out.nodes.size(); ++i)
        if (wide_index[i] >= 0) out.wide_node[wide_index[i]] = static_cast<int>(i);
    auto child_ref = [&](const BlasNode &c, const int index, int32_t &ref, int32_t &count) {
        if (c.count > 0) {
            ref = ~c.a;
            count = c.count;
        } else {
            ref = wide_index[index];
            count = 0;
        }
    };
    for (size_t i = 0; i < out.nodes.size(); ++i) {
        const BlasNode &n = out.nodes[i];
        if (n.count != 0) continue;
        WideNode &w = out.wide[wide_index[i]];
        const BlasNode &l = out.nodes[n.a], &r = out.nodes[n.a + 1];
        for (int k = 0; k < 3; ++k) {
            w.l_lo[k] = l.lo[k];
            w.l_hi[k] = l.hi[k];
            w.r_lo[k] = r.lo[k];
            w.r_hi[k] = r.hi[k];
        }
        child_ref(l, n.a, w.left, w.left_count);
        child_ref(r, n.a + 1, w.right, w.right_count);
    }
    out.wide_parent.assign(out.wide.size(), -1);
    for (size_t i = 0; i < out.wide.size(); ++i) {
        if (out.wide[i].left >= 0) out.wide_parent[out.wide[i].left] = (int)i;
        if (out.wide[i].right >= 0) out.wide_parent[out.wide[i].right] = (int)i;
    }
    // The per-lane layout: every node with its octant-ordered escapes (the next node after its subtree).
    out.skip.assign(out.nodes.size(), SkipNode{});
    for (size_t i = 0; i < out.nodes.size(); ++i) {
        const BlasNode &n = out.nodes[i];
        SkipNode &k = out.skip[i];
        for (int ax = 0; ax < 3; ++ax) {
            k.lo[ax] = n.lo[ax];
            k.hi[ax] = n.hi[ax];
        }
        k.a = n.a;
        if (n.count > 0) {
            k.meta = n.count;
        } else {
            const BlasNode &l = out.nodes[n.a], &r = out.nodes[n.a + 1];
            int axis = 0;
            float sep = -1.0f;
            for (int ax = 0; ax < 3; ++ax) {
                const float d = std::fabs((l.lo[ax] + l.hi[ax]) - (r.lo[ax] + r.hi[ax]));
                if (d > sep) {
                    sep = d;
                    axis = ax;
                }
            }
            const int flip = (l.lo[axis] + l.hi[axis]) > (r.lo[axis] + r.hi[axis]) ? 1 : 0;
            k.meta = -(1 + 2 * axis + flip);
        }
    }
    for (int o = 0; o < 8; ++o) {
        std::vector<std::pair<int, int>> todo;  // (node, its escape)
        for (const int root : out.roots) todo.push_back({root, -1});
        while (!todo.empty()) {
            const auto [node, esc] = todo.back();
            todo.pop_back();
            out.skip[node].escape[o] = esc;
            const int meta = out.skip[node].meta;
            if (meta > 0) continue;
            const int code = -meta - 1, axis = code >> 1, flip = code & 1;
            const bool positive = (o >> axis) & 1;
            const int left = out.skip[node].a, right = left + 1;
            const int first = (positive != (flip == 1)) ? left : right, second = first == left ? right : left;
            todo.push_back({second, esc});
            todo.push_back({first, second});
        }
    }
    for (int m = 0; m < in.mesh_count; ++m) out.tri_counts.push_back((in.index_offsets[m + 1] - in.index_offsets[m]) / 3);
    for (size_t m = 0; m < out.roots.size(); ++m) {
        const BlasNode &root = out.nodes[out.roots[m]];
        MeshRoot mr{};
        for (int k = 0; k < 3; ++k) {
            mr.lo[k] = root.lo[k];
            mr.hi[k] = root.hi[k];
        }
        child_ref(root, out.roots[m], mr.ref, mr.count);
        out.mesh_roots.push_back(mr);
    }
}

bool transform_ok(const float *m) {
    for (int i = 0; i < 12; ++i)
        if (!std::isfinite(m[i])) return false;
    const float c00 = m[5] * m[10] - m[6] * m[9];
    const float c01 = m[6] * m[8] - m[4] * m[10];
    const float c02 = m[4] * m[9] - m[5] * m[8];
    const float det = m[0] * c00 + m[1] * c01 + m[2] * c02;
    float mag = 0.0f;
    for (const int i : {0, 1, 2, 4, 5, 6, 8, 9, 10}) mag = std::fmax(mag, std::fabs(m[i]));
    return std::isfinite(det) && std::fabs(det) > 1e-12f * mag * mag * mag;
}

}  // namespace ngx::host
