// Deforming meshes: a mesh's bottom level refitted on the GPU to new vertex positions, keeping its tree.
//
// The tree's topology (built once on the host with the reference's binned SAH) is kept; every node's box is
// recomputed from the new vertices exactly as the host computes a box (the min and max over the triangles'
// vertices, then unions upward), so the boxes are the ones a host refit of the same tree would give. The
// packed triangles (v0, e1, e2) are recomputed the same way. Three launches: triangles and the arrival
// counters; leaf boxes climbed to the root, the second child to arrive at a node making its box; then the
// boxes copied into the packet walk's wide nodes and the mesh's root record and bounds. No reference
// counterpart: the reference builds bottom levels once, on the host.
#include "ngx_kernels.h"

namespace ngx {
namespace {

__global__ __launch_bounds__(256) void refit_prepare_kernel(const MeshRefitArgs A) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < A.node_count) A.arrivals[A.node_base + i] = 0;
    if (i < A.tri_count) {
        const size_t t = (size_t)A.tri_base + i;
        const float *a = A.vertices + 3 * (size_t)(A.tri_verts[3 * t] - A.vertex_base);
        const float *b = A.vertices + 3 * (size_t)(A.tri_verts[3 * t + 1] - A.vertex_base);
        const float *c = A.vertices + 3 * (size_t)(A.tri_verts[3 * t + 2] - A.vertex_base);
        float *o = A.tris + 9 * t;
        o[0] = a[0];
        o[1] = a[1];
        o[2] = a[2];
        o[3] = b[0] - a[0];
        o[4] = b[1] - a[1];
        o[5] = b[2] - a[2];
        o[6] = c[0] - a[0];
        o[7] = c[1] - a[1];
        o[8] = c[2] - a[2];
    }
}

__global__ __launch_bounds__(256) void refit_boxes_kernel(const MeshRefitArgs A) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= A.node_count) return;
    int node = A.node_base + i;
    SkipNode &leaf = A.skip[node];
    if (leaf.meta <= 0) return;  // interior nodes are made by the climb
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int t = leaf.a; t < leaf.a + leaf.meta; ++t)
        for (int k = 0; k < 3; ++k) {
            const float *p = A.vertices + 3 * (size_t)(A.tri_verts[3 * (size_t)t + k] - A.vertex_base);
            for (int ax = 0; ax < 3; ++ax) {
                lo[ax] = fminf(lo[ax], p[ax]);
                hi[ax] = fmaxf(hi[ax], p[ax]);
            }
        }
    for (int ax = 0; ax < 3; ++ax) {
        leaf.lo[ax] = lo[ax];
        leaf.hi[ax] = hi[ax];
    }
    while (true) {
        const int p = A.node_parent[node];
        if (p < 0) break;
        // the second child to arrive makes the parent's box; acquire-release orders the boxes written before
        const int prior = __hip_atomic_fetch_add(&A.arrivals[p], 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (prior == 0) break;
        SkipNode &n = A.skip[p];
        const SkipNode &l = A.skip[n.a], &r = A.skip[n.a + 1];
        for (int ax = 0; ax < 3; ++ax) {
            n.lo[ax] = fminf(l.lo[ax], r.lo[ax]);
            n.hi[ax] = fmaxf(l.hi[ax], r.hi[ax]);
        }
        node = p;
    }
}

__global__ __launch_bounds__(256) void refit_layout_kernel(const MeshRefitArgs A) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= A.node_count) return;
    const int node = A.node_base + i;
    const SkipNode &n = A.skip[node];
    const int w = A.node_wide[node];
    if (w >= 0) {
        const SkipNode &l = A.skip[n.a], &r = A.skip[n.a + 1];
        WideNode &o = A.wide[w];
        for (int ax = 0; ax < 3; ++ax) {
            o.l_lo[ax] = l.lo[ax];
            o.l_hi[ax] = l.hi[ax];
            o.r_lo[ax] = r.lo[ax];
            o.r_hi[ax] = r.hi[ax];
        }
    }
    if (node == A.root) {
        for (int ax = 0; ax < 3; ++ax) {
            A.mesh_root->lo[ax] = n.lo[ax];
            A.mesh_root->hi[ax] = n.hi[ax];
            A.mesh_bounds[ax] = n.lo[ax];
            A.mesh_bounds[3 + ax] = n.hi[ax];
        }
    }
}

}  // namespace

hipError_t launch_mesh_refit(const MeshRefitArgs &a, hipStream_t stream) {
    if (a.node_count <= 0 || a.tri_count <= 0) return hipErrorInvalidValue;
    const int n = a.node_count > a.tri_count ? a.node_count : a.tri_count;
    refit_prepare_kernel<<<(n + 255) / 256, 256, 0, stream>>>(a);
    refit_boxes_kernel<<<(a.node_count + 255) / 256, 256, 0, stream>>>(a);
    refit_layout_kernel<<<(a.node_count + 255) / 256, 256, 0, stream>>>(a);
    return hipGetLastError();
}

}  // namespace ngx
