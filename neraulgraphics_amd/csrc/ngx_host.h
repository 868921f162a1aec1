// Host-side pieces of the backend that need no device: bottom-level tree build and input checks.
#pragma once
#include <stdint.h>
#include <vector>
#include "ngx_device.h"

namespace ngx::host {

struct Box {
    float lo[3];
    float hi[3];
};

struct MeshInput {
    int mesh_count;
    const float *vertices;
    const int32_t *vertex_offsets;
    const uint32_t *indices;
    const int32_t *index_offsets;
    const float *normals;
    const uint8_t *has_normals;
};

struct BuiltMeshes {
    std::vector<BlasNode> nodes;      // the reference's layout, global indices
    std::vector<float> triangles;     // 12 floats each: v0, e1, e2
    std::vector<float> attributes;    // 16 floats each, empty unless some mesh has normals
    std::vector<int> roots;           // per mesh, in nodes
    std::vector<Box> bounds;          // per mesh, in its own space
    std::vector<WideNode> wide;       // the traversal layout of nodes' interior nodes
    std::vector<int> wide_parent;     // per wide node: its parent wide node, -1 at a mesh's root
    std::vector<int> wide_node;       // per wide node: its node in `nodes` (and `skip`)
    std::vector<MeshRoot> mesh_roots; // per mesh, for the tracer
    std::vector<SkipNode> skip;       // per node of `nodes`: the per-lane walk's layout
    std::vector<int> tri_counts;      // per mesh
    // for refitting a mesh to new vertices on the device (ngx_set_mesh_vertices)
    std::vector<int> tri_verts;       // per packed triangle: its three vertices, global indices
    std::vector<int> node_parent;     // per node of `nodes`: its parent, -1 at a mesh's root
    std::vector<int> node_wide;       // per node of `nodes`: its wide node, -1 for a leaf
    std::vector<int> node_base, node_count, tri_base;  // per mesh
};

// Binned SAH build (src/bvh.cpp:46-174). Returns the node count.
int sah_build(const Box *bounds, int n, int max_leaf, std::vector<BlasNode> &nodes, std::vector<uint32_t> &order);

// Every mesh's tree and packed triangles (src/mesh_build.cpp:91-157), plus the traversal layout.
void build_meshes(const MeshInput &in, BuiltMeshes &out);

// Finite and invertible, relative to the matrix's own size (src/mesh_build.cpp:159-183 acceptance).
bool transform_ok(const float *m12);

}  // namespace ngx::host
