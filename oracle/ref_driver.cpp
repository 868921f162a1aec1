// ref_driver.cpp -- TEST INFRASTRUCTURE. A thin C-ABI harness around the parts of the reference that
// compile from their own sources (see oracle/Makefile): src/bvh.cpp, src/mesh.cpp, and the header-only
// traversal core src/mesh_dev.h + src/device_scene.h + src/render_core.h (camera) compiled for the host
// with NG_BACKEND_CPU, exactly as src/cpu_backend.cpp compiles them. It defines no reference symbol of
// its own; it only calls the reference's functions so the oracle's restatement can be checked against
// them bit for bit. Built into oracle/_ref/, never shipped, never copied into the repository.
#define NG_BACKEND_CPU
#include <cstdint>
#include <cstring>
#include <vector>
#include <neraulgraphics/mesh.h>
#include "bvh.h"
#include "render_core.h"

using namespace neraulgraphics;

extern "C" {

// bvh::build (src/bvh.cpp:46-174). nodes: capacity 2n; returns the node count.
int ref_sah_build(const float *bounds6, int n, int max_leaf, bvh::Node *nodes, uint32_t *order) {
    std::vector<bvh::Aabb> b(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            b[i].lo[k] = bounds6[6 * i + k];
            b[i].hi[k] = bounds6[6 * i + 3 + k];
        }
    const bvh::Bvh t = bvh::build(b, max_leaf);
    std::memcpy(nodes, t.nodes.data(), t.nodes.size() * sizeof(bvh::Node));
    std::memcpy(order, t.order.data(), t.order.size() * sizeof(uint32_t));
    return static_cast<int>(t.nodes.size());
}

// bvh::refit + bvh::sah_cost (src/bvh.cpp:176-203) on a tree given as nodes + order.
float ref_sah_refit_cost(bvh::Node *nodes, int node_count, const uint32_t *order, int n, const float *bounds6) {
    bvh::Bvh t;
    t.nodes.assign(nodes, nodes + node_count);
    t.order.assign(order, order + n);
    std::vector<bvh::Aabb> b(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            b[i].lo[k] = bounds6[6 * i + k];
            b[i].hi[k] = bounds6[6 * i + 3 + k];
        }
    bvh::refit(t, b);
    std::memcpy(nodes, t.nodes.data(), t.nodes.size() * sizeof(bvh::Node));
    return bvh::sah_cost(t);
}

// make_icosphere / make_box (src/mesh.cpp:20-95). Call with null outputs to get the sizes.
void ref_make_icosphere(float radius, int subdivisions, float *vertices, uint32_t *indices, int *nv, int *ni) {
    const Mesh m = make_icosphere(radius, subdivisions);
    *nv = static_cast<int>(m.vertices.size());
    *ni = static_cast<int>(m.indices.size());
    if (vertices) std::memcpy(vertices, m.vertices.data(), m.vertices.size() * sizeof(Vec3));
    if (indices) std::memcpy(indices, m.indices.data(), m.indices.size() * sizeof(uint32_t));
}

void ref_make_box(float hx, float hy, float hz, float *vertices, uint32_t *indices, int *nv, int *ni) {
    const Mesh m = make_box(Vec3{hx, hy, hz});
    *nv = static_cast<int>(m.vertices.size());
    *ni = static_cast<int>(m.indices.size());
    if (vertices) std::memcpy(vertices, m.vertices.data(), m.vertices.size() * sizeof(Vec3));
    if (indices) std::memcpy(indices, m.indices.data(), m.indices.size() * sizeof(uint32_t));
}

// make_transform (src/mesh.cpp:20-34).
void ref_make_transform(const float *position, const float *quat_xyzw, float scale, float *m12) {
    const Transform t = make_transform(Vec3{position[0], position[1], position[2]},
        Quat{quat_xyzw[0], quat_xyzw[1], quat_xyzw[2], quat_xyzw[3]}, scale);
    std::memcpy(m12, t.m, 12 * sizeof(float));
}

// camera_from_matrix + camera_ray (src/render_core.h:264-273, src/device_scene.h:114-119): the primary
// ray through every pixel centre of every view. origins/dirs: [views][H][W][3].
void ref_camera_rays(const float *cams12, int views, float fov, int width, int height, float *origins, float *dirs,
    float *forward) {
    using namespace neraulgraphics::host;
    for (int v = 0; v < views; ++v) {
        const neraulgraphics::host::Camera c = camera_from_matrix(cams12 + 12 * v, fov, width, height);
        forward[3 * v] = c.forward.x;
        forward[3 * v + 1] = c.forward.y;
        forward[3 * v + 2] = c.forward.z;
        for (int y = 0; y < height; ++y)
            for (int x = 0; x < width; ++x) {
                const float3 d = camera_ray(c, static_cast<float>(x) + 0.5f, static_cast<float>(y) + 0.5f, width, height);
                const size_t i = (static_cast<size_t>(v) * height + y) * width + x;
                origins[3 * i] = c.position.x;
                origins[3 * i + 1] = c.position.y;
                origins[3 * i + 2] = c.position.z;
                dirs[3 * i] = d.x;
                dirs[3 * i + 1] = d.y;
                dirs[3 * i + 2] = d.z;
            }
    }
}

// mesh_intersect (src/mesh_dev.h:177-205) for a batch of rays, each in world ray_world[i], over a given
// top-level forest (TlasNode layout) and bottom level (Node layout, triangles v0,e1,e2 as float4 x 3).
// inv: [N][12] world-to-mesh rows; blas_root: [N]. tmax_in: the distance to beat per ray (1e30 for a
// plain query). Outputs: t (unchanged tmax when nothing is hit), instance, triangle (-1 when no hit), u, v.
void ref_mesh_intersect(const bvh::TlasNode *tlas, const int32_t *tlas_root, const float *inv, const int32_t *blas_root,
    int instance_count, const bvh::Node *blas, const float *triangles, int count, const int32_t *ray_world,
    const float *origins, const float *dirs, const float *tmax_in, float *t_out, int32_t *inst_out, int32_t *tri_out,
    float *u_out, float *v_out) {
    using namespace neraulgraphics::host;
    std::vector<neraulgraphics::host::Instance> records(static_cast<size_t>(instance_count));
    for (int i = 0; i < instance_count; ++i) {
        std::memset(static_cast<void *>(&records[i]), 0, sizeof(neraulgraphics::host::Instance));
        for (int row = 0; row < 3; ++row)
            records[i].inv[row] = make_float4(inv[12 * i + 4 * row], inv[12 * i + 4 * row + 1], inv[12 * i + 4 * row + 2],
                inv[12 * i + 4 * row + 3]);
        records[i].blas_root = blas_root[i];
        records[i].texture = records[i].normal_texture = records[i].orm_texture = -1;
    }
    neraulgraphics::host::MeshScene m{};
    m.tlas = tlas;
    m.tlas_root = tlas_root;
    m.instances = records.data();
    m.blas = blas;
    m.triangles = reinterpret_cast<const float4 *>(triangles);
    m.attributes = nullptr;
    m.textures = nullptr;
    m.instance_count = instance_count;
    m.instances_per_world = 0;
    for (int i = 0; i < count; ++i) {
        const float3 o = make_float3(origins[3 * i], origins[3 * i + 1], origins[3 * i + 2]);
        const float3 d = make_float3(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]);
        float tmax = tmax_in[i];
        int hi = -1, ht = -1;
        float hu = 0.0f, hv = 0.0f;
        const bool found = tlas_root[ray_world[i]] >= 0 && mesh_intersect(m, ray_world[i], o, d, tmax, hi, ht, hu, hv);
        t_out[i] = tmax;
        inst_out[i] = found ? hi : -1;
        tri_out[i] = found ? ht : -1;
        u_out[i] = hu;
        v_out[i] = hv;
    }
}

}  // extern "C"
