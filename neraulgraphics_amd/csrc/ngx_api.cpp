// The C ABI (include/ngx_backend.h): a context that owns every device buffer of a batch, checks input
// the way the reference's front-end does (src/batch.cpp:47-192), and queues the kernels of ngx_tlas.hip
// and ngx_trace.hip on the caller's stream.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>
#include "../../include/ngx_backend.h"
#include "ngx_host.h"
#include "ngx_kernels.h"

static_assert(ngx::BACKWARD_MAX_BODIES == NGX_BACKWARD_MAX_BODIES, "the header states the backward's body limit");
static_assert(ngx::MAX_TRIANGLES == NGX_MAX_TRIANGLES, "the header states the triangle limit");
static_assert(ngx::NOALLOC_SCRATCH_FLOATS == (size_t)NGX_NOALLOC_SCRATCH_FLOATS, "the header states the scratch array's size");

namespace {

thread_local std::string g_last_error;

struct Failure {
    ngx_status status;
    std::string message;
};

[[noreturn]] void fail(const ngx_status status, const std::string &message) { throw Failure{status, message}; }
void require(const bool ok, const std::string &message) {
    if (!ok) fail(NGX_INVALID_ARGUMENT, message);
}
void require_state(const bool ok, const std::string &message) {
    if (!ok) fail(NGX_BAD_STATE, message);
}
void hip_check(const hipError_t err, const char *what) {
    if (err == hipSuccess) return;
    (void)hipGetLastError();
    fail(err == hipErrorOutOfMemory ? NGX_OUT_OF_MEMORY : NGX_DEVICE_ERROR,
         std::string(what) + " failed: " + hipGetErrorString(err) + " (" + hipGetErrorName(err) + ")");
}
#define HIP_CHECK(call) hip_check((call), #call)

template <class Fn>
ngx_status guarded(Fn &&fn) {
    try {
        fn();
        g_last_error.clear();
        return NGX_OK;
    } catch (const Failure &f) {
        g_last_error = f.message;
        return f.status;
    } catch (const std::bad_alloc &) {
        g_last_error = "host out of memory";
        return NGX_OUT_OF_MEMORY;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return NGX_DEVICE_ERROR;
    }
}

bool finite_all(const float *v, const size_t n) {
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(v[i])) return false;
    return true;
}

bool quaternion_usable(const float *q) {
    return std::isfinite(q[0]) && std::isfinite(q[1]) && std::isfinite(q[2]) && std::isfinite(q[3]) &&
           q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3] > 0.0f;
}

// Allocates count elements for ptr unless it holds some already; false (and nothing held, no error left pending) when
// the device has no room.
template <class T>
bool try_device_alloc(T *&ptr, const size_t count) {
    if (ptr) return true;
    if (hipMalloc(reinterpret_cast<void **>(&ptr), count * sizeof(T)) == hipSuccess) return true;
    ptr = nullptr;
    (void)hipGetLastError();
    return false;
}

template <class T>
void device_alloc(T *&ptr, const size_t count) {
    if (ptr) {
        (void)hipFree(ptr);
        ptr = nullptr;
    }
    if (count == 0) return;
    HIP_CHECK(hipMalloc(reinterpret_cast<void **>(&ptr), count * sizeof(T)));
}

template <class T>
void upload(T *&ptr, const T *src, const size_t count) {
    device_alloc(ptr, count);
    if (count) HIP_CHECK(hipMemcpy(ptr, src, count * sizeof(T), hipMemcpyHostToDevice));
}

template <class T>
void release(T *&ptr) {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
}

}  // namespace

struct ngx_ctx {
    int width = 0, height = 0, worlds = 0, cpw = 1, views = 0;
    int per_world = 0, total = 0, max_count = 0, device = 0;
    int small_max = 0;                // the largest world built by one workgroup
    std::vector<int> big_worlds;      // worlds built by the device-wide path
    ngx::BigBuildScratch big{};
    std::vector<int> counts, world_off;
    int sphere_count = 0;
    float floor[4] = {0, 0, 0, 0};

    // scene, fixed at creation
    int *d_world_off = nullptr;
    uint8_t *d_root_at_end = nullptr;
    int *d_tmpl_mesh = nullptr;
    float *d_mesh_bounds = nullptr;
    ngx::WideNode *d_blas_wide = nullptr;
    int *d_blas_parent = nullptr;
    int *d_blas_wide_node = nullptr;
    // deforming meshes (ngx_set_mesh_vertices)
    float *d_vertices = nullptr;        // every mesh's vertices, 3 floats each
    int *d_tri_verts = nullptr, *d_node_parent = nullptr, *d_node_wide = nullptr, *d_node_arrivals = nullptr;
    std::vector<int> mesh_vertex_base, mesh_vertex_count, mesh_node_base, mesh_node_count, mesh_tri_base, mesh_tri_count;
    std::vector<uint8_t> mesh_has_normals;
    bool top_level_stale = false;       // a mesh moved since the last pose update
    ngx::SkipNode *d_skip = nullptr;
    int *d_mesh_skip_root = nullptr, *d_mesh_tris = nullptr;
    int max_mesh_tris = 0, min_mesh_tris = 0;
    ngx::MeshRoot *d_mesh_roots = nullptr;
    float *d_tris = nullptr;            // 9 floats per triangle
    float4 *d_attrs = nullptr;
    float4 *d_spheres = nullptr;

    // the instances' state, rewritten by every pose update
    float *d_matrices = nullptr;        // [N][12] as placed
    float *d_matrix_stage = nullptr;    // [W][P][12] host matrices land here
    ngx::InstanceRec *d_inst = nullptr;
    ngx::TlasNode *d_nodes = nullptr;
    ngx::WideNode *d_wide = nullptr;
    int *d_wide_parent = nullptr;
    int *d_roots = nullptr;

    // bodies
    bool bound = false;
    int bodies = 0;
    int *d_body_of_instance = nullptr;
    float *d_local_pose = nullptr, *d_scale = nullptr, *d_local_rot_scale = nullptr;
    float *d_body_stage = nullptr;
    const float *last_body_poses = nullptr;

    // simulator geoms: each template instance drawn at a geom's world frame, scaled
    bool geoms_bound = false;
    int geoms = 0;
    int *d_geom_of_instance = nullptr;
    float *d_geom_scale = nullptr, *d_geom_stage = nullptr;  // stage: [W][G][3] positions, then [W][G][9] rotations

    // shading of the RGB picture
    bool shading_set = false;
    ngx::ShadeScene shade{};
    float4 *d_inst_mat = nullptr, *d_sphere_mat = nullptr, *d_lights = nullptr;
    uint8_t *d_rgb = nullptr;
    bool shade_no_bounce = false;       // every surface's reflectivity below the path's cut-off: paths end at their first hit
    int last_rgb_passes = 0;            // how the last RGB picture was drawn: 1 or 2 passes (0 before any)
    int4 *d_rgb_hits = nullptr;         // the two-pass picture's hits, pixels
    float2 *d_rgb_uv = nullptr;         // and their barycentrics (meshes with vertex normals)

    // outputs and staging, grown as needed
    float *d_depth = nullptr, *d_normal = nullptr;
    int32_t *d_id = nullptr;
    float *d_cam_stage = nullptr;
    float *d_ray_o = nullptr, *d_ray_d = nullptr;
    long long ray_in_cap = 0;
    float *d_hit_dist = nullptr, *d_hit_normal = nullptr;
    int32_t *d_hit_id = nullptr;
    uint32_t *d_ray_keys = nullptr;     // the direction order's scratch: 2 x rays (cell and index in one word)
    int *d_ray_vals = nullptr;          // the direction order: rays
    long long ray_cap = 0;
    float *d_partial = nullptr;
    size_t partial_cap = 0;
    float *d_noalloc_scratch = nullptr;  // this device's fixed scratch array (not owned: it lives as long as the process)

    // build intermediates
    bool keep_build = false;
    uint64_t *d_dbg_keys = nullptr, *d_dbg_sorted_keys = nullptr;
    int *d_dbg_sorted_ids = nullptr;
    float *d_dbg_aabb = nullptr, *d_dbg_world_bounds = nullptr;

    // timing and tuning
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool timed = false;
    hipStream_t last_stream = nullptr;
    int tile_w = 8, xcd_remap = 1, packet = 1, pair_limit = 64, persistent = 1, lane_min_tris = 300, sort_rays = 1;
    int share_cap = ngx::SHARE_STACK;   // the two-child shared walk's LDS stack entries (ngx_debug_set_share_capacity)
    // kernel variants by workgroups per compute unit (0: each walk's default)
    int render_blocks = 0, rgb_blocks = 0, ray_blocks = 0, backward_blocks = 0;
    int rgb_two_pass = 1;
    int compute_units = 0;              // sizes the persistent launch
    int *d_tile_counters = nullptr;
    int *d_overflow = nullptr;          // set by a walk whose stack would pass its capacity; read by ngx_synchronize

    ~ngx_ctx() {
        int *ints[] = {d_world_off, d_tmpl_mesh, d_roots, d_body_of_instance, d_dbg_sorted_ids, d_blas_parent, d_blas_wide_node, d_wide_parent,
                       d_tri_verts, d_node_parent, d_node_wide, d_node_arrivals,
                       d_geom_of_instance, d_tile_counters, d_mesh_skip_root, d_mesh_tris, d_overflow};
        for (int *p : ints) if (p) (void)hipFree(p);
        float *floats[] = {d_vertices, d_mesh_bounds, d_matrices, d_matrix_stage, d_local_pose, d_scale, d_local_rot_scale, d_body_stage,
                           d_geom_scale, d_geom_stage, d_depth, d_normal, d_cam_stage, d_ray_o, d_ray_d, d_hit_dist, d_hit_normal, d_partial, d_dbg_aabb,
                           d_dbg_world_bounds};
        for (float *p : floats) if (p) (void)hipFree(p);
        void *others[] = {d_root_at_end, d_blas_wide, d_mesh_roots, d_tris, d_attrs, d_spheres, d_inst, d_nodes, d_wide,
                          d_id, d_hit_id, d_dbg_keys, d_dbg_sorted_keys, d_skip, d_ray_keys, d_ray_vals, d_inst_mat, d_sphere_mat,
                          d_lights, d_rgb, d_rgb_hits, d_rgb_uv, big.key[0], big.key[1], big.val[0], big.val[1],
                          big.box, big.node_box, big.left, big.right, big.parent, big.arrivals, big.bounds, big.tile_counts};
        for (void *p : others) if (p) (void)hipFree(p);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
    }

    ngx::TraceScene trace_scene() const {
        ngx::TraceScene s{};
        s.tlas_wide = d_wide;
        s.tlas_parent = d_wide_parent;
        s.blas_parent = d_blas_parent;
        s.blas_wide_node = d_blas_wide_node;
        s.roots = d_roots;
        s.total = total;
        s.inst = d_inst;
        s.mesh_roots = d_mesh_roots;
        s.blas_wide = d_blas_wide;
        s.tris = d_tris;
        s.attrs = d_attrs;
        s.sphere_count = sphere_count;
        s.spheres = d_spheres;
        s.floor_y = floor[0];
        s.floor_hx = floor[1];
        s.floor_zmin = floor[2];
        s.floor_zmax = floor[3];
        s.pair_limit = pair_limit;
        s.skip = d_skip ? d_skip + 1 : nullptr;
        s.mesh_skip_root = d_mesh_skip_root;
        s.mesh_tris = d_mesh_tris;
        s.lane_min_tris = lane_min_tris;
        s.overflow = d_overflow;
        s.share_cap = share_cap;
        return s;
    }

    ngx::TlasBuildArgs build_args() const {
        ngx::TlasBuildArgs a{};
        a.worlds = worlds;
        a.total = total;
        a.world_off = d_world_off;
        a.root_at_end = d_root_at_end;
        a.matrices_stride = per_world;
        a.bodies_per_world = bodies;
        a.body_of_instance = d_body_of_instance;
        a.local_pose = d_local_pose;
        a.scale = d_scale;
        a.geoms_per_world = geoms;
        a.geom_of_instance = d_geom_of_instance;
        a.geom_scale = d_geom_scale;
        a.matrices_out = d_matrices;
        a.tmpl_mesh = d_tmpl_mesh;
        a.mesh_bounds = d_mesh_bounds;
        a.mesh_tris = d_mesh_tris;
        a.mesh_skip_root = d_mesh_skip_root;
        a.inst = d_inst;
        a.nodes = d_nodes;
        a.wide = d_wide;
        a.wide_parent = d_wide_parent;
        a.roots = d_roots;
        if (keep_build) {
            a.dbg_keys = d_dbg_keys;
            a.dbg_sorted_keys = d_dbg_sorted_keys;
            a.dbg_sorted_ids = d_dbg_sorted_ids;
            a.dbg_aabb = d_dbg_aabb;
            a.dbg_world_bounds = d_dbg_world_bounds;
        }
        return a;
    }

    void set_device() const { HIP_CHECK(hipSetDevice(device)); }

    ngx::RenderArgs render_args(const float *cams, const float fov) const {
        ngx::RenderArgs a{};
        a.s = trace_scene();
        a.cams = cams;
        a.fov_scale = std::tan(fov * 0.5f);
        a.aspect = (float)width / (float)height;
        a.width = width;
        a.height = height;
        a.cameras_per_world = cpw;
        a.views = views;
        a.tile_w = tile_w;
        a.tile_h = 64 / tile_w;
        a.tiles_x = (width + a.tile_w - 1) / a.tile_w;
        a.tiles_per_view = a.tiles_x * ((height + a.tile_h - 1) / a.tile_h);
        a.total_tiles = (long long)a.tiles_per_view * views;
        a.xcd_remap = xcd_remap;
        a.walk = walk();
        a.tile_counters = persistent ? d_tile_counters : nullptr;
        return a;
    }

    // The kernels' walk: per-lane, packet, or packet with the bottom level of meshes of lane_min_tris or more
    // triangles walked lane by lane.
    int walk() const { return !packet ? 0 : (max_mesh_tris >= lane_min_tris ? 2 : 1); }
    // Ray queries and their backward: when every mesh is walked lane by lane, the walk built without the packet
    // bottom level (3), which holds fewer registers; the same results.
    int walk_rays() const { return packet && total > 0 && min_mesh_tris >= lane_min_tris ? 3 : walk(); }

    bool capturing(hipStream_t s) const {
        hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(s, &st) != hipSuccess) return false;
        return st != hipStreamCaptureStatusNone;
    }

    void build(const ngx::TlasBuildArgs &a, hipStream_t stream) const {
        HIP_CHECK(ngx::launch_tlas_build(a, small_max, stream));
        for (const int w : big_worlds) HIP_CHECK(ngx::launch_tlas_build_big(a, big, w, world_off[w], counts[w], stream));
    }

    void build_from_matrices(const float *matrices_wp12, hipStream_t stream) {
        if (total == 0) return;
        ngx::TlasBuildArgs a = build_args();
        a.source = ngx::SOURCE_MATRICES;
        a.matrices_in = matrices_wp12;
        build(a, stream);
        last_body_poses = nullptr;
        top_level_stale = false;
    }

    void build_from_bodies(const float *poses, hipStream_t stream) {
        if (total == 0) return;
        ngx::TlasBuildArgs a = build_args();
        a.source = ngx::SOURCE_BODIES;
        a.body_poses = poses;
        build(a, stream);
        last_body_poses = poses;
        top_level_stale = false;
    }

    void build_from_geoms(const float *pos3, const float *mat9, hipStream_t stream) {
        if (total == 0) return;
        ngx::TlasBuildArgs a = build_args();
        a.source = ngx::SOURCE_GEOMS;
        a.geom_pos = pos3;
        a.geom_mat = mat9;
        build(a, stream);
        last_body_poses = nullptr;
        top_level_stale = false;
    }
};

extern "C" {

const char *ngx_last_error(void) { return g_last_error.c_str(); }

ngx_status ngx_create(const ngx_scene_desc *scene, ngx_ctx **out) {
    return guarded([&] {
        require(scene != nullptr && out != nullptr, "scene and out must not be null");
        *out = nullptr;
        const ngx_scene_desc &s = *scene;
        require(s.width > 0 && s.height > 0,
                "frame size must be positive, got " + std::to_string(s.width) + "x" + std::to_string(s.height));
        require((long long)s.width * s.height <= INT32_MAX, "frame is too large");
        require(s.worlds >= 1, "world_count must be 1 or more, got " + std::to_string(s.worlds));
        require(s.cameras_per_world >= 1, "cameras_per_world must be 1 or more, got " + std::to_string(s.cameras_per_world));
        const long long views = (long long)s.worlds * s.cameras_per_world;
        require(views * s.height <= INT32_MAX, std::to_string(views) + " views of height " + std::to_string(s.height) + " are too many");
        require(s.sphere_count >= 0 && s.sphere_count <= NGX_MAX_SPHERES,
                "scene has " + std::to_string(s.sphere_count) + " spheres, at most " + std::to_string(NGX_MAX_SPHERES) + " fit");
        for (int i = 0; i < s.sphere_count; ++i) {
            require(finite_all(s.spheres + 4 * i, 4), "sphere " + std::to_string(i) + " has a NaN or infinite value");
            require(s.spheres[4 * i + 3] > 0.0f, "sphere " + std::to_string(i) + " radius must be positive");
        }
        const float fl[4] = {s.floor_y, s.floor_half_extent_x, s.floor_z_min, s.floor_z_max};
        require(finite_all(fl, 4), "floor has a NaN or infinite value");
        require(s.mesh_count >= 0 && s.instances_per_world >= 0, "mesh and instance counts must not be negative");
        long long all_triangles = 0;
        for (int m = 0; m < s.mesh_count; ++m) {
            const std::string name = "mesh " + std::to_string(m);
            const int nv = s.vertex_offsets[m + 1] - s.vertex_offsets[m];
            const int ni = s.index_offsets[m + 1] - s.index_offsets[m];
            require(nv >= 0 && ni > 0, name + " has no triangles");
            require(ni % 3 == 0, name + " has " + std::to_string(ni) + " indices, which is not a multiple of 3");
            require(ni / 3 <= 100000000, name + " has too many triangles");
            all_triangles += ni / 3;
            require(finite_all(s.vertices + 3 * (size_t)s.vertex_offsets[m], 3 * (size_t)nv), name + " has a NaN or infinite vertex");
            for (int i = 0; i < ni; ++i)
                require(s.indices[s.index_offsets[m] + i] < (uint32_t)nv,
                        name + " index " + std::to_string(i) + " is " + std::to_string(s.indices[s.index_offsets[m] + i]) +
                            ", but there are only " + std::to_string(nv) + " vertices");
            if (s.normals && s.mesh_has_normals && s.mesh_has_normals[m])
                require(finite_all(s.normals + 3 * (size_t)s.vertex_offsets[m], 3 * (size_t)nv), name + " has a NaN or infinite normal");
        }
        if (all_triangles > ngx::MAX_TRIANGLES)
            fail(NGX_UNSUPPORTED, "the meshes hold " + std::to_string(all_triangles) + " triangles; at most " +
                                      std::to_string(ngx::MAX_TRIANGLES) + " over all meshes are supported");
        for (int i = 0; i < s.instances_per_world; ++i) {
            const std::string name = "instance " + std::to_string(i);
            require(s.instance_mesh[i] >= 0 && s.instance_mesh[i] < s.mesh_count,
                    name + " uses mesh " + std::to_string(s.instance_mesh[i]) + ", but there are " + std::to_string(s.mesh_count) + " meshes");
            require(finite_all(s.instance_transforms + 12 * (size_t)i, 12), name + " transform has a NaN or infinite value");
            require(ngx::host::transform_ok(s.instance_transforms + 12 * (size_t)i),
                    name + " transform is singular (it flattens space, so it cannot be inverted)");
        }
        std::vector<int> counts(s.worlds, s.instances_per_world);
        if (s.instances_in_world)
            for (int w = 0; w < s.worlds; ++w) {
                require(s.instances_in_world[w] >= 0 && s.instances_in_world[w] <= s.instances_per_world,
                        "world " + std::to_string(w) + " holds " + std::to_string(s.instances_in_world[w]) +
                            " instances, but the template has " + std::to_string(s.instances_per_world));
                counts[w] = s.instances_in_world[w];
            }
        long long total = 0;
        int max_count = 0;
        for (const int c : counts) {
            total += c;
            max_count = std::max(max_count, c);
        }
        require(total <= INT32_MAX / 4, "too many instances over all worlds: " + std::to_string(total));

        auto ctx = std::make_unique<ngx_ctx>();
        ngx_ctx &c = *ctx;
        c.width = s.width;
        c.height = s.height;
        c.worlds = s.worlds;
        c.cpw = s.cameras_per_world;
        c.views = (int)views;
        c.per_world = s.instances_per_world;
        c.device = s.device_id;
        c.counts = counts;
        c.world_off.assign(s.worlds + 1, 0);
        for (int w = 0; w < s.worlds; ++w) c.world_off[w + 1] = c.world_off[w] + counts[w];
        c.total = (int)total;
        c.max_count = max_count;
        for (int w = 0; w < s.worlds; ++w) {
            if (counts[w] > ngx::WORKGROUP_BUILD_MAX) c.big_worlds.push_back(w);
            else c.small_max = std::max(c.small_max, counts[w]);
        }
        c.sphere_count = s.sphere_count;
        std::memcpy(c.floor, fl, sizeof(fl));
        c.set_device();
        {
            hipDeviceProp_t prop;
            HIP_CHECK(hipGetDeviceProperties(&prop, c.device));
            c.compute_units = std::max(1, prop.multiProcessorCount);
        }
        HIP_CHECK(ngx::noalloc_scratch(&c.d_noalloc_scratch));
        device_alloc(c.d_tile_counters, (size_t)ngx::TILE_COUNTER_INTS);
        device_alloc(c.d_overflow, 1);
        HIP_CHECK(hipMemset(c.d_overflow, 0, sizeof(int)));
        HIP_CHECK(hipEventCreate(&c.ev0));
        HIP_CHECK(hipEventCreate(&c.ev1));

        if (s.sphere_count > 0) upload(c.d_spheres, reinterpret_cast<const float4 *>(s.spheres), (size_t)s.sphere_count);
        if (c.total > 0) {
            ngx::host::MeshInput mi{s.mesh_count, s.vertices, s.vertex_offsets, s.indices, s.index_offsets, s.normals, s.mesh_has_normals};
            ngx::host::BuiltMeshes built;
            ngx::host::build_meshes(mi, built);
            {
                // the packet walk keeps one stack entry per level of a bottom-level tree: the build stops at 60
                int deepest = 0;
                for (size_t k = 0; k < built.node_parent.size(); ++k) {
                    int d = 0;
                    for (int p = built.node_parent[k]; p >= 0 && d <= ngx::WALK_STACK; p = built.node_parent[p]) ++d;
                    deepest = std::max(deepest, d);
                }
                if (deepest >= ngx::WALK_STACK)
                    fail(NGX_DEVICE_ERROR, "a bottom-level tree is " + std::to_string(deepest) + " levels deep; the walks hold " +
                                               std::to_string(ngx::WALK_STACK - 1));
            }
            upload(c.d_blas_wide, built.wide.data(), built.wide.size() ? built.wide.size() : 0);
            if (!c.d_blas_wide) device_alloc(c.d_blas_wide, 1);
            upload(c.d_blas_parent, built.wide_parent.data(), built.wide_parent.size());
            if (!c.d_blas_parent) device_alloc(c.d_blas_parent, 1);
            upload(c.d_blas_wide_node, built.wide_node.data(), built.wide_node.size());
            if (!c.d_blas_wide_node) device_alloc(c.d_blas_wide_node, 1);
            upload(c.d_mesh_roots, built.mesh_roots.data(), built.mesh_roots.size());
            {
                // one leading unused record: node i then sits at 64 (i + 1) bytes, so the child pairs (2k + 1, 2k + 2)
                // of a mesh whose nodes start at an even index share one 128-byte cache line
                std::vector<ngx::SkipNode> shifted(built.skip.size() + 1);
                std::copy(built.skip.begin(), built.skip.end(), shifted.begin() + 1);
                upload(c.d_skip, shifted.data(), shifted.size());
            }
            upload(c.d_mesh_skip_root, built.roots.data(), built.roots.size());
            upload(c.d_vertices, s.vertices, 3 * (size_t)s.vertex_offsets[s.mesh_count]);
            upload(c.d_tri_verts, built.tri_verts.data(), built.tri_verts.size());
            upload(c.d_node_parent, built.node_parent.data(), built.node_parent.size());
            upload(c.d_node_wide, built.node_wide.data(), built.node_wide.size());
            device_alloc(c.d_node_arrivals, built.nodes.size());
            for (int m = 0; m < s.mesh_count; ++m) {
                c.mesh_vertex_base.push_back(s.vertex_offsets[m]);
                c.mesh_vertex_count.push_back(s.vertex_offsets[m + 1] - s.vertex_offsets[m]);
                c.mesh_has_normals.push_back(s.normals && s.mesh_has_normals && s.mesh_has_normals[m] ? 1 : 0);
            }
            c.mesh_node_base = built.node_base;
            c.mesh_node_count = built.node_count;
            c.mesh_tri_base = built.tri_base;
            c.mesh_tri_count = built.tri_counts;
            upload(c.d_mesh_tris, built.tri_counts.data(), built.tri_counts.size());
            for (const int t : built.tri_counts) c.max_mesh_tris = std::max(c.max_mesh_tris, t);
            c.min_mesh_tris = *std::min_element(built.tri_counts.begin(), built.tri_counts.end());
            {
                // v0, e1, e2 without the padding float of each: 36 bytes a triangle on the device
                std::vector<float> packed;
                packed.reserve(built.triangles.size() / 12 * 9);
                for (size_t t = 0; t + 12 <= built.triangles.size(); t += 12)
                    for (int k = 0; k < 3; ++k) packed.insert(packed.end(), &built.triangles[t + 4 * k], &built.triangles[t + 4 * k + 3]);
                upload(c.d_tris, packed.data(), packed.size());
            }
            if (!built.attributes.empty())
                upload(c.d_attrs, reinterpret_cast<const float4 *>(built.attributes.data()), built.attributes.size() / 4);
            std::vector<float> mesh_bounds;
            for (const auto &b : built.bounds) mesh_bounds.insert(mesh_bounds.end(), {b.lo[0], b.lo[1], b.lo[2], b.hi[0], b.hi[1], b.hi[2]});
            upload(c.d_mesh_bounds, mesh_bounds.data(), mesh_bounds.size());
            upload(c.d_tmpl_mesh, s.instance_mesh, (size_t)s.instances_per_world);
            upload(c.d_world_off, c.world_off.data(), c.world_off.size());
            // Which end of its run carries a world's root in the reference's global tree: the end next to the
            // neighbouring world that shares fewer leading key bits with it (see ngx_tlas.hip).
            std::vector<uint8_t> root_at_end(s.worlds, 0);
            int prev = -1;
            std::vector<int> next_nonempty(s.worlds, -1);
            for (int w = s.worlds - 1, nx = -1; w >= 0; --w) {
                next_nonempty[w] = nx;
                if (counts[w] > 0) nx = w;
            }
            for (int w = 0; w < s.worlds; ++w) {
                if (counts[w] == 0) continue;
                const int p_prev = prev < 0 ? -1 : __builtin_clz((unsigned)(w ^ prev));
                const int p_next = next_nonempty[w] < 0 ? -1 : __builtin_clz((unsigned)(w ^ next_nonempty[w]));
                root_at_end[w] = p_next > p_prev ? 1 : 0;
                prev = w;
            }
            upload(c.d_root_at_end, root_at_end.data(), root_at_end.size());
            device_alloc(c.d_matrices, 12 * (size_t)c.total);
            device_alloc(c.d_matrix_stage, 12 * (size_t)s.worlds * s.instances_per_world);
            device_alloc(c.d_inst, (size_t)c.total);
            device_alloc(c.d_nodes, 2 * (size_t)c.total - 1);
            device_alloc(c.d_wide, std::max<size_t>(1, (size_t)c.total - 1));
            device_alloc(c.d_wide_parent, std::max<size_t>(1, (size_t)c.total - 1));
            if (!c.big_worlds.empty()) {
                ngx::BigBuildScratch &b = c.big;
                b.capacity = 0;
                for (const int w : c.big_worlds) b.capacity = std::max(b.capacity, counts[w]);
                const size_t cap = (size_t)b.capacity, tiles = (cap + 1023) / 1024;
                for (int k = 0; k < 2; ++k) {
                    device_alloc(b.key[k], cap);
                    device_alloc(b.val[k], cap);
                }
                device_alloc(b.box, 6 * cap);
                device_alloc(b.node_box, 6 * cap);
                device_alloc(b.left, cap);
                device_alloc(b.right, cap);
                device_alloc(b.parent, 2 * cap);
                device_alloc(b.arrivals, cap);
                device_alloc(b.bounds, 6);
                device_alloc(b.tile_counts, 16 * tiles);
            }
            // nodes that join worlds are never written by a build: give them a fixed value once
            HIP_CHECK(hipMemset(c.d_nodes, 0, sizeof(ngx::TlasNode) * (2 * (size_t)c.total - 1)));
            HIP_CHECK(hipMemset(c.d_wide, 0, sizeof(ngx::WideNode) * std::max<size_t>(1, (size_t)c.total - 1)));
        }
        device_alloc(c.d_roots, (size_t)s.worlds);
        HIP_CHECK(hipMemset(c.d_roots, 0xFF, sizeof(int) * s.worlds));
        if (c.total > 0) {
            // every world starts where the template puts its instances
            std::vector<float> start(12 * (size_t)s.worlds * s.instances_per_world);
            for (int w = 0; w < s.worlds; ++w)
                std::memcpy(&start[12 * (size_t)w * s.instances_per_world], s.instance_transforms, sizeof(float) * 12 * s.instances_per_world);
            HIP_CHECK(hipMemcpy(c.d_matrix_stage, start.data(), start.size() * sizeof(float), hipMemcpyHostToDevice));
            c.build_from_matrices(c.d_matrix_stage, nullptr);
            HIP_CHECK(hipDeviceSynchronize());
        }
        *out = ctx.release();
    });
}

void ngx_destroy(ngx_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    delete ctx;
}

ngx_status ngx_set_instance_matrices(ngx_ctx *ctx, const float *matrices, ngx_memory memory, void *stream) {
    return guarded([&] {
        require(ctx != nullptr, "context is null");
        if (ctx->per_world == 0) return;
        require(matrices != nullptr, "matrices must not be null");
        ctx->set_device();
        const hipStream_t s = static_cast<hipStream_t>(stream);
        const float *src = matrices;
        if (memory == NGX_MEM_HOST) {
            for (int w = 0; w < ctx->worlds; ++w)
                for (int i = 0; i < ctx->counts[w]; ++i)
                    if (!ngx::host::transform_ok(matrices + 12 * ((size_t)w * ctx->per_world + i))) {
                        const float *m = matrices + 12 * ((size_t)w * ctx->per_world + i);
                        const std::string name = "world " + std::to_string(w) + " instance " + std::to_string(i) + " transform";
                        require(finite_all(m, 12), name + " has a NaN or infinite value");
                        require(false, name + " is singular (it flattens space, so it cannot be inverted)");
                    }
            HIP_CHECK(hipMemcpyAsync(ctx->d_matrix_stage, matrices, sizeof(float) * 12 * (size_t)ctx->worlds * ctx->per_world,
                                     hipMemcpyHostToDevice, s));
            src = ctx->d_matrix_stage;
        }
        ctx->last_stream = s;
        ctx->build_from_matrices(src, s);
    });
}

ngx_status ngx_bind_bodies(ngx_ctx *ctx, int32_t bodies_per_world, const int32_t *body_of_instance, const float *local_pose7,
                           const float *scale3) {
    return guarded([&] {
        require(ctx != nullptr, "context is null");
        const int n = ctx->per_world;
        require(bodies_per_world >= 0, "bodies_per_world must be 0 or more");
        require(n == 0 || (body_of_instance && local_pose7 && scale3),
                "a body binding needs one body index, 7 pose floats and 3 scale floats for each of the " + std::to_string(n) +
                    " instances of a world");
        for (int i = 0; i < n; ++i) {
            const int b = body_of_instance[i];
            require(b >= -1 && b < bodies_per_world, "instance " + std::to_string(i) + " is on body " + std::to_string(b) +
                                                         ", but there are " + std::to_string(bodies_per_world) + " bodies per world");
            require(finite_all(local_pose7 + 7 * i, 7) && quaternion_usable(local_pose7 + 7 * i + 3),
                    "instance " + std::to_string(i) + " has a pose with a NaN or infinite value, or a zero quaternion");
            require(finite_all(scale3 + 3 * i, 3), "instance " + std::to_string(i) + " has a scale with a NaN or infinite value");
        }
        ctx->set_device();
        HIP_CHECK(hipDeviceSynchronize());
        upload(ctx->d_body_of_instance, body_of_instance, (size_t)n);
        upload(ctx->d_local_pose, local_pose7, 7 * (size_t)n);
        upload(ctx->d_scale, scale3, 3 * (size_t)n);
        // Each instance in its body's frame, Rl diag(s) | pl, for the backward pass.
        std::vector<float> lrs(12 * (size_t)n);
        for (int i = 0; i < n; ++i) {
            const float *q = local_pose7 + 7 * i + 3;
            double x = q[0], y = q[1], z = q[2], w = q[3];
            const double nn = std::sqrt(x * x + y * y + z * z + w * w);
            x /= nn; y /= nn; z /= nn; w /= nn;
            const double r[9] = {1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
                                 2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
                                 2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)};
            for (int row = 0; row < 3; ++row) {
                for (int col = 0; col < 3; ++col) lrs[12 * i + 4 * row + col] = (float)(r[3 * row + col] * scale3[3 * i + col]);
                lrs[12 * i + 4 * row + 3] = local_pose7[7 * i + row];
            }
        }
        upload(ctx->d_local_rot_scale, lrs.data(), lrs.size());
        device_alloc(ctx->d_body_stage, 7 * (size_t)ctx->worlds * bodies_per_world);
        ctx->bodies = bodies_per_world;
        ctx->bound = true;
    });
}

ngx_status ngx_set_body_poses(ngx_ctx *ctx, const float *poses7, ngx_memory memory, void *stream) {
    return guarded([&] {
        require(ctx != nullptr, "context is null");
        require(ctx->bound, "set_body_poses needs bind_bodies to be called first");
        const size_t bodies = (size_t)ctx->worlds * ctx->bodies;
        require(bodies == 0 || poses7 != nullptr, "body_poses must not be null");
        if (memory == NGX_MEM_HOST)
            for (size_t b = 0; b < bodies; ++b)
                require(finite_all(poses7 + 7 * b, 7) && quaternion_usable(poses7 + 7 * b + 3),
                        "body " + std::to_string(b) + " has a NaN or infinite value, or a zero quaternion");
        if (ctx->per_world == 0) return;
        ctx->set_device();
        const hipStream_t s = static_cast<hipStream_t>(stream);
        const float *src = poses7;
        if (memory == NGX_MEM_HOST && bodies > 0) {
            HIP_CHECK(hipMemcpyAsync(ctx->d_body_stage, poses7, 7 * sizeof(float) * bodies, hipMemcpyHostToDevice, s));
            src = ctx->d_body_stage;
        }
        ctx->last_stream = s;
        ctx->build_from_bodies(src, s);
    });
}

ngx_status ngx_bind_geoms(ngx_ctx *ctx, int32_t geoms_per_world, const int32_t *geom_of_instance, const float *scale3) {
    return guarded([&] {
        require(ctx != nullptr, "context is null");
        const int n = ctx->per_world;
        require(geoms_per_world >= 0, "geoms_per_world must be 0 or more");
        require(n == 0 || (geom_of_instance && scale3),
                "a geom binding needs one geom index and 3 scale floats for each of the " + std::to_string(n) +
                    " instances of a world");
        for (int i = 0; i < n; ++i) {
            const int g = geom_of_instance[i];
            require(g >= 0 && g < geoms_per_world, "instance " + std::to_string(i) + " is drawn at geom " + std::to_string(g) +
                                                       ", but there are " + std::to_string(geoms_per_world) + " geoms per world");
            require(finite_all(scale3 + 3 * i, 3), "instance " + std::to_string(i) + " has a scale with a NaN or infinite value");
        }
        ctx->set_device();
        HIP_CHECK(hipDeviceSynchronize());
        upload(ctx->d_geom_of_instance, geom_of_instance, (size_t)n);
        upload(ctx->d_geom_scale, scale3, 3 * (size_t)n);
        device_alloc(ctx->d_geom_stage, 12 * (size_t)ctx->worlds * geoms_per_world);
        ctx->geoms = geoms_per_world;
        ctx->geoms_bound = true;
    });
}

ngx_status ngx_set_geom_frames(ngx_ctx *ctx, const float *xpos3, const float *xmat9, ngx_memory memory, void *stream) {
    return guarded([&] {
        require(ctx != nullptr, "context is null");
        require(ctx->geoms_bound, "set_geom_frames needs bind_geoms to be called first");
        const size_t frames = (size_t)ctx->worlds * ctx->geoms;
        require(frames == 0 || (xpos3 != nullptr && xmat9 != nullptr), "geom positions and rotations must not be null");
        if (memory == NGX_MEM_HOST)
            for (size_t g = 0; g < frames; ++g)
                require(finite_all(xpos3 + 3 * g, 3) && finite_all(xmat9 + 9 * g, 9),
                        "geom frame " + std::to_string(g) + " has a NaN or infinite value");
        if (ctx->per_world == 0) return;
        ctx->set_device();
        const hipStream_t s = static_cast<hipStream_t>(stream);
        const float *pos = xpos3, *mat = xmat9;
        if (memory == NGX_MEM_HOST && frames > 0) {
            HIP_CHECK(hipMemcpyAsync(ctx->d_geom_stage, xpos3, 3 * sizeof(float) * frames, hipMemcpyHostToDevice, s));
            HIP_CHECK(hipMemcpyAsync(ctx->d_geom_stage + 3 * frames, xmat9, 9 * sizeof(float) * frames, hipMemcpyHostToDevice, s));
            pos = ctx->d_geom_stage;
            mat = ctx->d_geom_stage + 3 * frames;
        }
        ctx->last_stream = s;
        ctx->build_from_geoms(pos, mat, s);
    });
}

namespace {
// validate.cpp:23-31 for the classic fields
void check_material(const ngx_material &m, const std::string &name) {
    require(finite_all(m.diffuse_color, 3) && std::isfinite(m.diffuse_albedo) && std::isfinite(m.specular_albedo) &&
                std::isfinite(m.reflectivity) && std::isfinite(m.specular_exponent),
            name + " has a NaN or infinite value");
}
void pack_material(const ngx_material &m, float4 out[2]) {
    out[0] = make_float4(m.diffuse_color[0], m.diffuse_color[1], m.diffuse_color[2], m.diffuse_albedo);
    out[1] = make_float4(m.specular_albedo, m.reflectivity, m.specular_exponent, 0.0f);
}
}  // namespace

ngx_status ngx_set_mesh_vertices(ngx_ctx *ctx, const int32_t mesh, const float *vertices, const ngx_memory memory, void *stream) {
    return guarded([&] {
        require(ctx != nullptr, "context is null");
        require(mesh >= 0 && mesh < (int)ctx->mesh_vertex_count.size(),
                "mesh " + std::to_string(mesh) + " is not one of the scene's " + std::to_string(ctx->mesh_vertex_count.size()) +
                    " meshes (a scene with no instances keeps none)");
        if (ctx->mesh_has_normals[mesh])
            fail(NGX_UNSUPPORTED, "mesh " + std::to_string(mesh) + " has vertex normals, which a refit does not move");
        const size_t n = 3 * (size_t)ctx->mesh_vertex_count[mesh];
        require(n == 0 || vertices != nullptr, "vertices must not be null");
        if (memory == NGX_MEM_HOST) require(finite_all(vertices, n), "mesh " + std::to_string(mesh) + " has a NaN or infinite vertex");
        if (ctx->mesh_tri_count[mesh] == 0) return;
        ctx->set_device();
        const hipStream_t s = static_cast<hipStream_t>(stream);
        float *dst = ctx->d_vertices + 3 * (size_t)ctx->mesh_vertex_base[mesh];
        HIP_CHECK(hipMemcpyAsync(dst, vertices, n * sizeof(float),
                                 memory == NGX_MEM_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, s));
        ngx::MeshRefitArgs a{};
        a.vertices = dst;
        a.vertex_base = ctx->mesh_vertex_base[mesh];
        a.tri_verts = ctx->d_tri_verts;
        a.tri_base = ctx->mesh_tri_base[mesh];
        a.tri_count = ctx->mesh_tri_count[mesh];
        a.tris = ctx->d_tris;
        a.skip = ctx->d_skip + 1;
        a.node_parent = ctx->d_node_parent;
        a.node_wide = ctx->d_node_wide;
        a.wide = ctx->d_blas_wide;
        a.node_base = ctx->mesh_node_base[mesh];
        a.node_count = ctx->mesh_node_count[mesh];
        a.root = ctx->mesh_node_base[mesh];
        a.arrivals = ctx->d_node_arrivals;
        a.mesh_root = ctx->d_mesh_roots + mesh;
        a.mesh_bounds = ctx->d_mesh_bounds + 6 * (size_t)mesh;
        HIP_CHECK(ngx::launch_mesh_refit(a, s));
        ctx->top_level_stale = true;
    });
}

ngx_status ngx_set_shading(ngx_ctx *ctx, const ngx_shading *shading) {
    return guarded([&] {
        require(ctx != nullptr && shading != nullptr, "context and shading must not be null");
        const ngx_shading &h = *shading;
        require(ctx->per_world == 0 || h.instance_materials != nullptr, "instance_materials must hold one material per template instance");
        require(ctx->sphere_count == 0 || h.sphere_materials != nullptr, "sphere_materials must hold one material per sphere");
        for (int i = 0; i < ctx->per_world; ++i) check_material(h.instance_materials[i], "instance " + std::to_string(i));
        for (int i = 0; i < ctx->sphere_count; ++i) check_material(h.sphere_materials[i], "sphere " + std::to_string(i));
        check_material(h.floor_material, "floor");
        require(std::isfinite(h.floor_checker_scale), "floor_checker_scale must be finite");
        require(h.light_count >= 0 && h.light_count <= NGX_MAX_LIGHTS,
                "a scene holds at most " + std::to_string(NGX_MAX_LIGHTS) + " lights, got " + std::to_string(h.light_count));
        require(h.light_count == 0 || h.lights != nullptr, "lights must not be null");
        // validate.cpp:40-62
        for (int i = 0; i < h.light_count; ++i) {
            const ngx_light &l = h.lights[i];
            const std::string name = "light " + std::to_string(i);
            require(finite_all(l.position, 3) && std::isfinite(l.intensity) && std::isfinite(l.radius), name + " has a NaN or infinite value");
            require(l.radius >= 0.0f, name + " radius must be 0 or more, got " + std::to_string(l.radius));
            require(std::isfinite(l.falloff) && l.falloff >= 0.0f, name + " falloff must be 0 or more, got " + std::to_string(l.falloff));
            require(finite_all(l.direction, 3) && std::isfinite(l.spot_angle) && std::isfinite(l.spot_softness), name + " has a NaN or infinite value");
            require(l.spot_angle >= 0.0f && l.spot_angle < 3.14159f,
                    name + " spot_angle must be 0 (no cone) to below pi radians, got " + std::to_string(l.spot_angle));
            require(finite_all(l.color, 3) && l.color[0] >= 0.0f && l.color[1] >= 0.0f && l.color[2] >= 0.0f,
                    name + " color must be finite and not negative");
            require(l.spot_softness >= 0.0f && l.spot_softness <= 1.0f, name + " spot_softness must be 0 to 1, got " + std::to_string(l.spot_softness));
            require(l.spot_angle == 0.0f || l.direction[0] != 0.0f || l.direction[1] != 0.0f || l.direction[2] != 0.0f,
                    name + " is a spotlight (spot_angle > 0) but its direction is zero");
        }
        require(std::isfinite(h.ambient) && h.ambient >= 0.0f, "ambient must be 0 or more, got " + std::to_string(h.ambient));
        require(finite_all(h.sky_up, 3) && finite_all(h.sky_color_up, 3) && finite_all(h.sky_color_down, 3) && std::isfinite(h.sky_intensity),
                "the sky has a NaN or infinite value");
        const double un = std::sqrt((double)h.sky_up[0] * h.sky_up[0] + (double)h.sky_up[1] * h.sky_up[1] + (double)h.sky_up[2] * h.sky_up[2]);
        require(un > 0.0, "the sky's up axis must not be zero");
        ctx->set_device();
        HIP_CHECK(hipDeviceSynchronize());
        std::vector<float4> im(2 * (size_t)std::max(1, ctx->per_world)), sm(2 * (size_t)std::max(1, ctx->sphere_count)),
            lights(4 * (size_t)std::max(1, (int)h.light_count));
        for (int i = 0; i < ctx->per_world; ++i) pack_material(h.instance_materials[i], &im[2 * i]);
        for (int i = 0; i < ctx->sphere_count; ++i) pack_material(h.sphere_materials[i], &sm[2 * i]);
        for (int i = 0; i < h.light_count; ++i) {
            // the device form of a light (device_scene.h:77-84): unit direction, the cone as cosines
            const ngx_light &l = h.lights[i];
            float dx = l.direction[0], dy = l.direction[1], dz = l.direction[2];
            const float n = std::sqrt(dx * dx + dy * dy + dz * dz);
            if (n > 0.0f) {
                dx = dx / n;
                dy = dy / n;
                dz = dz / n;
            } else {
                dx = 0.0f;
                dy = 0.0f;
                dz = -1.0f;
            }
            const bool spot = l.spot_angle > 0.0f;
            const float cos_outer = spot ? std::cos(l.spot_angle) : -2.0f;
            const float cos_inner = spot ? std::cos(l.spot_angle * (1.0f - l.spot_softness)) : -2.0f;
            lights[4 * i] = make_float4(l.position[0], l.position[1], l.position[2], l.intensity);
            lights[4 * i + 1] = make_float4(l.radius, l.falloff, cos_outer, cos_inner);
            lights[4 * i + 2] = make_float4(dx, dy, dz, 0.0f);
            lights[4 * i + 3] = make_float4(l.color[0], l.color[1], l.color[2], 0.0f);
        }
        upload(ctx->d_inst_mat, im.data(), im.size());
        upload(ctx->d_sphere_mat, sm.data(), sm.size());
        upload(ctx->d_lights, lights.data(), lights.size());
        // a path's throughput starts at 1 and is multiplied by each surface's reflectivity; below the cut-off the
        // path ends, so with every reflectivity below it no path goes past its first hit
        bool no_bounce = h.floor_material.reflectivity < ngx::MIN_THROUGHPUT;
        for (int i = 0; i < ctx->per_world; ++i) no_bounce = no_bounce && h.instance_materials[i].reflectivity < ngx::MIN_THROUGHPUT;
        for (int i = 0; i < ctx->sphere_count; ++i) no_bounce = no_bounce && h.sphere_materials[i].reflectivity < ngx::MIN_THROUGHPUT;
        ctx->shade_no_bounce = no_bounce;
        ngx::ShadeScene &e = ctx->shade;
        e.instance_materials = ctx->d_inst_mat;
        e.sphere_materials = ctx->d_sphere_mat;
        pack_material(h.floor_material, e.floor_material);
        e.floor_checker_scale = h.floor_checker_scale;
        e.light_count = h.light_count;
        e.lights = ctx->d_lights;
        e.ambient = h.ambient;
        // the sky's up axis, normalized in FP64 as the reference's environment frame is (environment.cpp:30-33)
        e.sky_up = make_float3((float)(h.sky_up[0] / un), (float)(h.sky_up[1] / un), (float)(h.sky_up[2] / un));
        e.sky_color_up = make_float3(h.sky_color_up[0], h.sky_color_up[1], h.sky_color_up[2]);
        e.sky_color_down = make_float3(h.sky_color_down[0], h.sky_color_down[1], h.sky_color_down[2]);
        e.sky_intensity = h.sky_intensity;
        ctx->shading_set = true;
    });
}

ngx_status ngx_render_batch(ngx_ctx *ctx, const float *cams12, ngx_memory memory, float fov, int32_t spp, uint32_t outputs,
                            int32_t frame, void *stream, ngx_batch_frame *out) {
    return guarded([&] {
        require(ctx != nullptr && out != nullptr, "context and out must not be null");
        const unsigned all = NGX_OUTPUT_RGB | NGX_OUTPUT_DEPTH | NGX_OUTPUT_NORMAL | NGX_OUTPUT_OBJECT_ID;
        require(outputs != 0, "render needs at least one output");
        require((outputs & ~all) == 0, "unknown output bits in " + std::to_string(outputs));
        require(spp >= 1, "samples_per_pixel must be 1 or more, got " + std::to_string(spp));
        require(frame >= 0, "frame must be 0 or more, got " + std::to_string(frame));
        require(fov > 0.0f && fov < 3.14159265f, "fov must be between 0 and pi radians, got " + std::to_string(fov));
        require(cams12 != nullptr, "camera_matrices must not be null");
        require_state(!ctx->top_level_stale, "a mesh moved (ngx_set_mesh_vertices): set the instance matrices or body poses again first");
        if (memory == NGX_MEM_HOST) require(finite_all(cams12, 12 * (size_t)ctx->views), "a camera has a NaN or infinite value");
        if (outputs & NGX_OUTPUT_RGB)
            require_state(ctx->shading_set, "the RGB picture needs materials and lights: call ngx_set_shading first");
        ctx->set_device();
        const hipStream_t s = static_cast<hipStream_t>(stream);
        const size_t pixels = (size_t)ctx->views * ctx->width * ctx->height;
        if ((outputs & NGX_OUTPUT_DEPTH) && !ctx->d_depth) device_alloc(ctx->d_depth, pixels);
        if ((outputs & NGX_OUTPUT_NORMAL) && !ctx->d_normal) device_alloc(ctx->d_normal, 3 * pixels);
        if ((outputs & NGX_OUTPUT_OBJECT_ID) && !ctx->d_id) device_alloc(ctx->d_id, pixels);
        const float *cams = cams12;
        if (memory == NGX_MEM_HOST) {
            if (!ctx->d_cam_stage) device_alloc(ctx->d_cam_stage, 12 * (size_t)ctx->views);
            HIP_CHECK(hipMemcpyAsync(ctx->d_cam_stage, cams12, 12 * sizeof(float) * ctx->views, hipMemcpyHostToDevice, s));
            cams = ctx->d_cam_stage;
        }
        ngx::RenderArgs a = ctx->render_args(cams, fov);
        a.fov_scale = std::tan(fov * 0.5f);
        a.aspect = (float)ctx->width / (float)ctx->height;
        a.width = ctx->width;
        a.height = ctx->height;
        a.cameras_per_world = ctx->cpw;
        a.views = ctx->views;
        a.tile_w = ctx->tile_w;
        a.tile_h = 64 / ctx->tile_w;
        a.tiles_x = (ctx->width + a.tile_w - 1) / a.tile_w;
        a.tiles_per_view = a.tiles_x * ((ctx->height + a.tile_h - 1) / a.tile_h);
        a.total_tiles = (long long)a.tiles_per_view * ctx->views;
        a.depth = (outputs & NGX_OUTPUT_DEPTH) ? ctx->d_depth : nullptr;
        a.normal = (outputs & NGX_OUTPUT_NORMAL) ? ctx->d_normal : nullptr;
        a.object_id = (outputs & NGX_OUTPUT_OBJECT_ID) ? ctx->d_id : nullptr;
        a.xcd_remap = ctx->xcd_remap;
        a.walk = ctx->walk();
        a.blocks = ctx->render_blocks;
        uint8_t *rgb = nullptr;
        if (outputs & NGX_OUTPUT_RGB) {
            if (!ctx->d_rgb) device_alloc(ctx->d_rgb, 3 * pixels);
            rgb = ctx->d_rgb;
        }
        const bool timed = !ctx->capturing(s);
        if (timed) HIP_CHECK(hipEventRecord(ctx->ev0, s));
        if (a.depth || a.normal || a.object_id) HIP_CHECK(ngx::launch_render(a, s, ctx->persistent ? ctx->compute_units : 0));
        if (rgb) {
            ngx::RgbArgs g{};
            g.r = a;
            g.r.depth = nullptr;
            g.r.normal = nullptr;
            g.r.object_id = nullptr;
            g.r.tile_counters = ctx->persistent ? ctx->d_tile_counters : nullptr;
            g.shade = ctx->shade;
            g.rgb = rgb;
            g.spp = spp;
            g.frame = frame;
            g.blocks = ctx->rgb_blocks;
            bool two_pass = ctx->rgb_two_pass && spp == 1 && ctx->shade_no_bounce && ctx->persistent && ctx->compute_units > 0 &&
                            a.walk != 0;
            // the hit records are an optimization: where the device cannot hold them, the one-pass kernel draws the
            // same bytes
            if (two_pass && !try_device_alloc(ctx->d_rgb_hits, pixels)) two_pass = false;
            if (two_pass && ctx->d_attrs && !try_device_alloc(ctx->d_rgb_uv, pixels)) two_pass = false;
            if (two_pass) {
                g.hits = ctx->d_rgb_hits;
                g.hit_uv = ctx->d_attrs ? ctx->d_rgb_uv : nullptr;
                HIP_CHECK(ngx::launch_rgb_two_pass(g, ctx->render_blocks, s, ctx->compute_units));
            } else {
                HIP_CHECK(ngx::launch_rgb(g, s, ctx->persistent ? ctx->compute_units : 0));
            }
            ctx->last_rgb_passes = two_pass ? 2 : 1;
        }
        if (timed) HIP_CHECK(hipEventRecord(ctx->ev1, s));
        ctx->timed = timed;
        ctx->last_stream = s;
        *out = ngx_batch_frame{NGX_MEM_DEVICE, ctx->device, ctx->views, ctx->width, ctx->height, rgb, a.depth, a.normal, a.object_id};
    });
}

ngx_status ngx_cast_rays_batch(ngx_ctx *ctx, const float *origins, const float *directions, int32_t rays_per_world,
                               float max_distance, ngx_memory memory, void *stream, ngx_ray_hits *out) {
    return guarded([&] {
        require(ctx != nullptr && out != nullptr, "context and out must not be null");
        require(rays_per_world >= 0, "rays_per_world must be 0 or more, got " + std::to_string(rays_per_world));
        const long long count = (long long)rays_per_world * ctx->worlds;
        require(count <= INT32_MAX / 3, "too many rays: " + std::to_string(count));
        require(count == 0 || (origins != nullptr && directions != nullptr), "ray origins and directions must not be null");
        require(max_distance > 0.0f, "max_distance must be positive, got " + std::to_string(max_distance));
        require_state(!ctx->top_level_stale, "a mesh moved (ngx_set_mesh_vertices): set the instance matrices or body poses again first");
        if (memory == NGX_MEM_HOST) require(finite_all(origins, 3 * (size_t)count), "a ray origin has a NaN or infinite value");
        ctx->set_device();
        if (count == 0) {
            *out = ngx_ray_hits{NGX_MEM_DEVICE, ctx->device, 0, nullptr, nullptr, nullptr};
            return;
        }
        const hipStream_t s = static_cast<hipStream_t>(stream);
        const float *o = origins, *d = directions;
        if (memory == NGX_MEM_HOST) {
            if (count > ctx->ray_in_cap) {
                HIP_CHECK(hipStreamSynchronize(s));
                device_alloc(ctx->d_ray_o, 3 * (size_t)count);
                device_alloc(ctx->d_ray_d, 3 * (size_t)count);
                ctx->ray_in_cap = count;
            }
            HIP_CHECK(hipMemcpyAsync(ctx->d_ray_o, origins, 3 * sizeof(float) * count, hipMemcpyHostToDevice, s));
            HIP_CHECK(hipMemcpyAsync(ctx->d_ray_d, directions, 3 * sizeof(float) * count, hipMemcpyHostToDevice, s));
            o = ctx->d_ray_o;
            d = ctx->d_ray_d;
        }
        if (count > ctx->ray_cap) {
            HIP_CHECK(hipStreamSynchronize(s));
            device_alloc(ctx->d_hit_dist, (size_t)count);
            device_alloc(ctx->d_hit_normal, 3 * (size_t)count);
            device_alloc(ctx->d_hit_id, (size_t)count);
            device_alloc(ctx->d_ray_keys, 2 * (size_t)count);
            device_alloc(ctx->d_ray_vals, (size_t)count);
            ctx->ray_cap = count;
        }
        ngx::RayArgs a{};
        a.s = ctx->trace_scene();
        a.origins = o;
        a.dirs = d;
        a.rays_per_world = rays_per_world;
        a.count = count;
        a.max_distance = max_distance;
        a.distance = ctx->d_hit_dist;
        a.normal = ctx->d_hit_normal;
        a.object_id = ctx->d_hit_id;
        a.walk = ctx->walk_rays();
        a.blocks = ctx->ray_blocks;
        const bool timed = !ctx->capturing(s);
        if (timed) HIP_CHECK(hipEventRecord(ctx->ev0, s));
        if (ctx->sort_rays && a.walk != 0 && rays_per_world >= 1024 && rays_per_world <= ngx::RAY_ORDER_MAX_RAYS) {
            int *order = nullptr;
            HIP_CHECK(ngx::launch_ray_order(d, rays_per_world, ctx->worlds, ctx->d_ray_keys, ctx->d_ray_vals, &order, s));
            a.order = order;
            // contiguous ranges per XCD pay only with the direction order (lidar_sweeps: 23.4 -> 22.8 ms; in the
            // caller's order they cost 23.7 -> 30.0 ms)
            a.xcd_remap = ctx->xcd_remap;
        }
        HIP_CHECK(ngx::launch_rays(a, s));
        if (timed) HIP_CHECK(hipEventRecord(ctx->ev1, s));
        ctx->timed = timed;
        ctx->last_stream = s;
        *out = ngx_ray_hits{NGX_MEM_DEVICE, ctx->device, (int32_t)count, ctx->d_hit_dist, ctx->d_hit_normal, ctx->d_hit_id};
    });
}

namespace {

// The depth backward of ngx_pose_backward_depth and ngx_pose_backward_depth_noalloc. With no_alloc, the
// per-workgroup partial sums go to the device's fixed scratch array (looked up when the context was created)
// and nothing is allocated or waited for; otherwise to the context's buffer, grown as needed.
void pose_backward_depth(ngx_ctx *ctx, const float *body_poses, const float *cams12, const float fov, const float *depth,
                         const float *grad_depth, float *grad_out, void *stream, const bool no_alloc) {
    require(ctx != nullptr, "context is null");
    require(ctx->bound && ctx->bodies > 0, "the pose backward needs bodies: call bind_bodies first");
    if (ctx->bodies > ngx::BACKWARD_MAX_BODIES)
        fail(NGX_UNSUPPORTED, "the pose backward takes at most " + std::to_string(ngx::BACKWARD_MAX_BODIES) +
                                  " bodies per world; the binding has " + std::to_string(ctx->bodies));
    require(body_poses && cams12 && depth && grad_depth && grad_out, "device arrays must not be null");
    require(fov > 0.0f && fov < 3.14159265f, "fov must be between 0 and pi radians, got " + std::to_string(fov));
    // enough workgroups to fill the GPU: each view's tiles are split over `parts` workgroups of 4 waves
    const int tiles = ((ctx->width + 7) / 8) * ((ctx->height + 7) / 8);
    const int parts = std::max(1, std::min(16, tiles / 8));
    const size_t need = (size_t)ctx->views * parts * ctx->bodies * 7;
    if (no_alloc)
        require(need <= ngx::NOALLOC_SCRATCH_FLOATS,
                "the allocation-free backward holds " + std::to_string(ngx::NOALLOC_SCRATCH_FLOATS) + " floats of partial sums; " +
                    std::to_string(ctx->views) + " views x " + std::to_string(parts) + " parts x " + std::to_string(ctx->bodies) +
                    " bodies x 7 need " + std::to_string(need) + ": use ngx_pose_backward_depth");
    ctx->set_device();
    const hipStream_t s = static_cast<hipStream_t>(stream);
    float *partial = ctx->d_noalloc_scratch;
    if (!no_alloc) {
        if (need > ctx->partial_cap) {
            HIP_CHECK(hipStreamSynchronize(s));
            device_alloc(ctx->d_partial, need);
            ctx->partial_cap = need;
        }
        partial = ctx->d_partial;
    }
    ctx->build_from_bodies(body_poses, s);
    ngx::DepthBackwardArgs a{};
    a.s = ctx->trace_scene();
    a.cams = cams12;
    a.fov_scale = std::tan(fov * 0.5f);
    a.aspect = (float)ctx->width / (float)ctx->height;
    a.width = ctx->width;
    a.height = ctx->height;
    a.cameras_per_world = ctx->cpw;
    a.views = ctx->views;
    a.depth = depth;
    a.grad_depth = grad_depth;
    a.bodies_per_world = ctx->bodies;
    a.body_of_instance = ctx->d_body_of_instance;
    a.local_rot_scale = ctx->d_local_rot_scale;
    a.body_poses = body_poses;
    a.parts = parts;
    a.partial = partial;
    a.walk = ctx->walk();
    a.blocks = ctx->backward_blocks;
    // the backward kernel alone is what ngx_last_kernel_ms reports after this call
    const bool timed = !ctx->capturing(s);
    if (timed) HIP_CHECK(hipEventRecord(ctx->ev0, s));
    HIP_CHECK(ngx::launch_depth_backward(a, s));
    if (timed) HIP_CHECK(hipEventRecord(ctx->ev1, s));
    ctx->timed = timed;
    ngx::PoseGradFinishArgs f{ctx->worlds, ctx->bodies, ctx->cpw * parts, partial, body_poses, grad_out};
    HIP_CHECK(ngx::launch_pose_grad_finish(f, s));
    ctx->last_stream = s;
}

}  // namespace

ngx_status ngx_pose_backward_depth(ngx_ctx *ctx, const float *body_poses, const float *cams12, float fov, const float *depth,
                                   const float *grad_depth, float *grad_out, void *stream) {
    return guarded([&] { pose_backward_depth(ctx, body_poses, cams12, fov, depth, grad_depth, grad_out, stream, false); });
}

ngx_status ngx_pose_backward_depth_noalloc(ngx_ctx *ctx, const float *body_poses, const float *cams12, float fov, const float *depth,
                                           const float *grad_depth, float *grad_out, void *stream) {
    return guarded([&] { pose_backward_depth(ctx, body_poses, cams12, fov, depth, grad_depth, grad_out, stream, true); });
}

ngx_status ngx_pose_backward_range(ngx_ctx *ctx, const float *body_poses, const float *origins, const float *directions,
                                   int32_t rays_per_world, float max_distance, const float *distance, const float *grad_distance,
                                   float *grad_out, void *stream) {
    return guarded([&] {
        require(ctx != nullptr, "context is null");
        require(ctx->bound && ctx->bodies > 0, "the pose backward needs bodies: call bind_bodies first");
        if (ctx->bodies > ngx::BACKWARD_MAX_BODIES)
            fail(NGX_UNSUPPORTED, "the pose backward takes at most " + std::to_string(ngx::BACKWARD_MAX_BODIES) +
                                      " bodies per world; the binding has " + std::to_string(ctx->bodies));
        require(body_poses && origins && directions && distance && grad_distance && grad_out, "device arrays must not be null");
        require(rays_per_world >= 0, "rays_per_world must be 0 or more");
        require(max_distance > 0.0f, "max_distance must be positive, got " + std::to_string(max_distance));
        ctx->set_device();
        const hipStream_t s = static_cast<hipStream_t>(stream);
        const int chunks = std::max(1, std::min(64, rays_per_world / 4096));
        const size_t need = (size_t)ctx->worlds * chunks * ctx->bodies * 7;
        if (need > ctx->partial_cap) {
            HIP_CHECK(hipStreamSynchronize(s));
            device_alloc(ctx->d_partial, need);
            ctx->partial_cap = need;
        }
        ctx->build_from_bodies(body_poses, s);
        ngx::RangeBackwardArgs a{};
        a.s = ctx->trace_scene();
        a.origins = origins;
        a.dirs = directions;
        a.rays_per_world = rays_per_world;
        a.chunks = chunks;
        a.walk = ctx->walk_rays();
        a.max_distance = max_distance;
        a.distance = distance;
        a.grad_distance = grad_distance;
        a.bodies_per_world = ctx->bodies;
        a.body_of_instance = ctx->d_body_of_instance;
        a.local_rot_scale = ctx->d_local_rot_scale;
        a.body_poses = body_poses;
        a.worlds = ctx->worlds;
        a.partial = ctx->d_partial;
        const bool timed = !ctx->capturing(s);
        if (timed) HIP_CHECK(hipEventRecord(ctx->ev0, s));
        HIP_CHECK(ngx::launch_range_backward(a, s));
        if (timed) HIP_CHECK(hipEventRecord(ctx->ev1, s));
        ctx->timed = timed;
        ngx::PoseGradFinishArgs f{ctx->worlds, ctx->bodies, chunks, ctx->d_partial, body_poses, grad_out};
        HIP_CHECK(ngx::launch_pose_grad_finish(f, s));
        ctx->last_stream = s;
    });
}

ngx_status ngx_copy_to_host(ngx_ctx *ctx, const void *src, void *dst, size_t bytes) {
    return guarded([&] {
        require(ctx != nullptr, "context is null");
        ctx->set_device();
        if (bytes) HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    });
}

ngx_status ngx_synchronize(ngx_ctx *ctx) {
    return guarded([&] {
        require(ctx != nullptr, "context is null");
        ctx->set_device();
        HIP_CHECK(hipDeviceSynchronize());
        int overflow = 0;
        HIP_CHECK(hipMemcpy(&overflow, ctx->d_overflow, sizeof(int), hipMemcpyDeviceToHost));
        if (overflow) {
            HIP_CHECK(hipMemset(ctx->d_overflow, 0, sizeof(int)));
            fail(NGX_DEVICE_ERROR, "a tree walk ran out of stack (a tree deeper than " + std::to_string(ngx::WALK_STACK - 1) +
                                       " levels below a root): the images or ray hits since the last synchronize may be wrong");
        }
    });
}

float ngx_last_kernel_ms(ngx_ctx *ctx) {
    if (!ctx || !ctx->timed) return 0.0f;
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1) != hipSuccess) {
        (void)hipGetLastError();
        return 0.0f;
    }
    return ms;
}

int32_t ngx_world_count(const ngx_ctx *ctx) { return ctx ? ctx->worlds : 0; }

ngx_status ngx_frame_size(const ngx_ctx *ctx, int32_t *width, int32_t *height) {
    return guarded([&] {
        require(ctx != nullptr && width != nullptr && height != nullptr, "arguments must not be null");
        *width = ctx->width;
        *height = ctx->height;
    });
}
int32_t ngx_view_count(const ngx_ctx *ctx) { return ctx ? ctx->views : 0; }
int32_t ngx_instances_per_world(const ngx_ctx *ctx) { return ctx ? ctx->per_world : 0; }
int32_t ngx_instance_count(const ngx_ctx *ctx) { return ctx ? ctx->total : 0; }
int32_t ngx_device_id(const ngx_ctx *ctx) { return ctx ? ctx->device : 0; }

}  // extern "C"

namespace {

template <size_t N>
bool built_variant(const int blocks, const int (&list)[N]) {
    if (blocks == 0) return true;
    for (const int b : list)
        if (b == blocks) return true;
    return false;
}

ngx_trace_config current_config(const ngx_ctx *ctx) {
    return ngx_trace_config{ctx->tile_w,      ctx->xcd_remap,  ctx->packet,     ctx->pair_limit,
                            ctx->persistent,  ctx->lane_min_tris, ctx->sort_rays, ctx->render_blocks,
                            ctx->rgb_blocks,  ctx->ray_blocks, ctx->backward_blocks, ctx->rgb_two_pass};
}

// Times `runs` calls of `call` on the stream after one untimed call; returns ms per call, or -1 when a call fails
// (the failure's status in *failure).
template <class Call>
float time_calls(const hipStream_t s, const int reps, ngx_status *failure, Call &&call) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    float ms = -1.0f;
    *failure = call();
    if (*failure == NGX_OK) {
        (void)hipEventRecord(e0, s);
        for (int r = 0; r < reps && *failure == NGX_OK; ++r) *failure = call();
        (void)hipEventRecord(e1, s);
        (void)hipEventSynchronize(e1);
        if (*failure == NGX_OK) {
            (void)hipEventElapsedTime(&ms, e0, e1);
            ms /= (float)reps;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ms;
}

// Times each configuration with `call`; on a failure the context gets back `start` and the failure is raised.
template <class Call>
std::vector<float> time_each(ngx_ctx *ctx, const std::vector<ngx_trace_config> &cands, const ngx_trace_config &start,
                             const hipStream_t s, const int reps, Call &&call) {
    std::vector<float> times;
    for (const ngx_trace_config &c : cands) {
        ngx_set_trace_config(ctx, &c);
        ngx_status failure = NGX_OK;
        const float ms = time_calls(s, reps, &failure, call);
        if (failure != NGX_OK) {
            const std::string message = g_last_error;
            ngx_set_trace_config(ctx, &start);
            fail(failure, "autotune: " + message);
        }
        times.push_back(ms);
    }
    return times;
}

// Tries each configuration with `call` and keeps the fastest; on a failure the context gets back `start` and the
// failure is raised.
template <class Call>
void keep_fastest(ngx_ctx *ctx, const std::vector<ngx_trace_config> &cands, const ngx_trace_config &start, const hipStream_t s,
                  const int reps, ngx_trace_config &winner, float &best_time, Call &&call) {
    for (const ngx_trace_config &c : cands) {
        ngx_set_trace_config(ctx, &c);
        ngx_status failure = NGX_OK;
        const float ms = time_calls(s, reps, &failure, call);
        if (failure != NGX_OK) {
            const std::string message = g_last_error;
            ngx_set_trace_config(ctx, &start);
            fail(failure, "autotune: " + message);
        }
        if (best_time <= 0.0f || ms < best_time) {
            best_time = ms;
            winner = c;
        }
    }
}

}  // namespace

extern "C" {

ngx_status ngx_set_trace_config(ngx_ctx *ctx, const ngx_trace_config *config) {
    return guarded([&] {
        require(ctx != nullptr && config != nullptr, "context and config must not be null");
        const ngx_trace_config &c = *config;
        require(c.tile_w == 8 || c.tile_w == 16 || c.tile_w == 32 || c.tile_w == 64, "tile_w must be 8, 16, 32 or 64");
        require(c.pair_limit >= 1 && c.pair_limit <= 64, "pair_limit must be 1 to 64");
        require(c.lane_min_triangles >= 0, "lane_min_triangles must not be negative");
        require(built_variant(c.render_blocks, ngx::RENDER_BLOCKS), "render_blocks must be 0, 5, 6 or 7");
        require(built_variant(c.rgb_blocks, ngx::RGB_BLOCKS), "rgb_blocks must be 0, 4, 5, 6 or 7");
        require(built_variant(c.ray_blocks, ngx::RAY_BLOCKS), "ray_blocks must be 0, 5, 6 or 7");
        require(built_variant(c.backward_blocks, ngx::BACKWARD_BLOCKS), "backward_blocks must be 0, 4, 5 or 6");
        ctx->tile_w = c.tile_w;
        ctx->xcd_remap = c.xcd_remap ? 1 : 0;
        ctx->packet = c.packet ? 1 : 0;
        ctx->pair_limit = c.pair_limit;
        ctx->persistent = c.persistent ? 1 : 0;
        ctx->lane_min_tris = c.lane_min_triangles;
        ctx->sort_rays = c.sort_rays ? 1 : 0;
        ctx->render_blocks = c.render_blocks;
        ctx->rgb_blocks = c.rgb_blocks;
        ctx->ray_blocks = c.ray_blocks;
        ctx->backward_blocks = c.backward_blocks;
        ctx->rgb_two_pass = c.rgb_two_pass ? 1 : 0;
    });
}

ngx_status ngx_get_trace_config(const ngx_ctx *ctx, ngx_trace_config *config) {
    return guarded([&] {
        require(ctx != nullptr && config != nullptr, "context and config must not be null");
        *config = current_config(ctx);
    });
}

ngx_status ngx_autotune(ngx_ctx *ctx, const float *cams12, ngx_memory memory, float fov, uint32_t outputs, int32_t reps,
                        void *stream, ngx_trace_config *best, float *best_ms) {
    return guarded([&] {
        require(ctx != nullptr && cams12 != nullptr, "context and cameras must not be null");
        require(reps >= 1 && reps <= 100, "reps must be 1 to 100");
        const hipStream_t s = static_cast<hipStream_t>(stream);
        require(!ctx->capturing(s), "autotune cannot run while the stream is being captured");
        ctx->set_device();
        const ngx_trace_config start = current_config(ctx);
        auto render = [&] {
            ngx_batch_frame f;
            return ngx_render_batch(ctx, cams12, memory, fov, 1, outputs, 0, stream, &f);
        };
        // first the walk and tiles, with each walk's default kernel variants: the per-lane bottom level for no
        // mesh, meshes of 1000+ triangles, of 300+, every mesh
        const int never = 1 << 30;
        std::vector<ngx_trace_config> cands;
        for (const int tw : {8, 16, 32}) {
            for (const int lanes : {never, 1000, 300, 0})
                for (const int limit : {32, 64})
                    for (const int pers : {0, 1}) cands.push_back(ngx_trace_config{tw, 1, 1, limit, pers, lanes, start.sort_rays, 0, 0,
                                                                                   start.ray_blocks, start.backward_blocks,
                                                                                   start.rgb_two_pass});
            if (tw <= 16)
                cands.push_back(ngx_trace_config{tw, 1, 0, 64, 0, never, start.sort_rays, 0, 0, start.ray_blocks, start.backward_blocks,
                                                 start.rgb_two_pass});
        }
        // every candidate timed briefly, then the four fastest again over four times the renders: configurations a
        // percent or two apart are otherwise told apart by noise
        const std::vector<float> first = time_each(ctx, cands, start, s, reps, render);
        std::vector<int> rank(cands.size());
        for (size_t i = 0; i < rank.size(); ++i) rank[i] = (int)i;
        std::stable_sort(rank.begin(), rank.end(), [&](const int a, const int b) { return first[a] < first[b]; });
        std::vector<ngx_trace_config> finalists;
        for (size_t i = 0; i < rank.size() && i < 4; ++i) finalists.push_back(cands[rank[i]]);
        ngx_trace_config winner = start;
        float best_time = 0.0f;
        keep_fastest(ctx, finalists, start, s, 4 * reps, winner, best_time, render);
        // then, for the winner's walk, the kernel variants: workgroups per compute unit of the depth / normal / ID
        // kernel and of the RGB kernel (the per-lane walk has one variant of each)
        if (winner.packet) {
            const ngx_trace_config base = winner;
            cands.clear();
            if (outputs & (NGX_OUTPUT_DEPTH | NGX_OUTPUT_NORMAL | NGX_OUTPUT_OBJECT_ID))
                for (const int b : ngx::RENDER_BLOCKS) {
                    ngx_trace_config c = base;
                    c.render_blocks = b;
                    cands.push_back(c);
                }
            keep_fastest(ctx, cands, start, s, 2 * reps, winner, best_time, render);
            if (outputs & NGX_OUTPUT_RGB) {
                const ngx_trace_config base2 = winner;
                cands.clear();
                // the two-pass picture where it applies (one sample, no bounces), else the one-pass kernel alone
                for (const int two : {1, 0}) {
                    if (two && !ctx->shade_no_bounce) continue;
                    for (const int b : ngx::RGB_BLOCKS) {
                        ngx_trace_config c = base2;
                        c.rgb_blocks = b;
                        c.rgb_two_pass = two;
                        cands.push_back(c);
                    }
                }
                keep_fastest(ctx, cands, start, s, 2 * reps, winner, best_time, render);
            }
        }
        ngx_set_trace_config(ctx, &winner);
        if (best) *best = winner;
        if (best_ms) *best_ms = best_time;
    });
}

ngx_status ngx_autotune_rays(ngx_ctx *ctx, const float *origins, const float *directions, int32_t rays_per_world,
                             float max_distance, ngx_memory memory, int32_t reps, void *stream, ngx_trace_config *best,
                             float *best_ms) {
    return guarded([&] {
        require(ctx != nullptr && origins != nullptr && directions != nullptr, "context and rays must not be null");
        require(reps >= 1 && reps <= 100, "reps must be 1 to 100");
        const hipStream_t s = static_cast<hipStream_t>(stream);
        require(!ctx->capturing(s), "autotune cannot run while the stream is being captured");
        ctx->set_device();
        const ngx_trace_config start = current_config(ctx);
        std::vector<ngx_trace_config> cands;
        for (const int sorted : {1, 0})
            for (const int b : ngx::RAY_BLOCKS) {
                ngx_trace_config c = start;
                c.sort_rays = sorted;
                c.ray_blocks = b;
                cands.push_back(c);
            }
        ngx_trace_config winner = start;
        float best_time = 0.0f;
        keep_fastest(ctx, cands, start, s, reps, winner, best_time, [&] {
            ngx_ray_hits h;
            return ngx_cast_rays_batch(ctx, origins, directions, rays_per_world, max_distance, memory, stream, &h);
        });
        ngx_set_trace_config(ctx, &winner);
        if (best) *best = winner;
        if (best_ms) *best_ms = best_time;
    });
}

ngx_status ngx_autotune_backward(ngx_ctx *ctx, const float *body_poses, const float *cams12, float fov, const float *depth,
                                 const float *grad_depth, float *grad_out, int32_t reps, void *stream, ngx_trace_config *best,
                                 float *best_ms) {
    return guarded([&] {
        require(ctx != nullptr, "context is null");
        require(reps >= 1 && reps <= 100, "reps must be 1 to 100");
        const hipStream_t s = static_cast<hipStream_t>(stream);
        require(!ctx->capturing(s), "autotune cannot run while the stream is being captured");
        ctx->set_device();
        const ngx_trace_config start = current_config(ctx);
        std::vector<ngx_trace_config> cands;
        if (start.packet)
            for (const int b : ngx::BACKWARD_BLOCKS) {
                ngx_trace_config c = start;
                c.backward_blocks = b;
                cands.push_back(c);
            }
        else
            cands.push_back(start);
        ngx_trace_config winner = start;
        float best_time = 0.0f;
        keep_fastest(ctx, cands, start, s, reps, winner, best_time, [&] {
            return ngx_pose_backward_depth(ctx, body_poses, cams12, fov, depth, grad_depth, grad_out, stream);
        });
        ngx_set_trace_config(ctx, &winner);
        if (best) *best = winner;
        if (best_ms) *best_ms = best_time;
    });
}

ngx_status ngx_debug_keep_build(ngx_ctx *ctx, int32_t keep) {
    return guarded([&] {
        require(ctx != nullptr, "context is null");
        ctx->set_device();
        ctx->keep_build = keep != 0;
        if (ctx->keep_build && !ctx->d_dbg_keys && ctx->total > 0) {
            device_alloc(ctx->d_dbg_keys, (size_t)ctx->total);
            device_alloc(ctx->d_dbg_sorted_keys, (size_t)ctx->total);
            device_alloc(ctx->d_dbg_sorted_ids, (size_t)ctx->total);
            device_alloc(ctx->d_dbg_aabb, 6 * (size_t)ctx->total);
            device_alloc(ctx->d_dbg_world_bounds, 6 * (size_t)ctx->worlds);
        }
    });
}

ngx_status ngx_debug_read_build(ngx_ctx *ctx, void *nodes, int32_t *roots, float *inv, float *matrices, uint64_t *keys,
                                uint64_t *sorted_keys, int32_t *sorted_ids, float *aabb, float *world_bounds) {
    return guarded([&] {
        require(ctx != nullptr, "context is null");
        ctx->set_device();
        HIP_CHECK(hipDeviceSynchronize());
        const size_t n = (size_t)ctx->total;
        if (roots) HIP_CHECK(hipMemcpy(roots, ctx->d_roots, sizeof(int) * ctx->worlds, hipMemcpyDeviceToHost));
        if (n == 0) return;
        if (nodes) HIP_CHECK(hipMemcpy(nodes, ctx->d_nodes, sizeof(ngx::TlasNode) * (2 * n - 1), hipMemcpyDeviceToHost));
        if (matrices) HIP_CHECK(hipMemcpy(matrices, ctx->d_matrices, sizeof(float) * 12 * n, hipMemcpyDeviceToHost));
        if (inv) {
            std::vector<ngx::InstanceRec> rec(n);
            HIP_CHECK(hipMemcpy(rec.data(), ctx->d_inst, sizeof(ngx::InstanceRec) * n, hipMemcpyDeviceToHost));
            for (size_t g = 0; g < n; ++g)
                for (int row = 0; row < 3; ++row) {
                    inv[12 * g + 4 * row] = rec[g].inv[row].x;
                    inv[12 * g + 4 * row + 1] = rec[g].inv[row].y;
                    inv[12 * g + 4 * row + 2] = rec[g].inv[row].z;
                    inv[12 * g + 4 * row + 3] = rec[g].inv[row].w;
                }
        }
        const bool want_dbg = keys || sorted_keys || sorted_ids || aabb || world_bounds;
        if (want_dbg) require(ctx->keep_build && ctx->d_dbg_keys, "build intermediates were not kept: call ngx_debug_keep_build(ctx, 1) first");
        if (keys) HIP_CHECK(hipMemcpy(keys, ctx->d_dbg_keys, sizeof(uint64_t) * n, hipMemcpyDeviceToHost));
        if (sorted_keys) HIP_CHECK(hipMemcpy(sorted_keys, ctx->d_dbg_sorted_keys, sizeof(uint64_t) * n, hipMemcpyDeviceToHost));
        if (sorted_ids) HIP_CHECK(hipMemcpy(sorted_ids, ctx->d_dbg_sorted_ids, sizeof(int) * n, hipMemcpyDeviceToHost));
        if (aabb) HIP_CHECK(hipMemcpy(aabb, ctx->d_dbg_aabb, sizeof(float) * 6 * n, hipMemcpyDeviceToHost));
        if (world_bounds) HIP_CHECK(hipMemcpy(world_bounds, ctx->d_dbg_world_bounds, sizeof(float) * 6 * ctx->worlds, hipMemcpyDeviceToHost));
    });
}

ngx_status ngx_build_skip_host(const float *vertices, int32_t vertex_count, const uint32_t *indices, int32_t triangle_count,
                               void *skip_nodes, int32_t *node_count) {
    return guarded([&] {
        require(vertices && indices && skip_nodes && node_count, "arrays must not be null");
        require(triangle_count > 0 && vertex_count > 0, "a mesh needs triangles");
        for (int i = 0; i < 3 * triangle_count; ++i) require(indices[i] < (uint32_t)vertex_count, "index out of range");
        const int32_t vo[2] = {0, vertex_count}, io[2] = {0, 3 * triangle_count};
        const ngx::host::MeshInput mi{1, vertices, vo, indices, io, nullptr, nullptr};
        ngx::host::BuiltMeshes built;
        ngx::host::build_meshes(mi, built);
        std::memcpy(skip_nodes, built.skip.data(), sizeof(ngx::SkipNode) * built.skip.size());
        *node_count = (int32_t)built.skip.size();
    });
}

ngx_status ngx_build_blas_host(const float *vertices, int32_t vertex_count, const uint32_t *indices, int32_t triangle_count,
                               void *nodes, int32_t *node_count, float *triangles, uint32_t *order) {
    return guarded([&] {
        require(vertices && indices && nodes && node_count && triangles && order, "arrays must not be null");
        require(triangle_count > 0 && vertex_count > 0, "a mesh needs triangles");
        for (int i = 0; i < 3 * triangle_count; ++i) require(indices[i] < (uint32_t)vertex_count, "index out of range");
        std::vector<ngx::host::Box> tb(triangle_count);
        for (int t = 0; t < triangle_count; ++t) {
            ngx::host::Box b{{INFINITY, INFINITY, INFINITY}, {-INFINITY, -INFINITY, -INFINITY}};
            for (int k = 0; k < 3; ++k)
                for (int a = 0; a < 3; ++a) {
                    b.lo[a] = std::fmin(b.lo[a], vertices[3 * (size_t)indices[3 * t + k] + a]);
                    b.hi[a] = std::fmax(b.hi[a], vertices[3 * (size_t)indices[3 * t + k] + a]);
                }
            tb[t] = b;
        }
        std::vector<ngx::BlasNode> tree;
        std::vector<uint32_t> ord;
        const int nn = ngx::host::sah_build(tb.data(), triangle_count, 4, tree, ord);
        std::memcpy(nodes, tree.data(), sizeof(ngx::BlasNode) * nn);
        std::memcpy(order, ord.data(), sizeof(uint32_t) * ord.size());
        *node_count = nn;
        for (int k = 0; k < triangle_count; ++k) {
            const uint32_t t = ord[k];
            const float *a = vertices + 3 * (size_t)indices[3 * t], *b = vertices + 3 * (size_t)indices[3 * t + 1],
                        *c = vertices + 3 * (size_t)indices[3 * t + 2];
            float *o = triangles + 12 * (size_t)k;
            const float packed[12] = {a[0], a[1], a[2], 0.0f, b[0] - a[0], b[1] - a[1], b[2] - a[2], 0.0f,
                                      c[0] - a[0], c[1] - a[1], c[2] - a[2], 0.0f};
            std::memcpy(o, packed, sizeof(packed));
        }
    });
}

ngx_status ngx_debug_rgb_passes(const ngx_ctx *ctx, int32_t *passes) {
    return guarded([&] {
        require(ctx != nullptr && passes != nullptr, "context and passes must not be null");
        *passes = ctx->last_rgb_passes;
    });
}

ngx_status ngx_debug_set_share_capacity(ngx_ctx *ctx, int32_t entries) {
    return guarded([&] {
        require(ctx != nullptr, "context must not be null");
        require(entries >= 1 && entries <= ngx::SHARE_STACK,
                "entries must be 1 to " + std::to_string(ngx::SHARE_STACK) + ", got " + std::to_string(entries));
        ctx->share_cap = entries;
    });
}

#ifdef NGX_TRACE_STATS
// Instrumented builds only: the render kernel's traversal counters summed over all waves (device cameras).
ngx_status ngx_debug_render_stats(ngx_ctx *ctx, const float *cams_device, float fov, long long *sums) {
    return guarded([&] {
        require(ctx != nullptr && cams_device && sums, "arguments must not be null");
        ctx->set_device();
        const ngx::RenderArgs a = ctx->render_args(cams_device, fov);
        int *d = nullptr;
        device_alloc(d, (size_t)a.total_tiles * ngx::TRACE_STAT_COUNT);
        HIP_CHECK(ngx::launch_render_stats(a, d, nullptr));
        std::vector<int> h((size_t)a.total_tiles * ngx::TRACE_STAT_COUNT);
        HIP_CHECK(hipMemcpy(h.data(), d, h.size() * sizeof(int), hipMemcpyDeviceToHost));
        (void)hipFree(d);
        for (int k = 0; k < ngx::TRACE_STAT_COUNT; ++k) sums[k] = 0;
        for (size_t i = 0; i < h.size(); ++i) sums[i % ngx::TRACE_STAT_COUNT] += h[i];
        sums[ngx::TRACE_STAT_COUNT] = a.total_tiles;
    });
}
#endif

}  // extern "C"
