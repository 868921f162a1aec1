// Kernel argument blocks and launchers of the MI355X backend. Included by the HIP translation units
// that define the kernels and by the host code (ngx_api.cpp) that owns the buffers and launches them.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "ngx_device.h"

namespace ngx {

// The instances of a batch are world-major: world w holds template entries 0 .. count(w)-1 at global
// positions world_off[w] .. world_off[w+1]-1. The reference has count(w) equal for every world.
// Where the build takes instance placements from: given matrices, body poses composed with a body binding,
// or simulator geom frames (position and rotation matrix per geom) scaled by a geom binding.
enum { SOURCE_MATRICES = 0, SOURCE_BODIES = 1, SOURCE_GEOMS = 2 };

// The persistent launches' tile counters: each XCD's range of tiles is cut into COUNTER_SLICES slices with a
// counter each, every counter on a 256-byte line of its own (COUNTER_STRIDE ints apart), so the counter updates
// of the whole chip spread over 8 x COUNTER_SLICES lines instead of queueing on one.
#ifndef NGX_COUNTER_SLICES
#define NGX_COUNTER_SLICES 8
#endif
constexpr int COUNTER_SLICES = NGX_COUNTER_SLICES, COUNTER_STRIDE = 64;
constexpr int TILE_COUNTER_INTS = 8 * COUNTER_SLICES * COUNTER_STRIDE;

struct TlasBuildArgs {
    int worlds;
    int total;                 // N = world_off[worlds]
    const int *world_off;      // [worlds + 1]
    const uint8_t *root_at_end; // [worlds]: the world's subtree root is node world_off[w+1]-1, not world_off[w]

    int source;                // SOURCE_MATRICES, SOURCE_BODIES or SOURCE_GEOMS
    const float *matrices_in;  // world w's instance i at 12 * (w * matrices_stride + i)
    int matrices_stride;       // instances per world in matrices_in (the template size)
    const float *body_poses;   // [worlds][bodies_per_world][7]
    int bodies_per_world;
    const int *body_of_instance; // [P]
    const float *local_pose;   // [P][7]
    const float *scale;        // [P][3]
    const float *geom_pos;     // [worlds][geoms_per_world][3]: each geom's world position
    const float *geom_mat;     // [worlds][geoms_per_world][9]: each geom's world rotation, row-major
    int geoms_per_world;
    const int *geom_of_instance; // [P]
    const float *geom_scale;   // [P][3]
    float *matrices_out;       // [N][12]: every instance's matrix as placed

    const int *tmpl_mesh;      // [P]
    const float *mesh_bounds;  // [meshes][6]
    const int *mesh_tris;      // [meshes]: triangle counts
    const int *mesh_skip_root; // [meshes]: root of each mesh's skip layout

    InstanceRec *inst;         // [N]
    TlasNode *nodes;           // [2N-1], the reference layout
    WideNode *wide;            // [N-1], the traversal layout
    int *wide_parent;          // [N-1]: the wide node whose child each one is, -1 for a world's root
    int *roots;                // [worlds]: root in `nodes` (a leaf index >= N-1 for a one-instance world), -1 if empty

    // Optional build intermediates, for checking the build against the reference algorithm (null: skipped).
    uint64_t *dbg_keys;        // [N] by instance: world << 32 | morton
    uint64_t *dbg_sorted_keys; // [N] in sorted order
    int *dbg_sorted_ids;       // [N]
    float *dbg_aabb;           // [N][6]
    float *dbg_world_bounds;   // [worlds][6]: min and max of the instances' box centres
};

struct TraceScene {
    const WideNode *tlas_wide;
    const int *tlas_parent;    // per wide node, -1 at a world's root
    const int *blas_parent;    // per bottom-level wide node, -1 at a mesh's root
    const int *blas_wide_node; // per bottom-level wide node, its node in the reference layout (and the skip layout)
    const int *roots;
    int total;
    const InstanceRec *inst;
    const MeshRoot *mesh_roots;
    const WideNode *blas_wide;
    const float *tris;         // 9 floats per triangle: v0, e1, e2 (packed: the walks fetch 36 bytes a triangle)
    const float4 *attrs;       // 4 per triangle (normals + flags) or null
    int sphere_count;
    const float4 *spheres;     // centre xyz, radius
    float floor_y, floor_hx, floor_zmin, floor_zmax;
    int pair_limit;            // packet walk: (ray, triangle) pairs queued before they are tested, 1..64
    const SkipNode *skip;      // the bottom level for the per-lane walk, same indices as the reference layout
    const int *mesh_skip_root; // per mesh: its root in skip
    const int *mesh_tris;      // per mesh: triangle count
    int lane_min_tris;         // packet walk: meshes of at least this many triangles are walked lane by lane
    int *overflow;             // set to 1 by a walk whose stack or trail would pass its capacity (a tree deeper than
                               // the walks allow); the walk then drops what does not fit, so its images are wrong
    int share_cap;             // entries of the LDS stack the two-child shared walk gives subtrees to, 1..128 (128;
                               // lower only to test the lanes' own stacks, which take the rest)
};

// The walks' depth limits: a wave-uniform stack of WALK_STACK entries per level for the packet walk, a 64-bit trail per
// level for the per-lane walk.
constexpr int WALK_STACK = 64;
// Entries of the LDS stack on which the shared per-lane walks give subtrees to idle lanes.
constexpr int SHARE_STACK = 128;

// The batched walks keep a triangle or node index and a lane number in one 32-bit word (index << 6 | lane), so
// indices stay below 2^25; a tree over n triangles has fewer than 2n nodes, so at most 2^24 triangles over all meshes.
constexpr long long MAX_TRIANGLES = 1ll << 24;

struct RenderArgs {
    TraceScene s;
    const float *cams;         // [views][12]
    float fov_scale;           // tanf(fov / 2), computed on the host as the reference's CPU backend does
    float aspect;              // width / height
    int width, height, cameras_per_world, views;
    int tile_w, tile_h, tiles_x, tiles_per_view;
    long long total_tiles;
    float *depth;              // [views][H][W] or null
    float *normal;             // [views][H][W][3] or null
    int32_t *object_id;        // [views][H][W] or null
    int xcd_remap;             // 1: consecutive tiles stay on one XCD
    int walk;                  // 0 per-lane, 1 packet, 2 packet with fine meshes lane by lane
    int *tile_counters;        // persistent mode (non-null): TILE_COUNTER_INTS ints, zeroed by the launch
    int blocks;                // the kernel variant: workgroups per compute unit it is built for (0: the walk's default)
};

// The RGB picture's shading inputs, on the device: materials as two float4 (diffuse rgb, diffuse albedo;
// specular albedo, reflectivity, specular exponent, 0), lights as four (position, intensity; radius, falloff,
// cos_outer, cos_inner; direction, 0; colour, 0).
constexpr int MAX_LIGHTS = 8;
// A path of the RGB picture ends where its throughput (the product of the reflectivities met) falls below this
// (render_core.h:16, :238-239).
constexpr float MIN_THROUGHPUT = 1e-3f;
struct ShadeScene {
    const float4 *instance_materials;  // [template instances][2]
    const float4 *sphere_materials;    // [spheres][2]
    float4 floor_material[2];
    float floor_checker_scale;
    int light_count;
    const float4 *lights;              // [light_count][4]
    float ambient;
    float3 sky_up, sky_color_up, sky_color_down;
    float sky_intensity;
};

struct RgbArgs {
    RenderArgs r;                      // the frame, cameras, walk and tiles (its AOV outputs unused)
    ShadeScene shade;
    uint8_t *rgb;                      // [views][H][W][3]
    int spp, frame;
    int blocks;                        // the RGB kernel's variant (0: the walk's default)
    // the two-pass picture's hits: [views][H][W] (t's bits, what was hit, instance, triangle), and their
    // barycentrics when a mesh has vertex normals (else null)
    int4 *hits;
    float2 *hit_uv;
};

struct RayArgs {
    TraceScene s;
    const float *origins;      // [worlds][R][3]
    const float *dirs;
    int rays_per_world;
    long long count;
    float max_distance;
    float *distance;
    float *normal;
    int32_t *object_id;
    int walk;
    const int *order;          // null, or per world the order to trace its rays in (indices within the world)
    int xcd_remap;             // 1: each XCD takes a contiguous range of the waves
    int blocks;                // the kernel variant (0: the default)
};

// The pose backward of depth: dL/d(body pose) for every world, through each pixel's fixed hit triangle.
struct DepthBackwardArgs {
    TraceScene s;
    const float *cams;
    float fov_scale, aspect;
    int width, height, cameras_per_world, views;
    const float *depth;        // the forward's depth: bounds the re-trace, and skips misses
    const float *grad_depth;   // [views][H][W]
    int bodies_per_world;
    const int *body_of_instance;     // [P]
    const float *local_rot_scale;    // [P][12]: Rl diag(s) | pl, the instance in its body's frame
    const float *body_poses;   // [worlds][bodies][7], the poses the tree was built from
    int parts;                 // workgroups per view, each taking every parts-th group of tiles
    float *partial;            // [views][parts][bodies][7]
    int walk;                  // the re-trace's walk, as the forward's
    int blocks;                // the kernel variant (0: the default)
};

struct RangeBackwardArgs {
    TraceScene s;
    const float *origins, *dirs;
    int rays_per_world;
    int chunks;                // per world
    float max_distance;
    const float *distance;     // the forward's distances
    const float *grad_distance;
    int bodies_per_world;
    const int *body_of_instance;
    const float *local_rot_scale;
    const float *body_poses;
    int worlds;
    float *partial;            // [worlds][chunks][bodies][7]
    int walk;
};

struct PoseGradFinishArgs {
    int worlds, bodies_per_world, parts;  // parts summed per world, in order
    const float *partial;      // [worlds][parts][bodies][7]
    const float *body_poses;   // [worlds][bodies][7]
    float *grad;               // [worlds][bodies][7]
};

// The variants each kernel is built in, by workgroups of 256 threads per compute unit (its register budget
// follows: 512 / (blocks x 4 / 4 SIMDs) registers a lane). 0 in a trace configuration picks the default.
constexpr int RENDER_BLOCKS[] = {5, 6, 7}, RENDER_BLOCKS_DEFAULT = 7;
constexpr int RGB_BLOCKS[] = {4, 5, 6, 7}, RGB_BLOCKS_DEFAULT_PACKET = 4, RGB_BLOCKS_DEFAULT_HYBRID = 7;
// The two-pass picture's shading pass by default: 4 workgroups per compute unit, the build without spills (the reference's
// falling-bodies scene, hybrid walk: 4.97 ms a frame for 32 bodies at 128x128 against 5.84 at 5), and 5 for the build
// for one light, which holds less across its shadow walk (the same scene: 4.75 ms at 5, 5.01 at 4, 4.86 at 6)
constexpr int SHADE_BLOCKS_DEFAULT = 4, SHADE_ONE_LIGHT_BLOCKS_DEFAULT = 5;
constexpr int RAY_BLOCKS[] = {5, 6, 7}, RAY_BLOCKS_DEFAULT = 6;
constexpr int BACKWARD_BLOCKS[] = {4, 5, 6}, BACKWARD_BLOCKS_DEFAULT = 5;

// The pose backward sums gradients per body in LDS: at most this many bodies per world.
constexpr int BACKWARD_MAX_BODIES = 256;

// The fixed device array of the allocation-free depth backward (ngx_pose_backward_depth_noalloc): its partial
// sums, views x parts x bodies x 7 floats, must fit. noalloc_scratch gives the current device's address of it,
// looked up once per device; every call on one device shares it.
constexpr size_t NOALLOC_SCRATCH_FLOATS = (size_t)8 << 20;
hipError_t noalloc_scratch(float **out);

// Builds every world of at most WORKGROUP_BUILD_MAX instances (max_count: the largest such world), one
// workgroup per world; larger worlds are skipped.
hipError_t launch_tlas_build(const TlasBuildArgs &a, int max_count, hipStream_t stream);

// Scratch of the device-wide build of one large world, sized for `capacity` instances.
struct BigBuildScratch {
    int capacity;
    uint32_t *key[2], *val[2]; // [capacity] each
    float *box;                // [6][capacity]: instance boxes by template index, lo xyz then hi xyz
    float *node_box;           // [6][capacity]: internal node boxes by local index
    int *left, *right;         // [capacity]
    int *parent;               // [2 * capacity]
    int *arrivals;             // [capacity]
    int *bounds;               // [6]: centre bounds as ordered ints
    int *tile_counts;          // [16 * ceil(capacity / 1024)]
};

// Per world, a stable order of its rays by direction cell, for coherent wavefronts: keys is scratch of
// 2 x rays_per_world x worlds words, vals receives the order (rays_per_world x worlds), and *order is vals.
constexpr int RAY_ORDER_MAX_RAYS = 262144;
hipError_t launch_ray_order(const float *dirs, int rays_per_world, int worlds, uint32_t *keys, int *vals, int **order,
                            hipStream_t stream);

// Builds world `world` (instances first .. first+n-1, n > 1) with grid-wide kernels.
hipError_t launch_tlas_build_big(const TlasBuildArgs &a, const BigBuildScratch &s, int world, int first, int n, hipStream_t stream);
// The RGB picture of every view, one 64-pixel tile per wavefront.
// A mesh's bottom level refitted to new vertices (ngx_refit.hip).
struct MeshRefitArgs {
    const float *vertices;    // the mesh's vertices, 3 floats each
    int vertex_base;          // global index of its first vertex (tri_verts holds global indices)
    const int *tri_verts;     // per packed triangle: its three vertices
    int tri_base, tri_count;  // the mesh's packed triangles
    float *tris;              // packed triangles, 9 floats each
    SkipNode *skip;           // per node: the box written here first
    const int *node_parent;   // per node, -1 at the root
    const int *node_wide;     // per node: its wide node, -1 for a leaf
    WideNode *wide;
    int node_base, node_count, root;
    int *arrivals;            // per node, zeroed by the launch
    MeshRoot *mesh_root;      // the mesh's
    float *mesh_bounds;       // the mesh's 6 floats (lo, hi): the top-level build's local box
};
hipError_t launch_mesh_refit(const MeshRefitArgs &a, hipStream_t stream);

// The RGB picture; with a.r.tile_counters set and compute_units > 0, from a persistent grid as launch_render.
hipError_t launch_rgb(const RgbArgs &a, hipStream_t stream, int compute_units = 0);
// One sample a pixel, no surface reflective enough to bounce, persistent launch, packet or hybrid walk: the hit pass,
// then the shading pass; the same bytes as launch_rgb.
hipError_t launch_rgb_two_pass(const RgbArgs &a, int hit_blocks, hipStream_t stream, int compute_units);

// Renders every view; with a.tile_counters set and compute_units > 0, a fixed grid of as many workgroups as the
// device holds at once takes tiles from per-XCD counters until none are left (the counters are zeroed by this
// call first).
hipError_t launch_render(const RenderArgs &a, hipStream_t stream, int compute_units = 0);
hipError_t launch_rays(const RayArgs &a, hipStream_t stream);
hipError_t launch_depth_backward(const DepthBackwardArgs &a, hipStream_t stream);
hipError_t launch_range_backward(const RangeBackwardArgs &a, hipStream_t stream);
hipError_t launch_pose_grad_finish(const PoseGradFinishArgs &a, hipStream_t stream);
#ifdef NGX_TRACE_STATS
constexpr int TRACE_STAT_COUNT = 11;
hipError_t launch_render_stats(const RenderArgs &a, int *out, hipStream_t stream);
#endif

// The largest world the one-workgroup-per-world build takes.
constexpr int WORKGROUP_BUILD_MAX = 1024;

}  // namespace ngx
