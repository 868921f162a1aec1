/*
 * ngx_backend.h -- the C ABI of the MI355X batched sensor-rendering backend.
 *
 * This is the drop-in boundary for the reference's batched-world path. The reference's plug-in point is
 * the C++ interface detail::Backend (src/backend.h:36-85), built from a SceneInput (src/backend.h:19-34)
 * and reached from BatchRenderer (src/batch.cpp:91-192). Each entry point below replaces one method of
 * that interface; the citations say which. Plain C: pointers and sizes only, no exceptions across it.
 *
 * Conventions (the reference's, with the differences named):
 *  - Errors: every call returns an ngx_status; NGX_INVALID_ARGUMENT is the reference's
 *    std::invalid_argument (bad input, checked before anything changes, so a rejected call changes
 *    nothing), NGX_DEVICE_ERROR its CudaError (include/neraulgraphics/renderer.h:30-39).
 *    ngx_last_error() gives the message of the calling thread's last failure.
 *  - Memory: NGX_MEM_HOST arrays are validated as the reference validates host arrays
 *    (src/batch.cpp:91-192); NGX_MEM_DEVICE arrays (HIP device pointers) are trusted, as there.
 *  - Streams: every launch goes on the hip stream given (NULL: the default stream) and nothing waits
 *    for it, so a caller can capture a step into a HIP graph. ngx_synchronize() restores the
 *    reference's "finished on return" contract (src/cuda_backend.cu:403-405, :542, :583).
 *  - Results are in buffers the context owns, valid until the next call that writes them
 *    (include/neraulgraphics/batch.h:122-126, renderer.h:85-87).
 *  - A context is not thread-safe; the reference's Python edge serializes with a mutex per object
 *    (python/bindings.cpp:242-251).
 *
 * Additions with no counterpart in the reference:
 *  - per-world instance counts (instances_in_world): world w holds the first count[w] instances of the
 *    template; the reference has the same list in every world (include/neraulgraphics/batch.h:7-9);
 *  - ngx_pose_backward_depth / ngx_pose_backward_range: gradients of depth and range with respect to
 *    body poses;
 *  - ngx_set_mesh_vertices: deforming meshes, refitted on the device;
 *  - ngx_bind_geoms / ngx_set_geom_frames: the MuJoCo Warp adapter's compose step (a Warp kernel and a
 *    separate instance update in the reference) fused into the top-level build;
 *  - ngx_debug_*: the top-level build's intermediates, for checking it against the reference algorithm.
 */
#ifndef NGX_BACKEND_H
#define NGX_BACKEND_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ngx_ctx ngx_ctx;
typedef int32_t ngx_status;

enum {
    NGX_OK = 0,
    NGX_INVALID_ARGUMENT = 1,
    NGX_DEVICE_ERROR = 2,
    NGX_OUT_OF_MEMORY = 3,
    NGX_BAD_STATE = 4,
    NGX_UNSUPPORTED = 5
};

/* Memory (include/neraulgraphics/renderer.h:49). */
typedef enum { NGX_MEM_HOST = 0, NGX_MEM_DEVICE = 1 } ngx_memory;

/* Output bits (include/neraulgraphics/renderer.h:53-58). */
enum { NGX_OUTPUT_RGB = 1, NGX_OUTPUT_DEPTH = 2, NGX_OUTPUT_NORMAL = 4, NGX_OUTPUT_OBJECT_ID = 8 };

/* Object IDs (include/neraulgraphics/scene.h:19-22). */
enum { NGX_OBJECT_NONE = 0, NGX_OBJECT_SPHERE = 1, NGX_OBJECT_FLOOR = 33, NGX_OBJECT_INSTANCE = 34 };
#define NGX_MAX_SPHERES 32

/* The scene: src/backend.h:19-34 (SceneInput) with the meshes given raw, as BatchRenderer's constructor
 * takes them (include/neraulgraphics/batch.h:164-167); the bottom-level trees are built here, once. */
typedef struct {
    int32_t width, height, worlds, cameras_per_world;

    int32_t mesh_count;
    const float *vertices;          /* every mesh's vertices, 3 floats each, mesh after mesh */
    const int32_t *vertex_offsets;  /* mesh_count + 1: mesh m's vertices are vertex_offsets[m] .. [m+1]-1 */
    const uint32_t *indices;        /* 3 per triangle, indexing the mesh's own vertices */
    const int32_t *index_offsets;   /* mesh_count + 1 */
    const float *normals;           /* per vertex like vertices, or NULL */
    const uint8_t *mesh_has_normals;/* mesh_count flags (smooth shading), or NULL */

    int32_t instances_per_world;    /* the template: what a full world holds */
    const int32_t *instance_mesh;   /* instances_per_world mesh indices */
    const float *instance_transforms; /* instances_per_world x 12: where every world starts */
    const int32_t *instances_in_world;/* worlds counts in [0, instances_per_world], or NULL for all */

    int32_t sphere_count;           /* <= NGX_MAX_SPHERES */
    const float *spheres;           /* sphere_count x 4: centre x, y, z, radius */
    float floor_y, floor_half_extent_x, floor_z_min, floor_z_max; /* half_extent_x 0: no floor */

    int32_t device_id;              /* HIP device */
} ngx_scene_desc;

/* include/neraulgraphics/batch.h:126-136 */
typedef struct {
    int32_t memory;     /* ngx_memory: always NGX_MEM_DEVICE */
    int32_t device_id;
    int32_t views, width, height;
    const uint8_t *rgb; /* views x H x W x 3, or NULL when not asked for */
    const float *depth;
    const float *normal;
    const int32_t *object_id;
} ngx_batch_frame;

/* include/neraulgraphics/renderer.h:88-95 */
typedef struct {
    int32_t memory;
    int32_t device_id;
    int32_t count;
    const float *distance;
    const float *normal;
    const int32_t *object_id;
} ngx_ray_hits;

/* ---- the RGB picture (the classic plane: direct light from sphere lights with soft shadows, mirror
 *      reflections, an ambient term and a sky gradient; src/render_core.h:146-283) ---- */
#define NGX_MAX_LIGHTS 8

/* The classic fields of a material (include/neraulgraphics/scene.h:35-40). */
typedef struct {
    float diffuse_color[3];
    float diffuse_albedo, specular_albedo, reflectivity, specular_exponent;
} ngx_material;

/* include/neraulgraphics/scene.h:82-92: a sphere light (radius 0: a point), its falloff distance (0: none),
 * and a spotlight cone along direction of half-angle spot_angle (0: no cone), fading over its outer
 * spot_softness. */
typedef struct {
    float position[3];
    float intensity, radius, falloff;
    float direction[3];
    float spot_angle, spot_softness;
    float color[3];
} ngx_light;

/* Everything the picture needs beyond the geometry: the materials of the template's instances (one per
 * template entry, the same in every world), of the spheres and of the floor (its checker tint at
 * floor_checker_scale), the lights, the ambient level, and the sky gradient (include/neraulgraphics/scene.h:109-121;
 * the reference's defaults are up (0, 1, 0), up colour (0.10, 0.26, 0.60), down colour (0.35, 0.52, 0.70),
 * intensity 1, ambient 0). */
typedef struct {
    const ngx_material *instance_materials;  /* instances_per_world */
    const ngx_material *sphere_materials;    /* sphere_count */
    ngx_material floor_material;
    float floor_checker_scale;
    int32_t light_count;                     /* <= NGX_MAX_LIGHTS */
    const ngx_light *lights;
    float ambient;
    float sky_up[3], sky_color_up[3], sky_color_down[3], sky_intensity;
} ngx_shading;

/* make_cuda_backend (src/backend.h:88) with BatchRenderer's constructor checks (src/batch.cpp:47-75). More than
 * NGX_MAX_TRIANGLES triangles over all meshes are refused with NGX_UNSUPPORTED. */
#define NGX_MAX_TRIANGLES 16777216
ngx_status ngx_create(const ngx_scene_desc *scene, ngx_ctx **out);
void ngx_destroy(ngx_ctx *ctx);

/* Backend::set_instance_matrices (src/backend.h:62); host checks of src/batch.cpp:91-109.
 * worlds x instances_per_world x 12 floats, world-major; entries beyond a world's count are ignored. */
ngx_status ngx_set_instance_matrices(ngx_ctx *ctx, const float *matrices, ngx_memory memory, void *stream);

/* Backend::bind_bodies (src/backend.h:63); checks of src/batch.cpp:111-129. Host arrays, per template
 * instance: body index (-1..bodies_per_world-1), local pose (x, y, z, qx, qy, qz, qw), scale (x, y, z). */
ngx_status ngx_bind_bodies(ngx_ctx *ctx, int32_t bodies_per_world, const int32_t *body_of_instance,
                           const float *local_pose7, const float *scale3);

/* Backend::set_body_poses (src/backend.h:65); checks of src/batch.cpp:131-144.
 * worlds x bodies_per_world x 7 floats. Composed in FP64 as src/instances.cpp:137-164 does. */
ngx_status ngx_set_body_poses(ngx_ctx *ctx, const float *poses7, ngx_memory memory, void *stream);

/* Simulator geoms, the MuJoCo Warp flow (python/neraulgraphics/mujoco.py:177-302): binds each template
 * instance to geom geom_of_instance[i] (0..geoms_per_world-1) of its world, drawn with scale3[i] along the
 * mesh's own axes. Host arrays; checked before anything changes. Replaces the adapter's geom list and scales
 * (mujoco.py:239, :257); independent of the body binding. */
ngx_status ngx_bind_geoms(ngx_ctx *ctx, int32_t geoms_per_world, const int32_t *geom_of_instance, const float *scale3);

/* The adapter's per-step update (mujoco.py:276-302: a Warp kernel composing R diag(s) | p, then
 * set_instance_transforms) as one build: each instance's matrix is composed in FP32 from its geom's world
 * position (xpos3: worlds x geoms_per_world x 3 floats, MjData.geom_xpos) and row-major rotation (xmat9:
 * worlds x geoms_per_world x 9 floats, MjData.geom_xmat), both in `memory`, then the top level is rebuilt. */
ngx_status ngx_set_geom_frames(ngx_ctx *ctx, const float *xpos3, const float *xmat9, ngx_memory memory, void *stream);

/* No counterpart in the reference (its bottom levels are built once): moves the vertices of mesh `mesh` (as many
 * as it was created with, 3 floats each) and refits its bottom level on the device, keeping the tree. The top
 * level holds the mesh's old box until the next ngx_set_instance_matrices or ngx_set_body_poses, and renders and
 * ray queries return NGX_BAD_STATE until then. Meshes with vertex normals are refused (NGX_UNSUPPORTED). */
ngx_status ngx_set_mesh_vertices(ngx_ctx *ctx, int32_t mesh, const float *vertices, ngx_memory memory, void *stream);

/* Materials, lights, ambient and sky for the RGB picture: the reference takes them with the scene
 * (src/backend.h:19-34, :74-80); here they are set after creation, before the first RGB render
 * (NGX_BAD_STATE without them). Host arrays; checked before anything changes. */
ngx_status ngx_set_shading(ngx_ctx *ctx, const ngx_shading *shading);

/* Backend::render_batch (src/backend.h:67-68); checks of src/batch.cpp:146-162. views x 12 camera-to-world
 * floats (OpenGL convention). NGX_OUTPUT_RGB draws the classic picture (spp jittered samples a pixel, seeded
 * by pixel and frame, src/render_core.h:251-263), 3 bytes a pixel; depth, normal and ID do not depend on spp
 * or frame. */
ngx_status ngx_render_batch(ngx_ctx *ctx, const float *cams12, ngx_memory memory, float fov, int32_t spp,
                            uint32_t outputs, int32_t frame, void *stream, ngx_batch_frame *out);

/* Backend::cast_rays_batch (src/backend.h:69-70); checks of src/batch.cpp:180-192.
 * origins and directions: worlds x rays_per_world x 3 floats, both in `memory`. */
ngx_status ngx_cast_rays_batch(ngx_ctx *ctx, const float *origins, const float *directions, int32_t rays_per_world,
                               float max_distance, ngx_memory memory, void *stream, ngx_ray_hits *out);

/* No counterpart in the reference: dL/d(body poses) of a depth render, through each pixel's fixed hit
 * triangle. Device arrays: body_poses (worlds x B x 7, the poses to differentiate at; the top-level trees
 * are rebuilt from them first), cams (views x 12), depth (views x H x W, the forward's output: it bounds
 * the re-trace and marks misses), grad_depth (dL/d depth), grad_out (worlds x B x 7, overwritten).
 * The gradient is summed per body in LDS: a binding of more than NGX_BACKWARD_MAX_BODIES bodies per world is
 * refused with NGX_UNSUPPORTED, before anything changes. */
#define NGX_BACKWARD_MAX_BODIES 256
ngx_status ngx_pose_backward_depth(ngx_ctx *ctx, const float *body_poses, const float *cams12, float fov,
                                   const float *depth, const float *grad_depth, float *grad_out, void *stream);

/* No counterpart in the reference. For callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc): ngx_pose_backward_depth, same arguments and results, with
 * its per-workgroup partial sums in a fixed device array of NGX_NOALLOC_SCRATCH_FLOATS floats (32 MiB) that each GPU
 * holds from ngx_create until the process ends, instead of memory the context allocates; it calls no allocator
 * and does not wait on the stream. It needs views x parts x bodies_per_world x 7 floats, parts = min(16,
 * max(1, tiles / 8)) for a frame of `tiles` 8x8 tiles; a larger call is refused with NGX_INVALID_ARGUMENT
 * before anything changes. Every call on one GPU shares that GPU's array (it keeps the last call's contents;
 * each call overwrites what it reads): calls that overlap on one GPU, from other streams, threads or
 * contexts, race and give wrong gradients. */
#define NGX_NOALLOC_SCRATCH_FLOATS (8 * 1024 * 1024)
ngx_status ngx_pose_backward_depth_noalloc(ngx_ctx *ctx, const float *body_poses, const float *cams12, float fov,
                                           const float *depth, const float *grad_depth, float *grad_out, void *stream);

/* No counterpart in the reference: the same for a ray query's distances. Device arrays; the same body limit. */
ngx_status ngx_pose_backward_range(ngx_ctx *ctx, const float *body_poses, const float *origins, const float *directions,
                                   int32_t rays_per_world, float max_distance, const float *distance,
                                   const float *grad_distance, float *grad_out, void *stream);

/* Backend::copy_to_host (src/backend.h:54). Synchronous. */
ngx_status ngx_copy_to_host(ngx_ctx *ctx, const void *src, void *dst, size_t bytes);

/* Waits for everything queued on the context's last stream. Returns NGX_DEVICE_ERROR when a tree walk since the
 * last call ran out of stack (a tree deeper than the walks hold: 63 levels below a world's or a mesh's root; the
 * builds keep within that, so this marks a corrupted tree), in which case those results may be wrong. */
ngx_status ngx_synchronize(ngx_ctx *ctx);

/* Backend::last_kernel_ms (src/backend.h:83): device time of the last render or ray query's kernel (or, after a
 * pose backward, of its backward kernel), in ms, measured with events; read after ngx_synchronize. 0 after a call
 * made while its stream was being captured into a graph. */
float ngx_last_kernel_ms(ngx_ctx *ctx);

/* BatchRenderer's accessors (include/neraulgraphics/batch.h:214-221). */
int32_t ngx_world_count(const ngx_ctx *ctx);
int32_t ngx_view_count(const ngx_ctx *ctx);
int32_t ngx_instances_per_world(const ngx_ctx *ctx);
int32_t ngx_instance_count(const ngx_ctx *ctx);   /* over all worlds */
int32_t ngx_device_id(const ngx_ctx *ctx);
ngx_status ngx_frame_size(const ngx_ctx *ctx, int32_t *width, int32_t *height);

/* Message of the calling thread's last failed call ("" if none). */
const char *ngx_last_error(void);

/* Tuning of the trace kernels (no counterpart in the reference, whose tile is fixed at 8x8 threads):
 *  tile_w      a wavefront's pixels are a tile_w x (64 / tile_w) tile: 8, 16, 32 or 64;
 *  xcd_remap   1: consecutive workgroups, which the dispatcher deals round-robin over the eight XCDs, are
 *              renumbered so each XCD works through one contiguous range of tiles (one L2 per XCD);
 *  packet      1: the wavefront walks its 64 rays as a packet (the default); 0: each lane walks its own ray;
 *  pair_limit  packet walk: (ray, triangle) pairs queued before they are tested together, 1..64;
 *  persistent  1 (the default): a fixed grid of 8 workgroups per compute unit takes tiles from per-XCD
 *              counters; 0: one wavefront per tile, dealt by the dispatcher;
 *  lane_min_triangles  packet walk: the bottom level of meshes with at least this many triangles is walked
 *              lane by lane (each ray its own path) instead of as a packet (default 300);
 *  sort_rays   ray queries of 1024 to 262,144 rays a world: 1 (the default) traces each world's rays in an order
 *              of their directions, so a wavefront's rays are neighbours; results are in the caller's order;
 *  rgb_two_pass  1 (the default): the RGB picture at one sample a pixel, where no surface is reflective enough to
 *              bounce a ray (every reflectivity below 1e-3), with the persistent launch and the packet walk, is drawn
 *              in two passes: the camera rays' hits (16 bytes a pixel kept on the device, 24 with vertex normals),
 *              then their shading and shadow rays; the same bytes as the one pass (0), which is also taken where the
 *              device cannot hold the records. */
typedef struct {
    int32_t tile_w, xcd_remap, packet, pair_limit, persistent, lane_min_triangles, sort_rays;
    /* Kernel variants, by the workgroups of 256 threads per compute unit each is built for (the registers a lane
     * may hold follow from it; 0 picks the walk's default): the depth / normal / ID kernel 5, 6 or 7 (default 7),
     * the RGB kernel 4, 5, 6 or 7 (default 4 for the packet walk, 7 for the hybrid one; the two-pass picture's hit pass
     * takes render_blocks and its shading pass rgb_blocks, default 5 for a scene of one light and 4 otherwise), the
     * ray-query kernel 5, 6 or 7 (default 6), the depth backward 4, 5 or 6 (default 5). The per-lane walk has one
     * variant of each. */
    int32_t render_blocks, rgb_blocks, ray_blocks, backward_blocks;
    int32_t rgb_two_pass;
} ngx_trace_config;
ngx_status ngx_set_trace_config(ngx_ctx *ctx, const ngx_trace_config *config);
ngx_status ngx_get_trace_config(const ngx_ctx *ctx, ngx_trace_config *config);

/* Tries every trace configuration worth trying on this scene and these cameras (each rendered once to warm
 * up, then `reps` times timed with events on `stream`; the four fastest walk and tile choices again over 4 x reps
 * renders, then the kernel variants of the winner over 2 x reps), keeps the fastest, and returns it with its time
 * per render. Synchronous; not while `stream` is being captured. The outputs are left as the last render wrote
 * them. */
ngx_status ngx_autotune(ngx_ctx *ctx, const float *cams12, ngx_memory memory, float fov, uint32_t outputs, int32_t reps,
                        void *stream, ngx_trace_config *best, float *best_ms);

/* No counterpart in the reference: the same for ray queries (direction order on or off, the ray kernel's
 * variant) on the caller's rays, and for the depth backward (its kernel's variant) on the caller's arguments of
 * ngx_pose_backward_depth; the rest of the configuration is kept. */
ngx_status ngx_autotune_rays(ngx_ctx *ctx, const float *origins, const float *directions, int32_t rays_per_world,
                             float max_distance, ngx_memory memory, int32_t reps, void *stream, ngx_trace_config *best,
                             float *best_ms);
ngx_status ngx_autotune_backward(ngx_ctx *ctx, const float *body_poses, const float *cams12, float fov, const float *depth,
                                 const float *grad_depth, float *grad_out, int32_t reps, void *stream, ngx_trace_config *best,
                                 float *best_ms);

/* ---- build intermediates, for checking the top-level build against the reference algorithm ---- */

/* Keep (1) or drop (0) the intermediates on later builds: Morton keys, sort order, boxes. */
ngx_status ngx_debug_keep_build(ngx_ctx *ctx, int32_t keep);

/* Copies the last build to host arrays (NULL skips one):
 *  nodes (2N-1) x 8 x 4 bytes (TlasNode: lo xyz, left, hi xyz, right), roots (worlds),
 *  inv (N x 12 world-to-mesh rows), matrices (N x 12), and, when kept: keys (N, world << 32 | morton, by
 *  instance), sorted_keys (N), sorted_ids (N), aabb (N x 6), world_bounds (worlds x 6). */
ngx_status ngx_debug_read_build(ngx_ctx *ctx, void *nodes, int32_t *roots, float *inv, float *matrices, uint64_t *keys,
                                uint64_t *sorted_keys, int32_t *sorted_ids, float *aabb, float *world_bounds);

/* The per-lane bottom-level layout of one mesh, no device needed: skip_nodes (capacity 2 x triangles) x 64 bytes
 * (SkipNode: lo xyz, a, hi xyz, meta, escape[8]) in the same node order as ngx_build_blas_host's nodes. */
ngx_status ngx_build_skip_host(const float *vertices, int32_t vertex_count, const uint32_t *indices, int32_t triangle_count,
                               void *skip_nodes, int32_t *node_count);

/* Checks, on device device_id, the reciprocal the tracer uses for 1/x (the hardware estimate and one
 * fused Newton step) against IEEE division for all 2^32 float bit patterns: mismatches among
 * 2^-126 <= |x| < 2^126 (where the tracer uses it), among all finite nonzero x, and one mismatching
 * pattern of the first kind (0 if none). About a second of device time. */
ngx_status ngx_debug_check_reciprocal(int32_t device_id, uint64_t *mismatches_in_range, uint64_t *mismatches_anywhere,
                                      uint32_t *example);

/* No counterpart in the reference: the inclusive prefix sum over a wavefront that the batched walks use to place (ray,
 * triangle) pairs, run on device device_id over host arrays of 64 x waves values (each run of 64 scanned by one
 * wavefront), for checking its lane layout. At most 2^20 waves. */
ngx_status ngx_debug_wave_prefix_sum(int32_t device_id, const int32_t *values, int32_t waves, int32_t *sums);

/* No counterpart in the reference: the per-world order in which a ray query with sort_rays traces its rays, run on
 * device device_id over host arrays: directions (worlds x rays_per_world x 3) in, order (worlds x rays_per_world, each
 * world's ray indices) out. rays_per_world at most 262,144, at most 2^26 rays in all; for checking that the order is the
 * stable order of the rays' direction cells. */
ngx_status ngx_debug_ray_order(int32_t device_id, const float *directions, int32_t rays_per_world, int32_t worlds, int32_t *order);

/* No counterpart in the reference: how the last RGB picture of this context was drawn, 1 or 2 passes (see rgb_two_pass;
 * 0 before any), for checking which kernels ran. */
ngx_status ngx_debug_rgb_passes(const ngx_ctx *ctx, int32_t *passes);

/* No counterpart in the reference: how many entries (1 to 128, the default) of the LDS stack on which the shared
 * per-lane bottom-level walk gives far subtrees to idle lanes this context's later calls use; the lanes keep what does
 * not fit on their own stacks. Results do not depend on it (a ray's ties aside); a test lowers it to run that path. */
ngx_status ngx_debug_set_share_capacity(ngx_ctx *ctx, int32_t entries);

/* The host-side bottom-level build of one mesh (src/mesh_build.cpp:91-157 with src/bvh.cpp:46-174), no device
 * needed: nodes (capacity 2 x triangles) x 32 bytes in the reference's Node layout, triangles
 * (triangles x 12 floats, v0 e1 e2 padded), order (triangles). Returns the node count in *node_count. */
ngx_status ngx_build_blas_host(const float *vertices, int32_t vertex_count, const uint32_t *indices, int32_t triangle_count,
                               void *nodes, int32_t *node_count, float *triangles, uint32_t *order);

#ifdef __cplusplus
}
#endif

#endif /* NGX_BACKEND_H */
