/*
 * ngx_oracle.c -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
 *
 * A plain-C CPU restatement of the reference's batched sensor-rendering path, used only by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker. Nothing under
 * neraulgraphics_amd/ links, loads or calls this file.
 *
 * Compiled with -O2 -ffp-contract=off and no fast-math, so every float operation is one IEEE single
 * operation in source order, exactly as the reference's CPU backend (built without -march, no FMA on
 * x86-64) evaluates it.
 *
 * What each function restates (reference paths, file:line), and what pins it:
 *   orc_sah_build              src/bvh.cpp:46-174 (binned SAH, 16 bins, libstdc++ std::partition)
 *                              pinned: bit-exact vs the reference bvh.cpp compiled into oracle/_ref.
 *   orc_sah_refit, orc_sah_cost src/bvh.cpp:176-203
 *                              pinned: refit of unchanged boxes == build boxes; cost closed form on tiny trees.
 *   orc_build_meshes           src/mesh_build.cpp:91-157 (BLAS per mesh, triangles v0,e1,e2 in leaf order)
 *                              pinned: node arrays via _ref bvh.cpp; packing by reconstructing every triangle.
 *   orc_body_poses_to_matrices src/instances.cpp:126-164 (FP64, rounded once at the end)
 *                              pinned: the hand-composed matrices of tests/test_batch.cpp:233-258 and a
 *                              float64 library rotation (scipy) to within 1 float32 ulp.
 *   orc_lbvh_build             src/cuda_tlas.cuh:27-211 (place, Morton keys, stable radix sort, Karras,
 *                              leaves, bottom-up boxes, roots), generalised to per-world instance counts.
 *                              pinned: spread_bits == loop interleave; sort == numpy stable argsort;
 *                              topology invariants; brute-force split search on tiny inputs.
 *   orc_sah_tlas_*             src/instances.cpp:31-121 (per-world SAH TLAS, refit or rebuild, flatten)
 *                              pinned: forest node for node vs the reference's bvh.cpp build/refit/cost
 *                              (oracle/_ref) under the 1.3x rule on walks that straddle it; moving equals
 *                              building (tests/test_renderer.cpp:481-550).
 *   orc_render_batch           src/render_core.h:78-143, :264-273, :313-327; src/mesh_dev.h:58-205, :238-262;
 *                              src/device_scene.h:114-119; src/cpu_backend.cpp:94-120, :239-272
 *   orc_cast_rays_batch        src/render_core.h:331-349; src/cpu_backend.cpp:274-287
 *                              pinned (both): bit-exact per-ray hits vs the reference traversal core
 *                              (mesh_dev.h) compiled into oracle/_ref; known answers of tests/test_batch.cpp
 *                              and tests/test_renderer.cpp:739-775 (sphere, wall, floor depths and normals);
 *                              normals (mesh_surface, not linkable here) by the inverse-transpose face normals
 *                              of sheared boxes and the smooth normals of a unit icosphere.
 *   orc_render_rgb             src/render_core.h:146-283; src/mesh_dev.h:208-227; src/geom_dev.h:114-184;
 *                              src/device_scene.h:98-111
 *                              pinned: closed forms (sky, Lambert, shadow, mirror, Phong highlight, spotlight
 *                              cone, distance falloff) with each pixel's jitter from an independent PCG stream.
 *   orc_pose_grad_*            no reference (the backward pass does not exist there): FP64 closed form,
 *                              pinned by FP64 central finite differences of the ray-plane distance.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_BINS 16
#define ORC_MAX_DEPTH 60
#define ORC_TLAS_STACK 128
#define ORC_BLAS_STACK (ORC_MAX_DEPTH + 4)
#define ORC_NO_HIT 1e30f

#define OBJECT_NONE 0
#define OBJECT_SPHERE 1
#define OBJECT_FLOOR 33
#define OBJECT_INSTANCE 34

typedef struct { float lo[3]; int32_t a; float hi[3]; int32_t count; } OrcNode;        /* bvh.h:30-35 */
typedef struct { float lo[3]; int32_t left; float hi[3]; int32_t right; } OrcTlasNode; /* bvh.h:44-49 */
typedef struct { float lo[3]; float hi[3]; } OrcAabb;

/* ------------------------------------------------------------------ small helpers -- */

static float f_min(float a, float b) { return (b < a) ? b : a; } /* std::min */
static float f_max(float a, float b) { return (a < b) ? b : a; } /* std::max */

static OrcAabb empty_aabb(void) {
    OrcAabb b = {{INFINITY, INFINITY, INFINITY}, {-INFINITY, -INFINITY, -INFINITY}};
    return b;
}
static void grow(OrcAabb *box, const OrcAabb *o) { /* bvh.cpp:14-19 */
    for (int i = 0; i < 3; ++i) {
        box->lo[i] = f_min(box->lo[i], o->lo[i]);
        box->hi[i] = f_max(box->hi[i], o->hi[i]);
    }
}
static float surface_area(const OrcAabb *b) { /* bvh.cpp:21-27 */
    const float dx = b->hi[0] - b->lo[0], dy = b->hi[1] - b->lo[1], dz = b->hi[2] - b->lo[2];
    if (dx < 0.0f || dy < 0.0f || dz < 0.0f) return 0.0f;
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}

/* ------------------------------------------------------------ SAH build (bvh.cpp) -- */

typedef struct { int node; uint32_t begin, end; int depth; } Task;

static int bin_of(float c, float lo, float scale) {
    int b = (int)((c - lo) * scale);
    return b < ORC_BINS - 1 ? b : ORC_BINS - 1; /* std::min(BINS - 1, ...) */
}

/* libstdc++'s std::partition for bidirectional iterators (bits/stl_algo.h __partition): the order it
 * leaves is part of the tree the reference builds, so it is followed swap for swap. */
static uint32_t *partition_bins(uint32_t *first, uint32_t *last, const float (*centroid)[3], int axis,
                                float lo, float scale, int best_bin) {
#define PRED(p) (bin_of(centroid[(p)][axis], lo, scale) <= best_bin)
    for (;;) {
        for (;;) {
            if (first == last) return first;
            if (PRED(*first)) ++first; else break;
        }
        --last;
        for (;;) {
            if (first == last) return first;
            if (!PRED(*last)) --last; else break;
        }
        uint32_t t = *first; *first = *last; *last = t;
        ++first;
    }
#undef PRED
}

/* bounds: n boxes as 6 floats (lo xyz, hi xyz). nodes_out: capacity 2n (at least 1). order_out: n.
 * Returns the node count. */
int orc_sah_build(const float *bounds6, int n, int max_leaf_size, OrcNode *nodes, uint32_t *order) {
    if (n <= 0) return 0;
    const OrcAabb *bounds = (const OrcAabb *)bounds6;
    for (int i = 0; i < n; ++i) order[i] = (uint32_t)i;
    float (*centroid)[3] = (float (*)[3])malloc(sizeof(float) * 3 * (size_t)n);
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) centroid[i][k] = 0.5f * (bounds[i].lo[k] + bounds[i].hi[k]);
    int node_count = 1;
    memset(&nodes[0], 0, sizeof(OrcNode));
    Task *tasks = (Task *)malloc(sizeof(Task) * (size_t)(2 * n + 2));
    int ntask = 0;
    tasks[ntask++] = (Task){0, 0, (uint32_t)n, 0};
    while (ntask > 0) {
        const Task task = tasks[--ntask];
        const uint32_t count = task.end - task.begin;
        uint32_t *prims = order + task.begin;
        OrcAabb box = empty_aabb(), cbox = empty_aabb();
        for (uint32_t i = 0; i < count; ++i) {
            grow(&box, &bounds[prims[i]]);
            for (int k = 0; k < 3; ++k) {
                cbox.lo[k] = f_min(cbox.lo[k], centroid[prims[i]][k]);
                cbox.hi[k] = f_max(cbox.hi[k], centroid[prims[i]][k]);
            }
        }
        OrcNode node;
        memset(&node, 0, sizeof(node));
        for (int k = 0; k < 3; ++k) { node.lo[k] = box.lo[k]; node.hi[k] = box.hi[k]; }

        int best_axis = -1, best_bin = 0;
        float best_cost = INFINITY;
        if (count > 1) {
            for (int axis = 0; axis < 3; ++axis) {
                const float extent = cbox.hi[axis] - cbox.lo[axis];
                if (!(extent > 0.0f)) continue;
                const float scale = (float)ORC_BINS / extent;
                int bin_count[ORC_BINS] = {0};
                OrcAabb bin_box[ORC_BINS];
                for (int b = 0; b < ORC_BINS; ++b) bin_box[b] = empty_aabb();
                for (uint32_t i = 0; i < count; ++i) {
                    const int b = bin_of(centroid[prims[i]][axis], cbox.lo[axis], scale);
                    ++bin_count[b];
                    grow(&bin_box[b], &bounds[prims[i]]);
                }
                float right_area[ORC_BINS - 1];
                int right_count[ORC_BINS - 1];
                OrcAabb acc = empty_aabb();
                int running = 0;
                for (int i = ORC_BINS - 1; i >= 1; --i) {
                    grow(&acc, &bin_box[i]);
                    running += bin_count[i];
                    right_area[i - 1] = surface_area(&acc);
                    right_count[i - 1] = running;
                }
                acc = empty_aabb();
                running = 0;
                for (int i = 0; i < ORC_BINS - 1; ++i) {
                    grow(&acc, &bin_box[i]);
                    running += bin_count[i];
                    if (running == 0 || right_count[i] == 0) continue;
                    const float cost = (float)running * surface_area(&acc) + (float)right_count[i] * right_area[i];
                    if (cost < best_cost) { best_cost = cost; best_axis = axis; best_bin = i; }
                }
            }
        }
        const float leaf_cost = (float)count * surface_area(&box);
        const float split_cost = 1.0f * surface_area(&box) + best_cost;
        const int small_leaf = count <= (uint32_t)max_leaf_size;
        if (count == 1 || task.depth >= ORC_MAX_DEPTH || (small_leaf && (best_axis < 0 || leaf_cost <= split_cost))) {
            node.a = (int32_t)task.begin;
            node.count = (int32_t)count;
            nodes[task.node] = node;
            continue;
        }
        uint32_t mid;
        if (best_axis >= 0) {
            const float extent = cbox.hi[best_axis] - cbox.lo[best_axis];
            const float scale = (float)ORC_BINS / extent;
            uint32_t *p = partition_bins(prims, prims + count, (const float (*)[3])centroid, best_axis,
                                         cbox.lo[best_axis], scale, best_bin);
            mid = task.begin + (uint32_t)(p - prims);
            if (mid == task.begin || mid == task.end) mid = task.begin + count / 2;
        } else {
            mid = task.begin + count / 2;
        }
        const int left = node_count;
        node_count += 2;
        node.a = left;
        node.count = 0;
        nodes[task.node] = node;
        tasks[ntask++] = (Task){left + 1, mid, task.end, task.depth + 1};
        tasks[ntask++] = (Task){left, task.begin, mid, task.depth + 1};
    }
    free(tasks);
    free(centroid);
    return node_count;
}

/* bvh.cpp:176-191 */
void orc_sah_refit(OrcNode *nodes, int node_count, const uint32_t *order, const float *bounds6) {
    const OrcAabb *bounds = (const OrcAabb *)bounds6;
    for (int i = node_count; i-- > 0;) {
        OrcNode *node = &nodes[i];
        OrcAabb box = empty_aabb();
        if (node->count > 0) {
            for (int k = 0; k < node->count; ++k) grow(&box, &bounds[order[node->a + k]]);
        } else {
            for (int c = 0; c < 2; ++c) {
                const OrcNode *ch = &nodes[node->a + c];
                OrcAabb cb = {{ch->lo[0], ch->lo[1], ch->lo[2]}, {ch->hi[0], ch->hi[1], ch->hi[2]}};
                grow(&box, &cb);
            }
        }
        for (int k = 0; k < 3; ++k) { node->lo[k] = box.lo[k]; node->hi[k] = box.hi[k]; }
    }
}

/* bvh.cpp:193-203 */
float orc_sah_cost(const OrcNode *nodes, int node_count) {
    if (node_count <= 0) return 0.0f;
    double cost = 0.0;
    for (int i = 0; i < node_count; ++i) {
        OrcAabb b = {{nodes[i].lo[0], nodes[i].lo[1], nodes[i].lo[2]}, {nodes[i].hi[0], nodes[i].hi[1], nodes[i].hi[2]}};
        const float area = surface_area(&b);
        cost += nodes[i].count > 0 ? (double)area * nodes[i].count : (double)area * 1.0f;
    }
    OrcAabb r = {{nodes[0].lo[0], nodes[0].lo[1], nodes[0].lo[2]}, {nodes[0].hi[0], nodes[0].hi[1], nodes[0].hi[2]}};
    const float root_area = surface_area(&r);
    return root_area > 0.0f ? (float)(cost / root_area) : 0.0f;
}

/* ----------------------------------------------------------- meshes (mesh_build.cpp) -- */

/* Sizes of orc_build_meshes' outputs: nodes <= 2 * total triangles; triangles = 12 floats each. */
/* vertices: all meshes' vertices, 3 floats each; vert_off[m] .. vert_off[m+1];
 * indices: all meshes' indices (local to the mesh's vertices); idx_off[m] .. idx_off[m+1].
 * normals: per vertex like vertices, or NULL; has_normals[m] says which meshes have them.
 * Outputs: nodes (global indices), triangles (12 floats each), attributes (16 floats each, or not
 * written when attributes == NULL), roots[m], bounds[m][6]. Returns the node count. */
int orc_build_meshes(int mesh_count, const float *vertices, const int32_t *vert_off, const uint32_t *indices,
                     const int32_t *idx_off, const float *normals, const int32_t *has_normals,
                     OrcNode *nodes_out, float *tris_out, float *attr_out, int32_t *roots, float *bounds_out) {
    int node_total = 0, tri_total = 0;
    for (int m = 0; m < mesh_count; ++m) {
        const float *v = vertices + 3 * (size_t)vert_off[m];
        const uint32_t *ix = indices + idx_off[m];
        const int tri_count = (idx_off[m + 1] - idx_off[m]) / 3;
        float *tb = (float *)malloc(sizeof(float) * 6 * (size_t)tri_count);
        for (int t = 0; t < tri_count; ++t) {
            float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
            for (int k = 0; k < 3; ++k) {
                const float *p = v + 3 * (size_t)ix[3 * t + k];
                for (int a = 0; a < 3; ++a) { lo[a] = fminf(lo[a], p[a]); hi[a] = fmaxf(hi[a], p[a]); }
            }
            for (int a = 0; a < 3; ++a) { tb[6 * t + a] = lo[a]; tb[6 * t + 3 + a] = hi[a]; }
        }
        OrcNode *tree = (OrcNode *)malloc(sizeof(OrcNode) * (size_t)(2 * tri_count + 1));
        uint32_t *order = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)tri_count);
        const int nn = orc_sah_build(tb, tri_count, 4, tree, order);
        const int node_base = node_total, tri_base = tri_total;
        roots[m] = node_base;
        for (int a = 0; a < 3; ++a) { bounds_out[6 * m + a] = tree[0].lo[a]; bounds_out[6 * m + 3 + a] = tree[0].hi[a]; }
        for (int i = 0; i < nn; ++i) {
            OrcNode n = tree[i];
            n.a += n.count > 0 ? tri_base : node_base;
            nodes_out[node_total++] = n;
        }
        for (int k = 0; k < tri_count; ++k) {
            const uint32_t t = order[k];
            const float *a = v + 3 * (size_t)ix[3 * t], *b = v + 3 * (size_t)ix[3 * t + 1], *c = v + 3 * (size_t)ix[3 * t + 2];
            float *o = tris_out + 12 * (size_t)tri_total;
            o[0] = a[0]; o[1] = a[1]; o[2] = a[2]; o[3] = 0.0f;
            o[4] = b[0] - a[0]; o[5] = b[1] - a[1]; o[6] = b[2] - a[2]; o[7] = 0.0f;
            o[8] = c[0] - a[0]; o[9] = c[1] - a[1]; o[10] = c[2] - a[2]; o[11] = 0.0f;
            if (attr_out) {
                float *r = attr_out + 16 * (size_t)tri_total;
                memset(r, 0, sizeof(float) * 16);
                const int32_t flags = (has_normals && has_normals[m]) ? 1 : 0;
                if (flags) {
                    const float *nrm = normals + 3 * (size_t)vert_off[m];
                    for (int q = 0; q < 3; ++q)
                        for (int a2 = 0; a2 < 3; ++a2) r[4 * q + a2] = nrm[3 * (size_t)ix[3 * t + q] + a2];
                }
                memcpy(&r[15], &flags, sizeof(flags));
            }
            ++tri_total;
        }
        free(order);
        free(tree);
        free(tb);
    }
    return node_total;
}

/* ------------------------------------------------------------------ transforms -- */

/* mesh_build.cpp:159-183 and cuda_tlas.cuh:35-47: the same arithmetic. */
void orc_invert_affine(const float *m, float *out) {
    const float c00 = m[5] * m[10] - m[6] * m[9];
    const float c01 = m[6] * m[8] - m[4] * m[10];
    const float c02 = m[4] * m[9] - m[5] * m[8];
    const float det = m[0] * c00 + m[1] * c01 + m[2] * c02;
    const float inv = 1.0f / det;
    const float r[9] = {c00 * inv, (m[2] * m[9] - m[1] * m[10]) * inv, (m[1] * m[6] - m[2] * m[5]) * inv,
                        c01 * inv, (m[0] * m[10] - m[2] * m[8]) * inv, (m[2] * m[4] - m[0] * m[6]) * inv,
                        c02 * inv, (m[1] * m[8] - m[0] * m[9]) * inv, (m[0] * m[5] - m[1] * m[4]) * inv};
    for (int row = 0; row < 3; ++row) {
        for (int col = 0; col < 3; ++col) out[4 * row + col] = r[3 * row + col];
        out[4 * row + 3] = -(r[3 * row] * m[3] + r[3 * row + 1] * m[7] + r[3 * row + 2] * m[11]);
    }
}

/* mesh_build.cpp:159-183's acceptance test: finite, and |det| > 1e-12 * mag^3. */
int orc_transform_ok(const float *m) {
    for (int i = 0; i < 12; ++i) if (!isfinite(m[i])) return 0;
    const float c00 = m[5] * m[10] - m[6] * m[9];
    const float c01 = m[6] * m[8] - m[4] * m[10];
    const float c02 = m[4] * m[9] - m[5] * m[8];
    const float det = m[0] * c00 + m[1] * c01 + m[2] * c02;
    float mag = 0.0f;
    const int idx[9] = {0, 1, 2, 4, 5, 6, 8, 9, 10};
    for (int i = 0; i < 9; ++i) mag = fmaxf(mag, fabsf(m[idx[i]]));
    if (!isfinite(det) || fabsf(det) <= 1e-12f * mag * mag * mag) return 0;
    return 1;
}

/* The 8-corner world box, cuda_tlas.cuh:78-87 (starts at +-1e30) == mesh_build.cpp:185-197. */
static void transform_box(const float *m, const float *lb, float *lo, float *hi) {
    lo[0] = lo[1] = lo[2] = 1e30f;
    hi[0] = hi[1] = hi[2] = -1e30f;
    for (int corner = 0; corner < 8; ++corner) {
        const float p[3] = {(corner & 1) ? lb[3] : lb[0], (corner & 2) ? lb[4] : lb[1], (corner & 4) ? lb[5] : lb[2]};
        for (int a = 0; a < 3; ++a) {
            const float w = m[4 * a] * p[0] + m[4 * a + 1] * p[1] + m[4 * a + 2] * p[2] + m[4 * a + 3];
            lo[a] = fminf(lo[a], w);
            hi[a] = fmaxf(hi[a], w);
        }
    }
}

/* instances.cpp:126-133 */
static void quat_to_matrix_d(const float *q, double r[9]) {
    double x = q[0], y = q[1], z = q[2], w = q[3];
    const double n = sqrt(x * x + y * y + z * z + w * w);
    if (n > 0.0) { x /= n; y /= n; z /= n; w /= n; } else { x = y = z = 0.0; w = 1.0; }
    r[0] = 1 - 2 * (y * y + z * z); r[1] = 2 * (x * y - z * w);     r[2] = 2 * (x * z + y * w);
    r[3] = 2 * (x * y + z * w);     r[4] = 1 - 2 * (x * x + z * z); r[5] = 2 * (y * z - x * w);
    r[6] = 2 * (x * z - y * w);     r[7] = 2 * (y * z + x * w);     r[8] = 1 - 2 * (x * x + y * y);
}

/* instances.cpp:137-164, with per-world counts: world w holds template instances 0..counts[w]-1 at
 * global positions world_off[w] + i. out: 12 floats per global instance. */
void orc_body_poses_to_matrices(int worlds, const int32_t *world_off, int bodies_per_world,
                                const int32_t *body_of_instance, const float *local_pose, const float *scale,
                                const float *body_poses, float *out) {
    for (int w = 0; w < worlds; ++w) {
        const int count = world_off[w + 1] - world_off[w];
        for (int i = 0; i < count; ++i) {
            const float *local = &local_pose[7 * i];
            const float *sc = &scale[3 * i];
            double rl[9], r[9], p[3] = {local[0], local[1], local[2]};
            quat_to_matrix_d(local + 3, rl);
            if (body_of_instance[i] < 0) {
                for (int k = 0; k < 9; ++k) r[k] = rl[k];
            } else {
                const float *body = &body_poses[7 * ((size_t)w * bodies_per_world + body_of_instance[i])];
                double rb[9];
                quat_to_matrix_d(body + 3, rb);
                for (int row = 0; row < 3; ++row) {
                    for (int col = 0; col < 3; ++col)
                        r[3 * row + col] = rb[3 * row] * rl[col] + rb[3 * row + 1] * rl[3 + col] + rb[3 * row + 2] * rl[6 + col];
                    p[row] = body[row] + rb[3 * row] * local[0] + rb[3 * row + 1] * local[1] + rb[3 * row + 2] * local[2];
                }
            }
            float *m = out + 12 * ((size_t)world_off[w] + i);
            for (int row = 0; row < 3; ++row) {
                for (int col = 0; col < 3; ++col) m[4 * row + col] = (float)(r[3 * row + col] * sc[col]);
                m[4 * row + 3] = (float)p[row];
            }
        }
    }
}

/* ------------------------------------------------------ LBVH build (cuda_tlas.cuh) -- */

static int32_t to_ordered(float f) {
    int32_t i;
    memcpy(&i, &f, 4);
    return i ^ ((i >> 31) & 0x7fffffff);
}
static float from_ordered(int32_t i) {
    const int32_t b = i ^ ((i >> 31) & 0x7fffffff);
    float f;
    memcpy(&f, &b, 4);
    return f;
}

uint32_t orc_spread_bits(uint32_t v) { /* cuda_tlas.cuh:50-56 */
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

static int clz64(uint64_t x) { return x ? __builtin_clzll(x) : 64; }
static int clz32(uint32_t x) { return x ? __builtin_clz(x) : 32; }

static int common_prefix(const uint64_t *keys, int n, int i, int j) { /* cuda_tlas.cuh:119-124 */
    if (j < 0 || j >= n) return -1;
    const uint64_t a = keys[i], b = keys[j];
    if (a == b) return 64 + clz32((uint32_t)(i ^ j));
    return clz64(a ^ b);
}

/* Stable LSD radix sort of (key, id) pairs over bits [0, end_bit), 8 bits per pass: the contract of
 * cub::DeviceRadixSort::SortPairs as the reference calls it (cuda_tlas.cuh:241-243, :273-274). */
static void stable_radix_sort(uint64_t *keys, int32_t *ids, int n, int end_bit) {
    uint64_t *k2 = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(n > 0 ? n : 1));
    int32_t *i2 = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    for (int shift = 0; shift < end_bit; shift += 8) {
        size_t hist[257] = {0};
        for (int i = 0; i < n; ++i) ++hist[((keys[i] >> shift) & 0xFF) + 1];
        for (int d = 0; d < 256; ++d) hist[d + 1] += hist[d];
        for (int i = 0; i < n; ++i) {
            const size_t pos = hist[(keys[i] >> shift) & 0xFF]++;
            k2[pos] = keys[i];
            i2[pos] = ids[i];
        }
        memcpy(keys, k2, sizeof(uint64_t) * (size_t)n);
        memcpy(ids, i2, sizeof(int32_t) * (size_t)n);
    }
    free(k2);
    free(i2);
}

/* The whole build of cuda_tlas.cuh::Builder::build (:266-283) over every world at once.
 * matrices [N][12]; local_bounds [P][6] per template instance (template index i = g - world_off[w]);
 * world_off [W+1] exclusive prefix of per-world counts (the reference has counts all equal).
 * Outputs (caller-allocated):
 *   inv [N][12], aabb [N][6], world_bounds [W][6] (floats, decoded from the ordered ints),
 *   keys [N] (unsorted, indexed by instance), sorted_keys [N], sorted_ids [N],
 *   nodes [2N-1], parent [2N-1], range [N-1][2], roots [W] (-1 for a world with no instances). */
void orc_lbvh_build(int worlds, const int32_t *world_off, const float *matrices, const float *local_bounds,
                    float *inv, float *aabb, float *world_bounds, uint64_t *keys, uint64_t *sorted_keys,
                    int32_t *sorted_ids, OrcTlasNode *nodes, int32_t *parent, int32_t *range, int32_t *roots) {
    const int n = world_off[worlds];
    int32_t *wb = (int32_t *)malloc(sizeof(int32_t) * 6 * (size_t)worlds);
    int32_t *world_of = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    for (int w = 0; w < worlds; ++w) {
        for (int k = 0; k < 3; ++k) { wb[6 * w + k] = to_ordered(1e30f); wb[6 * w + 3 + k] = to_ordered(-1e30f); }
        for (int g = world_off[w]; g < world_off[w + 1]; ++g) world_of[g] = w;
        roots[w] = -1;
    }
    /* place_instances, cuda_tlas.cuh:69-96 */
    for (int g = 0; g < n; ++g) {
        const int w = world_of[g];
        const float *m = matrices + 12 * (size_t)g;
        orc_invert_affine(m, inv + 12 * (size_t)g);
        const float *lb = local_bounds + 6 * (size_t)(g - world_off[w]);
        float lo[3], hi[3];
        transform_box(m, lb, lo, hi);
        for (int a = 0; a < 3; ++a) {
            aabb[6 * (size_t)g + a] = lo[a];
            aabb[6 * (size_t)g + 3 + a] = hi[a];
            const int32_t c = to_ordered(0.5f * (lo[a] + hi[a]));
            if (c < wb[6 * w + a]) wb[6 * w + a] = c;
            if (c > wb[6 * w + 3 + a]) wb[6 * w + 3 + a] = c;
        }
    }
    for (int i = 0; i < 6 * worlds; ++i) world_bounds[i] = from_ordered(wb[i]);
    /* make_keys, cuda_tlas.cuh:99-115 */
    int32_t *ids = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    for (int g = 0; g < n; ++g) {
        const int w = world_of[g];
        uint32_t q[3];
        for (int a = 0; a < 3; ++a) {
            const float lo = from_ordered(wb[6 * w + a]), hi = from_ordered(wb[6 * w + 3 + a]);
            const float centre = 0.5f * (aabb[6 * (size_t)g + a] + aabb[6 * (size_t)g + 3 + a]);
            const float extent = hi - lo;
            const float t = extent > 1e-20f ? (centre - lo) / extent : 0.0f;
            q[a] = (uint32_t)(fminf(fmaxf(t, 0.0f), 1.0f) * 1023.0f);
        }
        const uint32_t morton = (orc_spread_bits(q[0]) << 2) | (orc_spread_bits(q[1]) << 1) | orc_spread_bits(q[2]);
        keys[g] = ((uint64_t)w << 32) | morton;
        ids[g] = g;
    }
    /* the sort, cuda_tlas.cuh:238-243 */
    int world_bits = 1;
    while ((1 << world_bits) < worlds) ++world_bits;
    memcpy(sorted_keys, keys, sizeof(uint64_t) * (size_t)n);
    memcpy(sorted_ids, ids, sizeof(int32_t) * (size_t)n);
    stable_radix_sort(sorted_keys, sorted_ids, n, 32 + world_bits);
    if (n == 0) { free(ids); free(world_of); free(wb); return; }
    /* memsets, make_leaves, build_radix_tree (cuda_tlas.cuh:128-175, :276-279) */
    for (int v = 0; v < 2 * n - 1; ++v) parent[v] = -1;
    int32_t *leaf_world = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    for (int k = 0; k < n; ++k) {
        const int id = sorted_ids[k];
        OrcTlasNode *leaf = &nodes[n - 1 + k];
        for (int a = 0; a < 3; ++a) { leaf->lo[a] = aabb[6 * (size_t)id + a]; leaf->hi[a] = aabb[6 * (size_t)id + 3 + a]; }
        leaf->left = ~id;
        leaf->right = -1;
        leaf_world[k] = world_of[id];
    }
    for (int i = 0; i < n - 1; ++i) {
        const uint64_t *K = sorted_keys;
        const int d = common_prefix(K, n, i, i + 1) > common_prefix(K, n, i, i - 1) ? 1 : -1;
        const int min_prefix = common_prefix(K, n, i, i - d);
        int max_length = 2;
        while (common_prefix(K, n, i, i + max_length * d) > min_prefix) max_length *= 2;
        int length = 0;
        for (int t = max_length / 2; t >= 1; t /= 2)
            if (common_prefix(K, n, i, i + (length + t) * d) > min_prefix) length += t;
        const int j = i + length * d;
        const int node_prefix = common_prefix(K, n, i, j);
        int s = 0;
        for (int t = (length + 1) / 2;; t = (t + 1) / 2) {
            if (common_prefix(K, n, i, i + (s + t) * d) > node_prefix) s += t;
            if (t == 1) break;
        }
        const int gamma = i + s * d + (d < 0 ? d : 0);
        const int lo = i < j ? i : j, hi = i < j ? j : i;
        const int left = (lo == gamma) ? n - 1 + gamma : gamma;
        const int right = (hi == gamma + 1) ? n - 1 + gamma + 1 : gamma + 1;
        nodes[i].left = left;
        nodes[i].right = right;
        parent[left] = i;
        parent[right] = i;
        range[2 * i] = lo;
        range[2 * i + 1] = hi;
    }
    /* fit_boxes (cuda_tlas.cuh:179-196): min and max do not depend on the order, so a post-order
     * walk from the root gives the same boxes as the arrival-counter climb. */
    if (n > 1) {
        int32_t *stack = (int32_t *)malloc(sizeof(int32_t) * (size_t)(2 * n));
        uint8_t *seen = (uint8_t *)calloc((size_t)n, 1);
        int sp = 0;
        stack[sp++] = 0;
        while (sp > 0) {
            const int v = stack[sp - 1];
            const int l = nodes[v].left, r = nodes[v].right;
            if (!seen[v]) {
                seen[v] = 1;
                if (l < n - 1) stack[sp++] = l;
                if (r < n - 1) stack[sp++] = r;
                continue;
            }
            --sp;
            for (int a = 0; a < 3; ++a) {
                nodes[v].lo[a] = fminf(nodes[l].lo[a], nodes[r].lo[a]);
                nodes[v].hi[a] = fmaxf(nodes[l].hi[a], nodes[r].hi[a]);
            }
        }
        free(seen);
        free(stack);
    }
    /* find_roots, cuda_tlas.cuh:199-211 */
    for (int v = 0; v < 2 * n - 1; ++v) {
        const int rx = v < n - 1 ? range[2 * v] : v - (n - 1);
        const int ry = v < n - 1 ? range[2 * v + 1] : v - (n - 1);
        const int world = leaf_world[rx];
        if (leaf_world[ry] != world) continue;
        const int p = parent[v];
        if (p < 0) { roots[world] = v; continue; }
        if (leaf_world[range[2 * p]] != leaf_world[range[2 * p + 1]]) roots[world] = v;
    }
    free(leaf_world);
    free(ids);
    free(world_of);
    free(wb);
}

/* ------------------------------------------- CPU TLAS (instances.cpp), per world -- */

typedef struct {
    int worlds;
    int32_t *world_off;     /* W+1 */
    OrcNode **trees;        /* per world */
    int *tree_nodes;
    uint32_t **orders;
    float *built_cost;
    float *boxes;           /* [N][6] */
    OrcTlasNode *nodes;     /* flattened forest */
    int node_count, node_cap;
    int32_t *roots;
} OrcSahTlas;

static int push_node(OrcSahTlas *s) {
    if (s->node_count == s->node_cap) {
        s->node_cap = s->node_cap ? 2 * s->node_cap : 1024;
        s->nodes = (OrcTlasNode *)realloc(s->nodes, sizeof(OrcTlasNode) * (size_t)s->node_cap);
    }
    memset(&s->nodes[s->node_count], 0, sizeof(OrcTlasNode));
    return s->node_count++;
}

static int append_run(OrcSahTlas *s, int w, const uint32_t *order, uint32_t first, uint32_t count) { /* :98-121 */
    const int self = push_node(s);
    const int base = s->world_off[w];
    if (count == 1) {
        const float *b = &s->boxes[6 * ((size_t)base + order[first])];
        for (int k = 0; k < 3; ++k) { s->nodes[self].lo[k] = b[k]; s->nodes[self].hi[k] = b[3 + k]; }
        s->nodes[self].left = ~(int32_t)(base + order[first]);
        s->nodes[self].right = -1;
        return self;
    }
    const uint32_t half = count / 2;
    const int left = append_run(s, w, order, first, half);
    const int right = append_run(s, w, order, first + half, count - half);
    for (int k = 0; k < 3; ++k) {
        s->nodes[self].lo[k] = f_min(s->nodes[left].lo[k], s->nodes[right].lo[k]);
        s->nodes[self].hi[k] = f_max(s->nodes[left].hi[k], s->nodes[right].hi[k]);
    }
    s->nodes[self].left = left;
    s->nodes[self].right = right;
    return self;
}

static void flatten(OrcSahTlas *s, int w) { /* instances.cpp:68-95 */
    const OrcNode *tree = s->trees[w];
    const int nn = s->tree_nodes[w];
    const int base = s->node_count;
    for (int i = 0; i < nn; ++i) push_node(s);
    s->roots[w] = base;
    const int g0 = s->world_off[w];
    for (int i = 0; i < nn; ++i) {
        const OrcNode n = tree[i];
        const int self = base + i;
        for (int k = 0; k < 3; ++k) { s->nodes[self].lo[k] = n.lo[k]; s->nodes[self].hi[k] = n.hi[k]; }
        if (n.count == 0) {
            s->nodes[self].left = base + n.a;
            s->nodes[self].right = base + n.a + 1;
        } else if (n.count == 1) {
            s->nodes[self].left = ~(int32_t)(g0 + s->orders[w][n.a]);
            s->nodes[self].right = -1;
        } else {
            const uint32_t first = (uint32_t)n.a, half = (uint32_t)n.count / 2;
            const int left = append_run(s, w, s->orders[w], first, half);
            const int right = append_run(s, w, s->orders[w], first + half, (uint32_t)n.count - half);
            s->nodes[self].left = left;
            s->nodes[self].right = right;
        }
    }
}

OrcSahTlas *orc_sah_tlas_create(int worlds, const int32_t *world_off) {
    OrcSahTlas *s = (OrcSahTlas *)calloc(1, sizeof(OrcSahTlas));
    s->worlds = worlds;
    s->world_off = (int32_t *)malloc(sizeof(int32_t) * (size_t)(worlds + 1));
    memcpy(s->world_off, world_off, sizeof(int32_t) * (size_t)(worlds + 1));
    s->trees = (OrcNode **)calloc((size_t)worlds, sizeof(OrcNode *));
    s->tree_nodes = (int *)calloc((size_t)worlds, sizeof(int));
    s->orders = (uint32_t **)calloc((size_t)worlds, sizeof(uint32_t *));
    s->built_cost = (float *)calloc((size_t)worlds, sizeof(float));
    s->boxes = (float *)malloc(sizeof(float) * 6 * (size_t)(world_off[worlds] > 0 ? world_off[worlds] : 1));
    s->roots = (int32_t *)malloc(sizeof(int32_t) * (size_t)worlds);
    for (int w = 0; w < worlds; ++w) s->roots[w] = -1;
    return s;
}

void orc_sah_tlas_destroy(OrcSahTlas *s) {
    if (!s) return;
    for (int w = 0; w < s->worlds; ++w) { free(s->trees[w]); free(s->orders[w]); }
    free(s->trees); free(s->tree_nodes); free(s->orders); free(s->built_cost);
    free(s->boxes); free(s->roots); free(s->nodes); free(s->world_off);
    free(s);
}

/* InstanceSet::update (instances.cpp:31-64): inverses, world boxes, then per world a refit or a
 * rebuild above 1.3x the cost it had when built, and the forest flattened afresh. inv_out [N][12]. */
void orc_sah_tlas_update(OrcSahTlas *s, const float *matrices, const float *local_bounds, float *inv_out) {
    const int worlds = s->worlds;
    for (int w = 0; w < worlds; ++w)
        for (int g = s->world_off[w]; g < s->world_off[w + 1]; ++g) {
            const float *m = matrices + 12 * (size_t)g;
            orc_invert_affine(m, inv_out + 12 * (size_t)g);
            const float *lb = local_bounds + 6 * (size_t)(g - s->world_off[w]);
            /* transform_aabb from empty_aabb (mesh_build.cpp:185-197) */
            float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
            for (int corner = 0; corner < 8; ++corner) {
                const float p[3] = {(corner & 1) ? lb[3] : lb[0], (corner & 2) ? lb[4] : lb[1], (corner & 4) ? lb[5] : lb[2]};
                for (int a = 0; a < 3; ++a) {
                    const float v = m[4 * a] * p[0] + m[4 * a + 1] * p[1] + m[4 * a + 2] * p[2] + m[4 * a + 3];
                    lo[a] = fminf(lo[a], v);
                    hi[a] = fmaxf(hi[a], v);
                }
            }
            for (int a = 0; a < 3; ++a) { s->boxes[6 * (size_t)g + a] = lo[a]; s->boxes[6 * (size_t)g + 3 + a] = hi[a]; }
        }
    s->node_count = 0;
    for (int w = 0; w < worlds; ++w) {
        const int count = s->world_off[w + 1] - s->world_off[w];
        if (count == 0) { s->roots[w] = -1; continue; }
        const float *boxes = s->boxes + 6 * (size_t)s->world_off[w];
        int rebuilt = s->trees[w] == NULL;
        if (!rebuilt) {
            orc_sah_refit(s->trees[w], s->tree_nodes[w], s->orders[w], boxes);
            rebuilt = orc_sah_cost(s->trees[w], s->tree_nodes[w]) > 1.3f * s->built_cost[w];
        }
        if (rebuilt) {
            if (!s->trees[w]) {
                s->trees[w] = (OrcNode *)malloc(sizeof(OrcNode) * (size_t)(2 * count + 1));
                s->orders[w] = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)count);
            }
            s->tree_nodes[w] = orc_sah_build(boxes, count, 1, s->trees[w], s->orders[w]);
            s->built_cost[w] = orc_sah_cost(s->trees[w], s->tree_nodes[w]);
        }
        flatten(s, w);
    }
}

int orc_sah_tlas_node_count(const OrcSahTlas *s) { return s->node_count; }
const OrcTlasNode *orc_sah_tlas_nodes(const OrcSahTlas *s) { return s->nodes; }
const int32_t *orc_sah_tlas_roots(const OrcSahTlas *s) { return s->roots; }

/* ------------------------------------------------------ traversal (mesh_dev.h) -- */

typedef struct { float x, y, z; } f3;
static f3 mk(float x, float y, float z) { f3 r = {x, y, z}; return r; }
static f3 add3(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static f3 sub3(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static f3 mul3(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
static float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static f3 cross3(f3 a, f3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static f3 normalize3(f3 v) { return mul3(v, 1.0f / sqrtf(dot3(v, v))); } /* geom_dev.h:45-51, host branch */

typedef struct {
    const OrcTlasNode *tlas;
    const int32_t *tlas_root;     /* per world, -1 for none */
    const float *inv;             /* [N][12] */
    const int32_t *blas_root;     /* [N] */
    const OrcNode *blas;
    const float *tris;            /* [T][12] */
    const float *attrs;           /* [T][16] or NULL */
    const int32_t *world_off;     /* [W+1] */
    int instance_count;
    /* spheres: [S][4] centre xyz, radius; floor: y, half_extent_x, z_min, z_max */
    int sphere_count;
    const float *spheres;
    float floor_y, floor_hx, floor_zmin, floor_zmax;
} OrcScene;

static f3 safe_inverse(f3 d) { /* mesh_dev.h:58-61 */
    f3 r;
    r.x = 1.0f / (fabsf(d.x) < 1e-12f ? copysignf(1e-12f, d.x) : d.x);
    r.y = 1.0f / (fabsf(d.y) < 1e-12f ? copysignf(1e-12f, d.y) : d.y);
    r.z = 1.0f / (fabsf(d.z) < 1e-12f ? copysignf(1e-12f, d.z) : d.z);
    return r;
}

static float box_entry(const float *lo, const float *hi, f3 o, f3 inv_d, float tmax) { /* mesh_dev.h:66-74 */
    const float tx0 = (lo[0] - o.x) * inv_d.x, tx1 = (hi[0] - o.x) * inv_d.x;
    const float ty0 = (lo[1] - o.y) * inv_d.y, ty1 = (hi[1] - o.y) * inv_d.y;
    const float tz0 = (lo[2] - o.z) * inv_d.z, tz1 = (hi[2] - o.z) * inv_d.z;
    const float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), 0.0f));
    const float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), tmax));
    return tn <= tf ? tn : ORC_NO_HIT;
}

static int triangle_hit(const float *tris, int i, f3 o, f3 d, float tmax, float *t, float *bu, float *bv) {
    const float *r = tris + 12 * (size_t)i; /* mesh_dev.h:80-99 */
    const f3 v0 = mk(r[0], r[1], r[2]), e1 = mk(r[4], r[5], r[6]), e2 = mk(r[8], r[9], r[10]);
    const f3 p = cross3(d, e2);
    const float det = dot3(e1, p);
    if (fabsf(det) < 1e-14f) return 0;
    const float inv_det = 1.0f / det;
    const f3 s = sub3(o, v0);
    const float u = dot3(s, p) * inv_det;
    if (u < 0.0f || u > 1.0f) return 0;
    const f3 q = cross3(s, e1);
    const float v = dot3(d, q) * inv_det;
    if (v < 0.0f || u + v > 1.0f) return 0;
    *t = dot3(e2, q) * inv_det;
    *bu = u;
    *bv = v;
    return *t > 0.0f && *t < tmax;
}

typedef struct { float t; int inst, tri; float u, v; int found; } OrcHit;

static void blas_walk(const OrcScene *m, int root, int inst, f3 lo, f3 ld, float *tmax, OrcHit *hit) {
    const OrcNode *nodes = m->blas; /* traverse<false>, mesh_dev.h:113-159 */
    const f3 inv_d = safe_inverse(ld);
    if (box_entry(nodes[root].lo, nodes[root].hi, lo, inv_d, *tmax) >= ORC_NO_HIT) return;
    int stack[ORC_BLAS_STACK];
    float stack_t[ORC_BLAS_STACK];
    int sp = 0, node = root;
    for (;;) {
        const OrcNode *n = &nodes[node];
        int descended = 0;
        if (n->count > 0) {
            for (int j = n->a; j < n->a + n->count; ++j) {
                float t, bu, bv;
                if (triangle_hit(m->tris, j, lo, ld, *tmax, &t, &bu, &bv)) {
                    *tmax = t; hit->inst = inst; hit->tri = j; hit->u = bu; hit->v = bv; hit->found = 1;
                }
            }
        } else {
            const int a = n->a, right = n->a + 1;
            const float dl = box_entry(nodes[a].lo, nodes[a].hi, lo, inv_d, *tmax);
            const float dr = box_entry(nodes[right].lo, nodes[right].hi, lo, inv_d, *tmax);
            if (dl < ORC_NO_HIT || dr < ORC_NO_HIT) {
                const int left_first = dl <= dr;
                const int near_node = left_first ? a : right, far_node = left_first ? right : a;
                const float far_t = left_first ? dr : dl;
                if (far_t < ORC_NO_HIT) {
                    if (sp >= ORC_BLAS_STACK) abort();
                    stack[sp] = far_node; stack_t[sp] = far_t; ++sp;
                }
                node = near_node;
                descended = 1;
            }
        }
        if (descended) continue;
        while (sp > 0 && stack_t[sp - 1] >= *tmax) --sp;
        if (sp == 0) return;
        node = stack[--sp];
    }
}

static void instance_leaf(const OrcScene *m, int id, f3 o, f3 d, float *tmax, OrcHit *hit) {
    const float *iv = m->inv + 12 * (size_t)id; /* mesh_dev.h:164-171, :186-200 */
    const f3 r0 = mk(iv[0], iv[1], iv[2]), r1 = mk(iv[4], iv[5], iv[6]), r2 = mk(iv[8], iv[9], iv[10]);
    const f3 lo = mk(dot3(r0, o) + iv[3], dot3(r1, o) + iv[7], dot3(r2, o) + iv[11]);
    const f3 ld = mk(dot3(r0, d), dot3(r1, d), dot3(r2, d));
    blas_walk(m, m->blas_root[id], id, lo, ld, tmax, hit);
}

/* mesh_intersect, mesh_dev.h:177-205: the closest hit among the instances of one world. */
static void mesh_intersect(const OrcScene *m, int world, f3 o, f3 d, float *tmax, OrcHit *hit) {
    const int root = m->tlas_root[world];
    if (root < 0) return;
    const OrcTlasNode *nodes = m->tlas;
    const f3 inv_d = safe_inverse(d);
    if (box_entry(nodes[root].lo, nodes[root].hi, o, inv_d, *tmax) >= ORC_NO_HIT) return;
    int stack[ORC_TLAS_STACK];
    float stack_t[ORC_TLAS_STACK];
    int sp = 0, node = root;
    for (;;) {
        const OrcTlasNode *n = &nodes[node];
        int descended = 0;
        if (n->left < 0) {
            instance_leaf(m, ~n->left, o, d, tmax, hit);
        } else {
            const int a = n->left, right = n->right;
            const float dl = box_entry(nodes[a].lo, nodes[a].hi, o, inv_d, *tmax);
            const float dr = box_entry(nodes[right].lo, nodes[right].hi, o, inv_d, *tmax);
            if (dl < ORC_NO_HIT || dr < ORC_NO_HIT) {
                const int left_first = dl <= dr;
                const int near_node = left_first ? a : right, far_node = left_first ? right : a;
                const float far_t = left_first ? dr : dl;
                if (far_t < ORC_NO_HIT) {
                    if (sp >= ORC_TLAS_STACK) abort();
                    stack[sp] = far_node; stack_t[sp] = far_t; ++sp;
                }
                node = near_node;
                descended = 1;
            }
        }
        if (descended) continue;
        while (sp > 0 && stack_t[sp - 1] >= *tmax) --sp;
        if (sp == 0) return;
        node = stack[--sp];
    }
}

static int sphere_hit(const float *s, f3 origin, f3 dir, float *t) { /* render_core.h:65-75 */
    const f3 L = sub3(mk(s[0], s[1], s[2]), origin);
    const float tca = dot3(L, dir);
    const float r2 = s[3] * s[3];
    const float d2 = dot3(L, L) - tca * tca;
    if (d2 > r2) return 0;
    const float thc = sqrtf(r2 - d2);
    *t = tca - thc;
    if (*t < 0.0f) *t = tca + thc;
    return *t >= 0.0f;
}

typedef struct { float t; f3 normal; int object; int inst, tri; float u, v; } OrcSceneHit;

/* scene_intersect (render_core.h:78-143) and mesh_surface's normal (mesh_dev.h:238-262). */
static int scene_intersect(const OrcScene *s, int world, f3 origin, f3 dir, OrcSceneHit *out) {
    float best = 1e30f;
    int best_sphere = -1;
    for (int i = 0; i < s->sphere_count; ++i) {
        float t;
        if (sphere_hit(s->spheres + 4 * i, origin, dir, &t) && t < best) { best = t; best_sphere = i; }
    }
    int hit_floor = 0;
    if (fabsf(dir.y) > 1e-3f) {
        const float d = (s->floor_y - origin.y) / dir.y;
        if (d > 0.0f && d < best) {
            const f3 pt = add3(origin, mul3(dir, d));
            if (fabsf(pt.x) < s->floor_hx && pt.z > s->floor_zmin && pt.z < s->floor_zmax) { best = d; hit_floor = 1; }
        }
    }
    OrcHit mh = {0};
    if (s->instance_count > 0) mesh_intersect(s, world, origin, dir, &best, &mh);
    if (mh.found) { best_sphere = -1; hit_floor = 0; }
    if (best_sphere < 0 && !hit_floor && !mh.found) return 0;
    out->t = best;
    out->inst = -1;
    const f3 point = add3(origin, mul3(dir, best));
    if (mh.found) {
        const float *iv = s->inv + 12 * (size_t)mh.inst;
        const f3 r0 = mk(iv[0], iv[1], iv[2]), r1 = mk(iv[4], iv[5], iv[6]), r2 = mk(iv[8], iv[9], iv[10]);
        const float *tr = s->tris + 12 * (size_t)mh.tri;
        const f3 nl = cross3(mk(tr[4], tr[5], tr[6]), mk(tr[8], tr[9], tr[10]));
        f3 geometric = normalize3(add3(add3(mul3(r0, nl.x), mul3(r1, nl.y)), mul3(r2, nl.z)));
        if (dot3(geometric, dir) > 0.0f) geometric = mk(-geometric.x, -geometric.y, -geometric.z);
        out->normal = geometric;
        if (s->attrs) {
            const float *a = s->attrs + 16 * (size_t)mh.tri;
            int32_t flags;
            memcpy(&flags, &a[15], 4);
            if (flags & 1) {
                const float w0 = 1.0f - mh.u - mh.v;
                const f3 n = add3(add3(mul3(mk(a[0], a[1], a[2]), w0), mul3(mk(a[4], a[5], a[6]), mh.u)),
                                  mul3(mk(a[8], a[9], a[10]), mh.v));
                if (dot3(n, n) > 1e-12f) {
                    const f3 smooth = normalize3(add3(add3(mul3(r0, n.x), mul3(r1, n.y)), mul3(r2, n.z)));
                    out->normal = dot3(smooth, geometric) < 0.0f ? mk(-smooth.x, -smooth.y, -smooth.z) : smooth;
                }
            }
        }
        out->object = OBJECT_INSTANCE + (mh.inst - s->world_off[world]);
        out->inst = mh.inst;
        out->tri = mh.tri;
        out->u = mh.u;
        out->v = mh.v;
    } else if (hit_floor) {
        out->object = OBJECT_FLOOR;
        out->normal = mk(0.0f, 1.0f, 0.0f);
    } else {
        const float *sp = s->spheres + 4 * best_sphere;
        out->normal = normalize3(sub3(point, mk(sp[0], sp[1], sp[2])));
        out->object = OBJECT_SPHERE + best_sphere;
    }
    return 1;
}

typedef struct { f3 position, right, up, forward; float fov_scale, aspect; } OrcCamera;

static OrcCamera camera_from_matrix(const float *m, float fov_scale, int width, int height) { /* render_core.h:264-273 */
    OrcCamera c;
    c.position = mk(m[3], m[7], m[11]);
    c.right = normalize3(mk(m[0], m[4], m[8]));
    c.up = normalize3(mk(m[1], m[5], m[9]));
    c.forward = normalize3(mk(-m[2], -m[6], -m[10]));
    c.fov_scale = fov_scale;
    c.aspect = (float)width / (float)height;
    return c;
}

static f3 camera_ray(const OrcCamera *c, float px, float py, int width, int height) { /* device_scene.h:114-119 */
    const float sx = (2.0f * px / (float)width - 1.0f) * c->fov_scale * c->aspect;
    const float sy = -(2.0f * py / (float)height - 1.0f) * c->fov_scale;
    return normalize3(add3(add3(mul3(c->right, sx), mul3(c->up, sy)), c->forward));
}

/* --------------------------------------------------------- parallel chunks (cpu_backend.cpp:94-120) -- */

typedef void (*chunk_fn)(void *ctx, int begin, int end);
typedef struct { chunk_fn fn; void *ctx; int total, chunk; int next; } Pool;

static void *pool_work(void *arg) {
    Pool *p = (Pool *)arg;
    for (;;) {
        const int begin = __atomic_fetch_add(&p->next, p->chunk, __ATOMIC_RELAXED);
        if (begin >= p->total) return NULL;
        p->fn(p->ctx, begin, begin + p->chunk < p->total ? begin + p->chunk : p->total);
    }
}

static void parallel_chunks(int total, int chunk, int threads, chunk_fn fn, void *ctx) {
    Pool p = {fn, ctx, total, chunk, 0};
    int items = (total + chunk - 1) / chunk;
    if (threads > items) threads = items;
    if (threads < 1) threads = 1;
    pthread_t *pool = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)threads);
    for (int i = 1; i < threads; ++i) pthread_create(&pool[i], NULL, pool_work, &p);
    pool_work(&p);
    for (int i = 1; i < threads; ++i) pthread_join(pool[i], NULL);
    free(pool);
}

/* ------------------------------------------------------- render and ray queries -- */

typedef struct {
    const OrcScene *s;
    const OrcCamera *cams;
    int width, height, cpw;
    float *depth, *normal;
    int32_t *object_id;
    int32_t *hit_inst, *hit_tri; /* optional, per pixel */
} RenderCtx;

static void render_rows(void *vctx, int begin, int end) { /* cpu_backend.cpp:255-268, aov_pixel render_core.h:313-327 */
    RenderCtx *c = (RenderCtx *)vctx;
    for (int item = begin; item < end; ++item) {
        const int view = item / c->height, y = item % c->height, world = view / c->cpw;
        const OrcCamera *cam = &c->cams[view];
        for (int x = 0; x < c->width; ++x) {
            const size_t idx = (size_t)view * c->width * c->height + (size_t)y * c->width + x;
            const f3 dir = camera_ray(cam, (float)x + 0.5f, (float)y + 0.5f, c->width, c->height);
            OrcSceneHit hit;
            const int found = scene_intersect(c->s, world, cam->position, dir, &hit);
            if (c->depth) c->depth[idx] = found ? hit.t * dot3(dir, cam->forward) : INFINITY;
            if (c->normal) {
                const f3 n = found ? hit.normal : mk(0.0f, 0.0f, 0.0f);
                c->normal[3 * idx] = n.x; c->normal[3 * idx + 1] = n.y; c->normal[3 * idx + 2] = n.z;
            }
            if (c->object_id) c->object_id[idx] = found ? hit.object : OBJECT_NONE;
            if (c->hit_inst) c->hit_inst[idx] = found ? hit.inst : -1;
            if (c->hit_tri) c->hit_tri[idx] = found && hit.inst >= 0 ? hit.tri : -1;
        }
    }
}

static OrcScene make_scene(const OrcTlasNode *tlas, const int32_t *roots, const float *inv, const int32_t *blas_root,
                           const OrcNode *blas, const float *tris, const float *attrs, int worlds, const int32_t *world_off,
                           int sphere_count, const float *spheres, const float *floor4) {
    OrcScene s;
    s.tlas = tlas; s.tlas_root = roots; s.inv = inv; s.blas_root = blas_root; s.blas = blas; s.tris = tris;
    s.attrs = attrs; s.world_off = world_off; s.instance_count = world_off[worlds];
    s.sphere_count = sphere_count; s.spheres = spheres;
    s.floor_y = floor4 ? floor4[0] : 0.0f;
    s.floor_hx = floor4 ? floor4[1] : 0.0f;
    s.floor_zmin = floor4 ? floor4[2] : 0.0f;
    s.floor_zmax = floor4 ? floor4[3] : 0.0f;
    return s;
}

/* ------------------------------------------------------------------------ RGB picture -- */

/* geom_dev.h:114-132: the PCG hash stream, seeded by pixel and frame */
static uint32_t pcg_hash(uint32_t x) {
    const uint32_t state = x * 747796405u + 2891336453u;
    const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
static float rand_float(uint32_t *state) {
    *state = pcg_hash(*state);
    return (float)(*state) * 2.3283064365386963e-10f;
}
static float saturatef(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); } /* geom_dev.h:77-83, host branch */
static f3 mulv(f3 a, f3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
static f3 reflect3(f3 i, f3 n) { return sub3(i, mul3(n, 2.0f * dot3(i, n))); } /* geom_dev.h:97 */

typedef struct {
    const OrcScene *s;
    const float *inst_mat, *sphere_mat, *floor_mat; /* 8 floats each */
    float floor_checker;
    int light_count;
    const float *lights;                            /* 16 floats each, the device form */
    float ambient;
    f3 sky_up, sky_cu, sky_cd;
    float sky_intensity;
} OrcShade;

static f3 sky(const OrcShade *e, f3 dir) { /* render_core.h:184-188, gradient */
    const float t = saturatef(dot3(dir, e->sky_up) * 0.5f + 0.5f);
    return mul3(add3(e->sky_cd, mul3(sub3(e->sky_cu, e->sky_cd), t)), e->sky_intensity);
}

static f3 offset_origin(f3 p, f3 n, f3 dir) { /* render_core.h:194-196 */
    return add3(p, mul3(n, dot3(dir, n) < 0.0f ? -1e-3f : 1e-3f));
}

static int scene_occluded(const OrcScene *s, int world, f3 o, f3 d, float max_dist) { /* render_core.h:146-160 */
    for (int i = 0; i < s->sphere_count; ++i) {
        float t;
        if (sphere_hit(s->spheres + 4 * i, o, d, &t) && t < max_dist) return 1;
    }
    if (fabsf(d.y) > 1e-3f) {
        const float dist = (s->floor_y - o.y) / d.y;
        if (dist > 0.0f && dist < max_dist) {
            const f3 pt = add3(o, mul3(d, dist));
            if (fabsf(pt.x) < s->floor_hx && pt.z > s->floor_zmin && pt.z < s->floor_zmax) return 1;
        }
    }
    if (s->instance_count == 0) return 0;
    /* mesh_occluded (mesh_dev.h:208-227): a hit closer than max_dist exists exactly when the closest hit
       limited to max_dist is found */
    float tmax = max_dist;
    OrcHit h = {0};
    mesh_intersect(s, world, o, d, &tmax, &h);
    return h.found;
}

static float light_factor(const float *l, f3 towards, float d) { /* device_scene.h:98-111 */
    float f = 1.0f;
    if (l[5] > 0.0f) {
        const float q = d / l[5];
        f = 1.0f / (1.0f + q * q);
    }
    if (l[6] > -1.5f) {
        const float c = -(l[8] * towards.x + l[9] * towards.y + l[10] * towards.z);
        const float t = fminf(fmaxf((c - l[6]) / fmaxf(l[7] - l[6], 1e-6f), 0.0f), 1.0f);
        f *= t * t * (3.0f - 2.0f * t);
    }
    return f;
}

static f3 trace_rgb(const OrcShade *e, int world, f3 origin, f3 dir, uint32_t *rng) { /* render_core.h:200-241 */
    f3 radiance = mk(0.0f, 0.0f, 0.0f);
    float throughput = 1.0f;
    for (int depth = 0;; ++depth) {
        OrcSceneHit hit;
        if (depth > 4 || !scene_intersect(e->s, world, origin, dir, &hit)) {
            radiance = add3(radiance, mul3(sky(e, dir), throughput));
            break;
        }
        const f3 point = add3(origin, mul3(dir, hit.t));
        const float *m;
        float dc[3];
        if (hit.inst >= 0) {
            m = e->inst_mat + 8 * (size_t)(hit.inst - e->s->world_off[world]);
            dc[0] = m[0]; dc[1] = m[1]; dc[2] = m[2];
        } else if (hit.object == OBJECT_FLOOR) {
            m = e->floor_mat;
            const int cx = (int)floorf(point.x * e->floor_checker), cz = (int)floorf(point.z * e->floor_checker);
            const int odd = (cx + cz) & 1;
            dc[0] = 1.0f * 0.3f; dc[1] = (odd ? 1.0f : 0.7f) * 0.3f; dc[2] = (odd ? 1.0f : 0.3f) * 0.3f;
        } else {
            m = e->sphere_mat + 8 * (size_t)(hit.object - OBJECT_SPHERE);
            dc[0] = m[0]; dc[1] = m[1]; dc[2] = m[2];
        }
        f3 diffuse = mk(0.0f, 0.0f, 0.0f), specular = mk(0.0f, 0.0f, 0.0f);
        for (int i = 0; i < e->light_count; ++i) {
            const float *l = e->lights + 16 * i;
            /* sample_sphere_light (geom_dev.h:134-152) */
            const float z = 1.0f - 2.0f * rand_float(rng);
            const float rad = sqrtf(fmaxf(0.0f, 1.0f - z * z));
            const float phi = 6.28318530718f * rand_float(rng);
            const f3 lp = add3(mk(l[0], l[1], l[2]), mul3(mk(rad * cosf(phi), rad * sinf(phi), z), l[4]));
            const f3 to_light = sub3(lp, point);
            const float light_distance = sqrtf(dot3(to_light, to_light));
            const f3 light_dir = mul3(to_light, 1.0f / light_distance);
            const float strength = l[3] * light_factor(l, light_dir, light_distance);
            if (strength <= 0.0f) continue;
            if (scene_occluded(e->s, world, offset_origin(point, hit.normal, light_dir), light_dir, light_distance)) continue;
            const f3 lit = mul3(mk(l[12], l[13], l[14]), strength);
            diffuse = add3(diffuse, mul3(lit, fmaxf(0.0f, dot3(light_dir, hit.normal))));
            specular = add3(specular, mul3(lit, powf(fmaxf(0.0f, dot3(reflect3(light_dir, hit.normal), dir)), m[6])));
        }
        const f3 dcv = mk(dc[0], dc[1], dc[2]);
        const f3 direct = add3(mulv(dcv, mul3(diffuse, m[3])), mul3(specular, m[4]));
        radiance = add3(radiance, mul3(add3(direct, mul3(dcv, e->ambient * m[3])), throughput));
        throughput *= m[5];
        if (throughput < 1e-3f) break;
        dir = reflect3(dir, hit.normal);
        origin = offset_origin(point, hit.normal, dir);
    }
    return radiance;
}

static float srgb_channel(float v) { /* geom_dev.h:166-172, host branch */
    const float x = saturatef(v);
    return x <= 0.0031308f ? 12.92f * x : 1.055f * powf(x, 1.0f / 2.4f) - 0.055f;
}

typedef struct {
    const OrcShade *e;
    const OrcCamera *cams;
    int width, height, cpw, spp, frame;
    uint8_t *rgb;
} RgbCtx;

static void rgb_rows(void *vctx, int begin, int end) { /* render_pixel_rgb8 and sample_pixel, render_core.h:244-283 */
    RgbCtx *c = (RgbCtx *)vctx;
    for (int item = begin; item < end; ++item) {
        const int view = item / c->height, y = item % c->height, world = view / c->cpw;
        const OrcCamera *cam = &c->cams[view];
        for (int x = 0; x < c->width; ++x) {
            uint32_t rng = pcg_hash((uint32_t)(y * c->width + x) ^ pcg_hash((uint32_t)c->frame + 1u));
            f3 sum = mk(0.0f, 0.0f, 0.0f);
            for (int sample = 0; sample < c->spp; ++sample) {
                const float jx = rand_float(&rng);
                const float jy = rand_float(&rng);
                sum = add3(sum, trace_rgb(c->e, world, cam->position, camera_ray(cam, (float)x + jx, (float)y + jy, c->width, c->height), &rng));
            }
            const f3 v = mul3(sum, 1.0f / (float)c->spp);
            uint8_t *out = c->rgb + 3 * ((size_t)view * c->width * c->height + (size_t)y * c->width + x);
            out[0] = (uint8_t)(saturatef(srgb_channel(v.x)) * 255.0f + 0.5f);
            out[1] = (uint8_t)(saturatef(srgb_channel(v.y)) * 255.0f + 0.5f);
            out[2] = (uint8_t)(saturatef(srgb_channel(v.z)) * 255.0f + 0.5f);
        }
    }
}

/* CpuBackend::render_batch (cpu_backend.cpp:239-272) for the RGB picture. Materials: 8 floats each (diffuse
 * rgb, diffuse albedo, specular albedo, reflectivity, specular exponent, 0) per template instance, per sphere
 * and for the floor; lights: 14 floats each as the C ABI's ngx_light (position, intensity, radius, falloff,
 * direction, spot angle, spot softness, colour), turned into the device form here as device_scene.h:77-84
 * does; sky: up (any length), colour up, colour down, intensity. */
void orc_render_rgb(const OrcTlasNode *tlas, const int32_t *roots, const float *inv, const int32_t *blas_root,
                    const OrcNode *blas, const float *tris, const float *attrs, int worlds, const int32_t *world_off,
                    int sphere_count, const float *spheres, const float *floor4,
                    const float *inst_mat, const float *sphere_mat, const float *floor_mat, float floor_checker,
                    int light_count, const float *lights14, float ambient, const float *sky10,
                    const float *cams, float fov_scale, int width, int height, int cameras_per_world, int spp, int frame,
                    uint8_t *rgb, int threads) {
    const OrcScene s = make_scene(tlas, roots, inv, blas_root, blas, tris, attrs, worlds, world_off, sphere_count, spheres, floor4);
    float dev[8 * 16];
    for (int i = 0; i < light_count && i < 8; ++i) {
        const float *l = lights14 + 14 * i;
        float *o = dev + 16 * i;
        float dx = l[6], dy = l[7], dz = l[8];
        const float n = sqrtf(dx * dx + dy * dy + dz * dz);
        if (n > 0.0f) { dx = dx / n; dy = dy / n; dz = dz / n; } else { dx = 0.0f; dy = 0.0f; dz = -1.0f; }
        const int spot = l[9] > 0.0f;
        o[0] = l[0]; o[1] = l[1]; o[2] = l[2]; o[3] = l[3];
        o[4] = l[4]; o[5] = l[5];
        o[6] = spot ? cosf(l[9]) : -2.0f;
        o[7] = spot ? cosf(l[9] * (1.0f - l[10])) : -2.0f;
        o[8] = dx; o[9] = dy; o[10] = dz; o[11] = 0.0f;
        o[12] = l[11]; o[13] = l[12]; o[14] = l[13]; o[15] = 0.0f;
    }
    OrcShade e;
    e.s = &s;
    e.inst_mat = inst_mat; e.sphere_mat = sphere_mat; e.floor_mat = floor_mat; e.floor_checker = floor_checker;
    e.light_count = light_count; e.lights = dev; e.ambient = ambient;
    const double un = sqrt((double)sky10[0] * sky10[0] + (double)sky10[1] * sky10[1] + (double)sky10[2] * sky10[2]);
    e.sky_up = mk((float)(sky10[0] / un), (float)(sky10[1] / un), (float)(sky10[2] / un));
    e.sky_cu = mk(sky10[3], sky10[4], sky10[5]);
    e.sky_cd = mk(sky10[6], sky10[7], sky10[8]);
    e.sky_intensity = sky10[9];
    const int views = worlds * cameras_per_world;
    OrcCamera *cc = (OrcCamera *)malloc(sizeof(OrcCamera) * (size_t)(views > 0 ? views : 1));
    for (int v = 0; v < views; ++v) cc[v] = camera_from_matrix(cams + 12 * (size_t)v, fov_scale, width, height);
    RgbCtx ctx = {&e, cc, width, height, cameras_per_world, spp, frame, rgb};
    parallel_chunks(views * height, 4, threads, rgb_rows, &ctx);
    free(cc);
}

/* CpuBackend::render_batch (cpu_backend.cpp:239-272) for depth, normal and object ID. fov_scale is
 * tanf(fov / 2), computed by the caller as the reference's host does. Any output may be NULL. */
void orc_render_batch(const OrcTlasNode *tlas, const int32_t *roots, const float *inv, const int32_t *blas_root,
                      const OrcNode *blas, const float *tris, const float *attrs, int worlds, const int32_t *world_off,
                      int sphere_count, const float *spheres, const float *floor4,
                      const float *cams, float fov_scale, int width, int height, int cameras_per_world,
                      float *depth, float *normal, int32_t *object_id, int32_t *hit_inst, int32_t *hit_tri, int threads) {
    const OrcScene s = make_scene(tlas, roots, inv, blas_root, blas, tris, attrs, worlds, world_off, sphere_count, spheres, floor4);
    const int views = worlds * cameras_per_world;
    OrcCamera *cc = (OrcCamera *)malloc(sizeof(OrcCamera) * (size_t)(views > 0 ? views : 1));
    for (int v = 0; v < views; ++v) cc[v] = camera_from_matrix(cams + 12 * (size_t)v, fov_scale, width, height);
    RenderCtx ctx = {&s, cc, width, height, cameras_per_world, depth, normal, object_id, hit_inst, hit_tri};
    parallel_chunks(views * height, 4, threads, render_rows, &ctx);
    free(cc);
}

typedef struct {
    const OrcScene *s;
    const float *o, *d;
    float max_distance;
    int rays_per_world;
    float *distance, *normal;
    int32_t *object_id, *hit_inst, *hit_tri;
} RayCtx;

static void ray_chunk(void *vctx, int begin, int end) { /* cast_ray, render_core.h:331-349 */
    RayCtx *c = (RayCtx *)vctx;
    for (int i = begin; i < end; ++i) {
        const int world = i / c->rays_per_world;
        const f3 o = mk(c->o[3 * (size_t)i], c->o[3 * (size_t)i + 1], c->o[3 * (size_t)i + 2]);
        f3 d = mk(c->d[3 * (size_t)i], c->d[3 * (size_t)i + 1], c->d[3 * (size_t)i + 2]);
        const float len2 = dot3(d, d);
        OrcSceneHit hit;
        int found = len2 > 1e-30f && len2 < 1e30f;
        if (found) {
            d = mul3(d, 1.0f / sqrtf(len2));
            found = scene_intersect(c->s, world, o, d, &hit) && hit.t <= c->max_distance;
        }
        c->distance[i] = found ? hit.t : INFINITY;
        const f3 n = found ? hit.normal : mk(0.0f, 0.0f, 0.0f);
        c->normal[3 * (size_t)i] = n.x; c->normal[3 * (size_t)i + 1] = n.y; c->normal[3 * (size_t)i + 2] = n.z;
        c->object_id[i] = found ? hit.object : OBJECT_NONE;
        if (c->hit_inst) c->hit_inst[i] = found ? hit.inst : -1;
        if (c->hit_tri) c->hit_tri[i] = found && hit.inst >= 0 ? hit.tri : -1;
    }
}

/* CpuBackend::cast_rays_batch (cpu_backend.cpp:274-287). */
void orc_cast_rays_batch(const OrcTlasNode *tlas, const int32_t *roots, const float *inv, const int32_t *blas_root,
                         const OrcNode *blas, const float *tris, const float *attrs, int worlds, const int32_t *world_off,
                         int sphere_count, const float *spheres, const float *floor4,
                         const float *origins, const float *directions, int rays_per_world, float max_distance,
                         float *distance, float *normal, int32_t *object_id, int32_t *hit_inst, int32_t *hit_tri, int threads) {
    const OrcScene s = make_scene(tlas, roots, inv, blas_root, blas, tris, attrs, worlds, world_off, sphere_count, spheres, floor4);
    RayCtx ctx = {&s, origins, directions, max_distance, rays_per_world, distance, normal, object_id, hit_inst, hit_tri};
    parallel_chunks(worlds * rays_per_world, 256, threads, ray_chunk, &ctx);
}

/* The primary ray of a pixel, for callers that need it (the FP64 gradient check). */
void orc_camera_ray(const float *cam12, float fov_scale, int width, int height, int x, int y, float *origin, float *dir,
                    float *forward) {
    const OrcCamera c = camera_from_matrix(cam12, fov_scale, width, height);
    const f3 d = camera_ray(&c, (float)x + 0.5f, (float)y + 0.5f, width, height);
    origin[0] = c.position.x; origin[1] = c.position.y; origin[2] = c.position.z;
    dir[0] = d.x; dir[1] = d.y; dir[2] = d.z;
    forward[0] = c.forward.x; forward[1] = c.forward.y; forward[2] = c.forward.z;
}

/* ------------------------------------------------- FP64 pose gradient (no reference) -- */

/* Rotation of a quaternion (x, y, z, w), normalized, in double; and d R / d q for the raw q. */
static void quat_rot_and_grad(const double *q, double R[9], double dR[4][9]) {
    const double s2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
    const double s = sqrt(s2);
    const double x = q[0] / s, y = q[1] / s, z = q[2] / s, w = q[3] / s;
    R[0] = 1 - 2 * (y * y + z * z); R[1] = 2 * (x * y - z * w);     R[2] = 2 * (x * z + y * w);
    R[3] = 2 * (x * y + z * w);     R[4] = 1 - 2 * (x * x + z * z); R[5] = 2 * (y * z - x * w);
    R[6] = 2 * (x * z - y * w);     R[7] = 2 * (y * z + x * w);     R[8] = 1 - 2 * (x * x + y * y);
    /* dR/d(unit components) */
    const double du[4][9] = {
        {0, 2 * y, 2 * z, 2 * y, -4 * x, -2 * w, 2 * z, 2 * w, -4 * x},
        {-4 * y, 2 * x, 2 * w, 2 * x, 0, 2 * z, -2 * w, 2 * z, -4 * y},
        {-4 * z, -2 * w, 2 * x, 2 * w, -4 * z, 2 * y, 2 * x, 2 * y, 0},
        {0, -2 * z, 2 * y, 2 * z, 0, -2 * x, -2 * y, 2 * x, 0}};
    const double qh[4] = {x, y, z, w};
    /* d unit_k / d q_j = (delta_kj - qh_k qh_j) / s */
    for (int j = 0; j < 4; ++j)
        for (int e = 0; e < 9; ++e) {
            double acc = 0.0;
            for (int k = 0; k < 4; ++k) acc += du[k][e] * (((k == j) ? 1.0 : 0.0) - qh[k] * qh[j]) / s;
            dR[j][e] = acc;
        }
}

static void quat_rot_d(const double *q, double R[9]) {
    double dR[4][9];
    quat_rot_and_grad(q, R, dR);
}

/* World-space vertices of a triangle in double from the body pose (7 doubles), the instance's local pose
 * and scale, and the triangle's local v0, e1, e2 (local vertices v0, v0 + e1, v0 + e2). */
static void world_vertices(const double *bp, const float *lp, const float *sc, int has_body, const float *tri, double P[3][3]) {
    double rl[9], rb[9], r[9], p[3];
    const double lq[4] = {lp[3], lp[4], lp[5], lp[6]};
    quat_rot_d(lq, rl);
    if (has_body) {
        const double bq[4] = {bp[3], bp[4], bp[5], bp[6]};
        quat_rot_d(bq, rb);
        for (int row = 0; row < 3; ++row) {
            for (int col = 0; col < 3; ++col)
                r[3 * row + col] = rb[3 * row] * rl[col] + rb[3 * row + 1] * rl[3 + col] + rb[3 * row + 2] * rl[6 + col];
            p[row] = bp[row] + rb[3 * row] * lp[0] + rb[3 * row + 1] * lp[1] + rb[3 * row + 2] * lp[2];
        }
    } else {
        for (int k = 0; k < 9; ++k) r[k] = rl[k];
        for (int k = 0; k < 3; ++k) p[k] = lp[k];
    }
    double v[3][3];
    for (int a = 0; a < 3; ++a) {
        v[0][a] = tri[a];
        v[1][a] = (double)tri[a] + (double)tri[4 + a];
        v[2][a] = (double)tri[a] + (double)tri[8 + a];
    }
    for (int k = 0; k < 3; ++k)
        for (int row = 0; row < 3; ++row)
            P[k][row] = r[3 * row] * sc[0] * v[k][0] + r[3 * row + 1] * sc[1] * v[k][1] + r[3 * row + 2] * sc[2] * v[k][2] + p[row];
}

/* Ray parameter of the plane through P in double. */
static double plane_t(const double P[3][3], const double *o, const double *d) {
    const double e1[3] = {P[1][0] - P[0][0], P[1][1] - P[0][1], P[1][2] - P[0][2]};
    const double e2[3] = {P[2][0] - P[0][0], P[2][1] - P[0][1], P[2][2] - P[0][2]};
    const double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    const double num = n[0] * (P[0][0] - o[0]) + n[1] * (P[0][1] - o[1]) + n[2] * (P[0][2] - o[2]);
    const double den = n[0] * d[0] + n[1] * d[1] + n[2] * d[2];
    return num / den;
}

/* The gradient of sum_i weight[i] * scale[i] * t_i with respect to every body pose, where ray i hits the
 * fixed triangle hit_tri[i] of instance hit_inst[i] (-1: no mesh hit, no contribution) and t_i is its
 * ray parameter; weight is dL/d(depth or range) and scale converts t to the output (dot(dir, forward)
 * for depth, 1 for range). mode 0: closed form; mode 1: central differences with step h.
 * grad: [W][B][7] doubles, zeroed by the caller. Instances whose body is -1 get nothing. */
void orc_pose_grad(int mode, double h, int count, const double *origins, const double *dirs, const double *weight,
                   const double *out_scale, const int32_t *hit_inst, const int32_t *hit_tri, const int32_t *ray_world,
                   const int32_t *world_off, int bodies_per_world, const int32_t *body_of_instance, const float *local_pose,
                   const float *scale, const float *body_poses, const float *tris, double *grad) {
    for (int i = 0; i < count; ++i) {
        const int g = hit_inst[i];
        if (g < 0 || weight[i] == 0.0) continue;
        const int w = ray_world[i];
        const int tmpl = g - world_off[w];
        const int b = body_of_instance[tmpl];
        if (b < 0) continue;
        const float *lp = local_pose + 7 * (size_t)tmpl, *sc = scale + 3 * (size_t)tmpl;
        const float *tri = tris + 12 * (size_t)hit_tri[i];
        const float *bpf = body_poses + 7 * ((size_t)w * bodies_per_world + b);
        double bp[7];
        for (int k = 0; k < 7; ++k) bp[k] = bpf[k];
        double *gout = grad + 7 * ((size_t)w * bodies_per_world + b);
        const double *o = origins + 3 * (size_t)i, *d = dirs + 3 * (size_t)i;
        const double wgt = weight[i] * out_scale[i];
        if (mode == 1) {
            for (int k = 0; k < 7; ++k) {
                double bpp[7], bpm[7], P[3][3];
                memcpy(bpp, bp, sizeof(bp));
                memcpy(bpm, bp, sizeof(bp));
                bpp[k] += h;
                bpm[k] -= h;
                world_vertices(bpp, lp, sc, 1, tri, P);
                const double tp = plane_t(P, o, d);
                world_vertices(bpm, lp, sc, 1, tri, P);
                const double tm = plane_t(P, o, d);
                gout[k] += wgt * (tp - tm) / (2.0 * h);
            }
            continue;
        }
        /* closed form: t moves by n.dX / n.d, where X is the material point hit and n the plane's normal */
        double P[3][3];
        world_vertices(bp, lp, sc, 1, tri, P);
        const double e1[3] = {P[1][0] - P[0][0], P[1][1] - P[0][1], P[1][2] - P[0][2]};
        const double e2[3] = {P[2][0] - P[0][0], P[2][1] - P[0][1], P[2][2] - P[0][2]};
        const double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        const double t = plane_t(P, o, d);
        const double X[3] = {o[0] + t * d[0], o[1] + t * d[1], o[2] + t * d[2]};
        const double nd = n[0] * d[0] + n[1] * d[1] + n[2] * d[2];
        const double a[3] = {wgt * n[0] / nd, wgt * n[1] / nd, wgt * n[2] / nd};
        double R[9], dR[4][9];
        quat_rot_and_grad(bp + 3, R, dR);
        /* body-frame point y = R^T (X - p) */
        const double xp[3] = {X[0] - bp[0], X[1] - bp[1], X[2] - bp[2]};
        const double y[3] = {R[0] * xp[0] + R[3] * xp[1] + R[6] * xp[2], R[1] * xp[0] + R[4] * xp[1] + R[7] * xp[2],
                             R[2] * xp[0] + R[5] * xp[1] + R[8] * xp[2]};
        for (int k = 0; k < 3; ++k) gout[k] += a[k];
        for (int j = 0; j < 4; ++j) {
            double acc = 0.0;
            for (int row = 0; row < 3; ++row)
                acc += a[row] * (dR[j][3 * row] * y[0] + dR[j][3 * row + 1] * y[1] + dR[j][3 * row + 2] * y[2]);
            gout[3 + j] += acc;
        }
    }
}
