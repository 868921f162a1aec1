// Top-level tree build for every world of a batch, one workgroup per world, everything in LDS.
//
// The reference builds one linear BVH over all worlds at once (src/cuda_tlas.cuh:215-306): seven
// kernels plus a device-wide radix sort of (world << 32 | morton) keys, then finds each world's root in
// the global tree. Because the world number sits above the Morton code, the sort never moves an
// instance out of its world, and the Karras hierarchy (cuda_tlas.cuh:128-159) over a world's run of
// keys depends only on those keys and on the positions used to tell equal keys apart. So this kernel
// builds each world's subtree independently, in one launch for the whole batch:
//
//   place     body pose (FP64, instances.cpp:137-164) -> matrix, inverse (cuda_tlas.cuh:35-47),
//             8-corner world box, the world's centre bounds (cuda_tlas.cuh:69-96)
//   keys      30-bit Morton code of the box centre (cuda_tlas.cuh:99-115)
//   sort      stable LSD radix sort, 4-bit digits, ranks from wave ballots (multisplit), staged in LDS
//   hierarchy Karras 2012 per node, equal keys told apart by their GLOBAL positions exactly as the
//             reference's common_prefix does (cuda_tlas.cuh:119-124)
//   boxes     bottom-up union with LDS arrival counters (the fit_boxes climb, cuda_tlas.cuh:179-196)
//   write     the reference's node layout at the reference's node indices, plus the 64-byte wide
//             nodes the tracer reads
//
// Node numbering is the reference's: internal nodes are numbered by sorted position, leaf k is node
// N-1+k. Inside world w (positions a..b) every internal node a < i < b is the same as in the global
// tree. The world's own root covers exactly a..b; it carries index a or index b, whichever the global
// tree does not give to the ancestor that joins this world to its neighbours, which the host decides
// once from the neighbouring world numbers (root_at_end). The ancestors joining worlds are never
// traversed (a ray starts at its world's root) and are not built.
#include "ngx_kernels.h"

namespace ngx {
namespace {

__device__ __forceinline__ int to_ordered(const float f) {
    const int i = __float_as_int(f);
    return i ^ ((i >> 31) & 0x7fffffff);
}
__device__ __forceinline__ float from_ordered(const int i) { return __int_as_float(i ^ ((i >> 31) & 0x7fffffff)); }

__device__ __forceinline__ uint32_t spread_bits(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

// Rotation of a quaternion x, y, z, w in FP64, normalized first; zero gives the identity (instances.cpp:126-133).
__device__ __forceinline__ void quat_to_matrix_d(const float *q, double r[9]) {
    double x = q[0], y = q[1], z = q[2], w = q[3];
    const double n = sqrt(x * x + y * y + z * z + w * w);
    if (n > 0.0) {
        x /= n; y /= n; z /= n; w /= n;
    } else {
        x = y = z = 0.0; w = 1.0;
    }
    r[0] = 1 - 2 * (y * y + z * z); r[1] = 2 * (x * y - z * w);     r[2] = 2 * (x * z + y * w);
    r[3] = 2 * (x * y + z * w);     r[4] = 1 - 2 * (x * x + z * z); r[5] = 2 * (y * z - x * w);
    r[6] = 2 * (x * z - y * w);     r[7] = 2 * (y * z + x * w);     r[8] = 1 - 2 * (x * x + y * y);
}

// An instance's matrix from its body's pose, its own offset and its scale, in FP64 and rounded once, as the
// reference's host does (instances.cpp:137-164); the reference's GPU path composes in FP32.
__device__ void compose_body_pose(const float *body, const float *local, const float *scale, float *m) {
    double rl[9], r[9], p[3] = {local[0], local[1], local[2]};
    quat_to_matrix_d(local + 3, rl);
    if (body == nullptr) {
        for (int k = 0; k < 9; ++k) r[k] = rl[k];
    } else {
        double rb[9];
        quat_to_matrix_d(body + 3, rb);
        for (int row = 0; row < 3; ++row) {
            for (int col = 0; col < 3; ++col)
                r[3 * row + col] = rb[3 * row] * rl[col] + rb[3 * row + 1] * rl[3 + col] + rb[3 * row + 2] * rl[6 + col];
            p[row] = (double)body[row] + rb[3 * row] * (double)local[0] + rb[3 * row + 1] * (double)local[1] +
                     rb[3 * row + 2] * (double)local[2];
        }
    }
    for (int row = 0; row < 3; ++row) {
        for (int col = 0; col < 3; ++col) m[4 * row + col] = (float)(r[3 * row + col] * (double)scale[col]);
        m[4 * row + 3] = (float)p[row];
    }
}

__device__ __forceinline__ void invert_affine(const float *m, float *out) {
    const float c00 = m[5] * m[10] - m[6] * m[9];
    const float c01 = m[6] * m[8] - m[4] * m[10];
    const float c02 = m[4] * m[9] - m[5] * m[8];
    const float inv = 1.0f / (m[0] * c00 + m[1] * c01 + m[2] * c02);
    const float r[9] = {c00 * inv, (m[2] * m[9] - m[1] * m[10]) * inv, (m[1] * m[6] - m[2] * m[5]) * inv,
                        c01 * inv, (m[0] * m[10] - m[2] * m[8]) * inv, (m[2] * m[4] - m[0] * m[6]) * inv,
                        c02 * inv, (m[1] * m[8] - m[0] * m[9]) * inv, (m[0] * m[5] - m[1] * m[4]) * inv};
    for (int row = 0; row < 3; ++row) {
        for (int col = 0; col < 3; ++col) out[4 * row + col] = r[3 * row + col];
        out[4 * row + 3] = -(r[3 * row] * m[3] + r[3 * row + 1] * m[7] + r[3 * row + 2] * m[11]);
    }
}

__device__ __forceinline__ int wave_min(int v) {
    for (int o = 32; o >= 1; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ int wave_max(int v) {
    for (int o = 32; o >= 1; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

// Exclusive scan of data[0 .. count) in place, count <= blockDim.x; wave_tot has one slot per wave.
__device__ void block_exclusive_scan(int *data, const int count, int *wave_tot) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int x = tid < count ? data[tid] : 0;
    const int incl = wave_prefix_sum(x);
    if (lane == 63) wave_tot[wave] = incl;
    __syncthreads();
    int base = 0;
    for (int w = 0; w < wave; ++w) base += wave_tot[w];
    if (tid < count) data[tid] = base + incl - x;
    __syncthreads();
}

// ---- pieces shared by the workgroup build and the device-wide build

// Places instance i of world w (global index g): its matrix as given or composed from its body, the
// inverse the tracer reads, and its world box (the 8 corners of the mesh box, cuda_tlas.cuh:69-96).
__device__ void place_instance(const TlasBuildArgs &A, const int w, const int i, const int g, float lo[3], float hi[3]) {
    float m[12];
    if (A.source == SOURCE_BODIES) {
        const int b = A.body_of_instance[i];
        const float *body = b < 0 ? nullptr : A.body_poses + 7 * ((size_t)w * A.bodies_per_world + b);
        compose_body_pose(body, A.local_pose + 7 * i, A.scale + 3 * i, m);
    } else if (A.source == SOURCE_GEOMS) {
        // R diag(s) | p from FP32 inputs, one FP32 product an element, as the reference's MuJoCo Warp compose
        // kernel does (python/neraulgraphics/mujoco.py:276-287).
        const size_t gw = (size_t)w * A.geoms_per_world + A.geom_of_instance[i];
        const float *p = A.geom_pos + 3 * gw, *r = A.geom_mat + 9 * gw, *sc = A.geom_scale + 3 * i;
        for (int row = 0; row < 3; ++row) {
            for (int col = 0; col < 3; ++col) m[4 * row + col] = r[3 * row + col] * sc[col];
            m[4 * row + 3] = p[row];
        }
    } else {
        const float4 *mi = reinterpret_cast<const float4 *>(A.matrices_in + 12 * ((size_t)w * A.matrices_stride + i));
        const float4 r0 = mi[0], r1 = mi[1], r2 = mi[2];
        m[0] = r0.x; m[1] = r0.y; m[2] = r0.z; m[3] = r0.w;
        m[4] = r1.x; m[5] = r1.y; m[6] = r1.z; m[7] = r1.w;
        m[8] = r2.x; m[9] = r2.y; m[10] = r2.z; m[11] = r2.w;
    }
    float4 *mo = reinterpret_cast<float4 *>(A.matrices_out + 12 * (size_t)g);
    mo[0] = make_float4(m[0], m[1], m[2], m[3]);
    mo[1] = make_float4(m[4], m[5], m[6], m[7]);
    mo[2] = make_float4(m[8], m[9], m[10], m[11]);
    float inv[12];
    invert_affine(m, inv);
    const int mesh = A.tmpl_mesh[i];
    InstanceRec rec;
    rec.inv[0] = make_float4(inv[0], inv[1], inv[2], inv[3]);
    rec.inv[1] = make_float4(inv[4], inv[5], inv[6], inv[7]);
    rec.inv[2] = make_float4(inv[8], inv[9], inv[10], inv[11]);
    rec.mesh = mesh;
    rec.tmpl = i;
    rec.mesh_tris = A.mesh_tris[mesh];
    rec.skip_root = A.mesh_skip_root[mesh];
    A.inst[g] = rec;

    const float *lb = A.mesh_bounds + 6 * mesh;
    for (int ax = 0; ax < 3; ++ax) { lo[ax] = 1e30f; hi[ax] = -1e30f; }
    for (int corner = 0; corner < 8; ++corner) {
        const float p[3] = {(corner & 1) ? lb[3] : lb[0], (corner & 2) ? lb[4] : lb[1], (corner & 4) ? lb[5] : lb[2]};
        for (int ax = 0; ax < 3; ++ax) {
            const float v = m[4 * ax] * p[0] + m[4 * ax + 1] * p[1] + m[4 * ax + 2] * p[2] + m[4 * ax + 3];
            lo[ax] = fminf(lo[ax], v);
            hi[ax] = fmaxf(hi[ax], v);
        }
    }
    if (A.dbg_aabb) {
        float *o = A.dbg_aabb + 6 * (size_t)g;
        for (int ax = 0; ax < 3; ++ax) { o[ax] = lo[ax]; o[3 + ax] = hi[ax]; }
    }
}

// 30-bit Morton code of a box centre within the world's centre bounds (ordered ints), cuda_tlas.cuh:99-115.
__device__ __forceinline__ uint32_t morton_code(const float lo[3], const float hi[3], const int *world_bounds) {
    uint32_t q[3];
    for (int ax = 0; ax < 3; ++ax) {
        const float blo = from_ordered(world_bounds[ax]), bhi = from_ordered(world_bounds[3 + ax]);
        const float centre = 0.5f * (lo[ax] + hi[ax]);
        const float extent = bhi - blo;
        const float t = extent > 1e-20f ? (centre - blo) / extent : 0.0f;
        q[ax] = (uint32_t)(fminf(fmaxf(t, 0.0f), 1.0f) * 1023.0f);
    }
    return (spread_bits(q[0]) << 2) | (spread_bits(q[1]) << 1) | spread_bits(q[2]);
}

// Karras 2012 for internal node i of a world's n sorted keys at global positions a..a+n-1 (cuda_tlas.cuh:
// 119-159): equal keys are told apart by their global positions, keys of other worlds are out of range.
// Children: local node index, or n-1+k for the k-th leaf.
template <class Keys>
__device__ __forceinline__ void karras_node(const int i, const int n, const int a, const Keys &key, int &l, int &r) {
    auto delta = [&](const int p, const int q) -> int {
        if (q < 0 || q >= n) return -1;
        const uint32_t kp = key[p], kq = key[q];
        if (kp == kq) return 64 + __clz((a + p) ^ (a + q));
        return 32 + __clz(kp ^ kq);
    };
    const int d = delta(i, i + 1) > delta(i, i - 1) ? 1 : -1;
    const int min_prefix = delta(i, i - d);
    int max_length = 2;
    while (delta(i, i + max_length * d) > min_prefix) max_length *= 2;
    int length = 0;
    for (int t = max_length / 2; t >= 1; t /= 2)
        if (delta(i, i + (length + t) * d) > min_prefix) length += t;
    const int j = i + length * d;
    const int node_prefix = delta(i, j);
    int s = 0;
    for (int t = (length + 1) / 2;; t = (t + 1) / 2) {
        if (delta(i, i + (s + t) * d) > node_prefix) s += t;
        if (t == 1) break;
    }
    const int gamma = i + s * d + (d < 0 ? d : 0);
    const int lo = min(i, j), hi = max(i, j);
    l = (lo == gamma) ? n - 1 + gamma : gamma;
    r = (hi == gamma + 1) ? n - 1 + gamma + 1 : gamma + 1;
}

// Writes internal node i of world w (runs at a, n instances) in both layouts: box(c, ax, &lo, &hi) gives
// local child c's box, id(k) the template index of the k-th sorted leaf, parent its local parent (-1 at the
// world's root).
template <class BoxOf, class IdOf>
__device__ __forceinline__ void write_interior(const TlasBuildArgs &A, const int i, const int n, const int a, const int root_index,
                                               const int l, const int r, const int parent, const BoxOf &box, const IdOf &id) {
    const int N = A.total;
    const int gi = i == 0 ? root_index : a + i;
    float llo[3], lhi[3], rlo[3], rhi[3];
    for (int ax = 0; ax < 3; ++ax) {
        box(l, ax, llo[ax], lhi[ax]);
        box(r, ax, rlo[ax], rhi[ax]);
    }
    TlasNode node;
    for (int ax = 0; ax < 3; ++ax) {
        node.lo[ax] = fminf(llo[ax], rlo[ax]);
        node.hi[ax] = fmaxf(lhi[ax], rhi[ax]);
    }
    node.left = l >= n - 1 ? N - 1 + a + (l - (n - 1)) : a + l;
    node.right = r >= n - 1 ? N - 1 + a + (r - (n - 1)) : a + r;
    A.nodes[gi] = node;
    WideNode wn;
    for (int ax = 0; ax < 3; ++ax) {
        wn.l_lo[ax] = llo[ax]; wn.l_hi[ax] = lhi[ax];
        wn.r_lo[ax] = rlo[ax]; wn.r_hi[ax] = rhi[ax];
    }
    wn.left = l >= n - 1 ? ~(a + id(l - (n - 1))) : a + l;
    wn.right = r >= n - 1 ? ~(a + id(r - (n - 1))) : a + r;
    wn.left_count = wn.right_count = 1;
    A.wide[gi] = wn;
    A.wide_parent[gi] = parent < 0 ? -1 : (parent == 0 ? root_index : a + parent);
}

template <int MAXN>
struct BuildShared {
    static constexpr int NW = MAXN / 64;
    float box_lo[3][MAXN], box_hi[3][MAXN];    // instance boxes by template index
    float node_lo[3][MAXN], node_hi[3][MAXN];  // interior nodes by local index
    uint32_t key[2][MAXN];
    uint32_t val[2][MAXN];
    int left[MAXN], right[MAXN], arrivals[MAXN];
    int parent[2 * MAXN];
    int hist[16 * NW];
    int wave_tot[NW];
    int bounds[6][NW];
    int world_bounds[6];
};

template <int MAXN>
__global__ __launch_bounds__(MAXN) void tlas_build_kernel(const TlasBuildArgs A) {
    __shared__ BuildShared<MAXN> sh;
    constexpr int NW = MAXN / 64;
    const int w = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int a = A.world_off[w];
    const int n = A.world_off[w + 1] - a;
    const int N = A.total;
    if (n == 0) {
        if (tid == 0) A.roots[w] = -1;
        return;
    }
    if (n > MAXN) return;  // built by the device-wide path
    const bool has = tid < n;

    // ---- place: matrix, inverse, box, vote on the centre bounds
    int cen[3] = {to_ordered(1e30f), to_ordered(1e30f), to_ordered(1e30f)};
    int cmax[3] = {to_ordered(-1e30f), to_ordered(-1e30f), to_ordered(-1e30f)};
    if (has) {
        float lo[3], hi[3];
        place_instance(A, w, tid, a + tid, lo, hi);
        for (int ax = 0; ax < 3; ++ax) {
            sh.box_lo[ax][tid] = lo[ax];
            sh.box_hi[ax][tid] = hi[ax];
            const int c = to_ordered(0.5f * (lo[ax] + hi[ax]));
            cen[ax] = c;
            cmax[ax] = c;
        }
    }
    for (int ax = 0; ax < 3; ++ax) {
        const int mn = wave_min(cen[ax]), mx = wave_max(cmax[ax]);
        if (lane == 0) { sh.bounds[ax][wave] = mn; sh.bounds[3 + ax][wave] = mx; }
    }
    __syncthreads();
    if (tid < 6) {
        int v = sh.bounds[tid][0];
        for (int k = 1; k < NW; ++k) v = tid < 3 ? min(v, sh.bounds[tid][k]) : max(v, sh.bounds[tid][k]);
        sh.world_bounds[tid] = v;
        if (A.dbg_world_bounds) A.dbg_world_bounds[6 * w + tid] = from_ordered(v);
    }
    __syncthreads();

    // ---- keys
    if (has) {
        const float lo[3] = {sh.box_lo[0][tid], sh.box_lo[1][tid], sh.box_lo[2][tid]};
        const float hi[3] = {sh.box_hi[0][tid], sh.box_hi[1][tid], sh.box_hi[2][tid]};
        const uint32_t morton = morton_code(lo, hi, sh.world_bounds);
        sh.key[0][tid] = morton;
        sh.val[0][tid] = (uint32_t)tid;
        if (A.dbg_keys) A.dbg_keys[a + tid] = ((uint64_t)w << 32) | morton;
    }

    // ---- stable LSD radix sort of the n keys, 4 bits a pass; 30-bit keys need passes at bits 0..28
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int pass = 0; pass < 8; ++pass) {
        const int src = pass & 1, dst = src ^ 1, shift = 4 * pass;
        for (int i = tid; i < 16 * NW; i += MAXN) sh.hist[i] = 0;
        __syncthreads();
        const uint32_t k = has ? sh.key[src][tid] : 0u, v = has ? sh.val[src][tid] : 0u;
        const uint32_t digit = (k >> shift) & 15u;
        uint64_t peers = ballot(has);
        for (int bit = 0; bit < 4; ++bit) {
            const bool set = (digit >> bit) & 1u;
            const uint64_t vote = ballot(has && set);
            peers &= set ? vote : ~vote;
        }
        const int rank = popc64(peers & below);
        if (has && rank == 0) sh.hist[digit * NW + wave] = popc64(peers);
        __syncthreads();
        block_exclusive_scan(sh.hist, 16 * NW, sh.wave_tot);
        if (has) {
            const int dest = sh.hist[digit * NW + wave] + rank;
            sh.key[dst][dest] = k;
            sh.val[dst][dest] = v;
        }
        __syncthreads();
    }
    // eight passes end in buffer 0
    const uint32_t *key = sh.key[0];
    const uint32_t *val = sh.val[0];
    if (has) {
        if (A.dbg_sorted_keys) A.dbg_sorted_keys[a + tid] = ((uint64_t)w << 32) | key[tid];
        if (A.dbg_sorted_ids) A.dbg_sorted_ids[a + tid] = a + (int)val[tid];
    }

    // ---- leaves, and the one-instance world
    if (has) {
        const int id = (int)val[tid];
        TlasNode leaf;
        for (int ax = 0; ax < 3; ++ax) { leaf.lo[ax] = sh.box_lo[ax][id]; leaf.hi[ax] = sh.box_hi[ax][id]; }
        leaf.left = ~(a + id);
        leaf.right = -1;
        A.nodes[(size_t)(N - 1) + a + tid] = leaf;
    }
    if (n == 1) {
        if (tid == 0) A.roots[w] = N - 1 + a;
        return;
    }

    // ---- Karras hierarchy over local positions 0..n-1 (global position a + i)
    for (int i = tid; i < 2 * n - 1; i += MAXN) sh.parent[i] = -1;
    if (tid < n - 1) sh.arrivals[tid] = 0;
    __syncthreads();
    if (tid < n - 1) {
        int l, r;
        karras_node(tid, n, a, key, l, r);
        sh.left[tid] = l;
        sh.right[tid] = r;
        sh.parent[l] = tid;
        sh.parent[r] = tid;
    }
    __syncthreads();

    // ---- boxes, bottom up: the second of two children to arrive at a node computes its box
    if (has) {
        int node = n - 1 + tid;
        while (true) {
            const int p = sh.parent[node];
            if (p < 0) break;
            const int prior = __hip_atomic_fetch_add(&sh.arrivals[p], 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (prior == 0) break;
            const int l = sh.left[p], r = sh.right[p];
            for (int ax = 0; ax < 3; ++ax) {
                const float llo = l >= n - 1 ? sh.box_lo[ax][val[l - (n - 1)]] : sh.node_lo[ax][l];
                const float lhi = l >= n - 1 ? sh.box_hi[ax][val[l - (n - 1)]] : sh.node_hi[ax][l];
                const float rlo = r >= n - 1 ? sh.box_lo[ax][val[r - (n - 1)]] : sh.node_lo[ax][r];
                const float rhi = r >= n - 1 ? sh.box_hi[ax][val[r - (n - 1)]] : sh.node_hi[ax][r];
                sh.node_lo[ax][p] = fminf(llo, rlo);
                sh.node_hi[ax][p] = fmaxf(lhi, rhi);
            }
            node = p;
        }
    }
    __syncthreads();

    // ---- write the interior nodes in both layouts
    const int root_index = A.root_at_end[w] ? a + n - 1 : a;
    if (tid == 0) A.roots[w] = root_index;
    if (tid < n - 1) {
        auto box = [&](const int c, const int ax, float &lo, float &hi) {
            lo = c >= n - 1 ? sh.box_lo[ax][val[c - (n - 1)]] : sh.node_lo[ax][c];
            hi = c >= n - 1 ? sh.box_hi[ax][val[c - (n - 1)]] : sh.node_hi[ax][c];
        };
        auto id = [&](const int k) { return (int)val[k]; };
        write_interior(A, tid, n, a, root_index, sh.left[tid], sh.right[tid], sh.parent[tid], box, id);
    }
}

// ------------------------------------------------------------------ device-wide build of a large world --
//
// A world with more instances than one workgroup sorts in LDS is built by a sequence of grid-wide kernels
// with the same arithmetic: place (and the centre bounds by ordered-int atomics, cuda_tlas.cuh:69-96),
// keys, a stable LSD radix sort of 4-bit digits over tiles of 1024 keys (per-tile digit counts, one scan
// of the counts in digit-major order, a stable scatter that ranks each tile's keys with the same wave
// ballots as the workgroup sort), Karras per node, the bottom-up union with arrival counters in global
// memory, and the same node writes.

constexpr int BIG_TILE = 1024;

__global__ __launch_bounds__(256) void big_init_kernel(const BigBuildScratch S, const int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < 2 * n - 1) S.parent[i] = -1;
    if (i < n - 1) S.arrivals[i] = 0;
    if (i < 3) {
        S.bounds[i] = to_ordered(1e30f);
        S.bounds[3 + i] = to_ordered(-1e30f);
    }
}

__global__ __launch_bounds__(256) void big_place_kernel(const TlasBuildArgs A, const BigBuildScratch S, const int w, const int a,
                                                        const int n) {
    __shared__ int part[6][4];
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int cmin[3] = {to_ordered(1e30f), to_ordered(1e30f), to_ordered(1e30f)};
    int cmax[3] = {to_ordered(-1e30f), to_ordered(-1e30f), to_ordered(-1e30f)};
    if (i < n) {
        float lo[3], hi[3];
        place_instance(A, w, i, a + i, lo, hi);
        for (int ax = 0; ax < 3; ++ax) {
            S.box[(size_t)ax * S.capacity + i] = lo[ax];
            S.box[(size_t)(3 + ax) * S.capacity + i] = hi[ax];
            cmin[ax] = cmax[ax] = to_ordered(0.5f * (lo[ax] + hi[ax]));
        }
    }
    for (int ax = 0; ax < 3; ++ax) {
        const int mn = wave_min(cmin[ax]), mx = wave_max(cmax[ax]);
        if (lane == 0) {
            part[ax][wave] = mn;
            part[3 + ax][wave] = mx;
        }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int t = threadIdx.x;
        int v = part[t][0];
        for (int k = 1; k < 4; ++k) v = t < 3 ? min(v, part[t][k]) : max(v, part[t][k]);
        if (t < 3) atomicMin(&S.bounds[t], v);
        else atomicMax(&S.bounds[t], v);
    }
}

__global__ __launch_bounds__(256) void big_keys_kernel(const TlasBuildArgs A, const BigBuildScratch S, const int w, const int a,
                                                       const int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < 6 && A.dbg_world_bounds) A.dbg_world_bounds[6 * w + i] = from_ordered(S.bounds[i]);
    if (i >= n) return;
    float lo[3], hi[3];
    for (int ax = 0; ax < 3; ++ax) {
        lo[ax] = S.box[(size_t)ax * S.capacity + i];
        hi[ax] = S.box[(size_t)(3 + ax) * S.capacity + i];
    }
    const uint32_t morton = morton_code(lo, hi, S.bounds);
    S.key[0][i] = morton;
    S.val[0][i] = (uint32_t)i;
    if (A.dbg_keys) A.dbg_keys[a + i] = ((uint64_t)w << 32) | morton;
}

// The digit of each key of a 1024-key tile, its rank among the equal digits of its wave, and how many
// of its wave share the digit (valid for the rank-0 lane).
__device__ __forceinline__ void tile_digit(const bool has, const uint32_t k, const int shift, uint32_t &digit, int &rank, int &count) {
    const int lane = threadIdx.x & 63;
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    digit = (k >> shift) & 15u;
    uint64_t peers = ballot(has);
    for (int bit = 0; bit < 4; ++bit) {
        const bool set = (digit >> bit) & 1u;
        const uint64_t vote = ballot(has && set);
        peers &= set ? vote : ~vote;
    }
    rank = popc64(peers & below);
    count = popc64(peers);
}

__global__ __launch_bounds__(BIG_TILE) void big_sort_count_kernel(const BigBuildScratch S, const int n, const int src, const int shift,
                                                                  const int tiles) {
    constexpr int NW = BIG_TILE / 64;
    __shared__ int hist[16 * NW];
    const int t = blockIdx.x, tid = threadIdx.x, wave = tid >> 6;
    const int i = t * BIG_TILE + tid;
    const bool has = i < n;
    if (tid < 16 * NW) hist[tid] = 0;
    __syncthreads();
    uint32_t digit;
    int rank, count;
    tile_digit(has, has ? S.key[src][i] : 0u, shift, digit, rank, count);
    if (has && rank == 0) hist[digit * NW + wave] = count;
    __syncthreads();
    if (tid < 16) {
        int sum = 0;
        for (int wv = 0; wv < NW; ++wv) sum += hist[tid * NW + wv];
        S.tile_counts[(size_t)tid * tiles + t] = sum;
    }
}

// Exclusive scan of tile_counts[0 .. total) in place by one workgroup, 1024 entries at a time.
__global__ __launch_bounds__(1024) void big_sort_scan_kernel(const BigBuildScratch S, const int total) {
    __shared__ int wave_tot[16];
    __shared__ int carry;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < total; base += 1024) {
        const int c = carry;
        const int idx = base + tid;
        const int x = idx < total ? S.tile_counts[idx] : 0;
        int incl = wave_prefix_sum(x);
        if (lane == 63) wave_tot[wave] = incl;
        __syncthreads();
        for (int wv = 0; wv < wave; ++wv) incl += wave_tot[wv];
        __syncthreads();
        if (idx < total) S.tile_counts[idx] = c + incl - x;
        if (tid == 1023) carry = c + incl;
        __syncthreads();
    }
}

__global__ __launch_bounds__(BIG_TILE) void big_sort_scatter_kernel(const BigBuildScratch S, const int n, const int src, const int shift,
                                                                    const int tiles) {
    constexpr int NW = BIG_TILE / 64;
    __shared__ int hist[16 * NW];
    __shared__ int wave_tot[NW];
    const int t = blockIdx.x, tid = threadIdx.x, wave = tid >> 6;
    const int i = t * BIG_TILE + tid;
    const bool has = i < n;
    const uint32_t k = has ? S.key[src][i] : 0u, v = has ? S.val[src][i] : 0u;
    if (tid < 16 * NW) hist[tid] = 0;
    __syncthreads();
    uint32_t digit;
    int rank, count;
    tile_digit(has, k, shift, digit, rank, count);
    if (has && rank == 0) hist[digit * NW + wave] = count;
    __syncthreads();
    block_exclusive_scan(hist, 16 * NW, wave_tot);
    if (has) {
        const int local = hist[digit * NW + wave] - hist[digit * NW] + rank;
        const int dest = S.tile_counts[(size_t)digit * tiles + t] + local;
        S.key[src ^ 1][dest] = k;
        S.val[src ^ 1][dest] = v;
    }
}

__global__ __launch_bounds__(256) void big_karras_kernel(const TlasBuildArgs A, const BigBuildScratch S, const int w, const int a,
                                                         const int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint32_t *key = S.key[0];
    const int id = (int)S.val[0][i];
    if (A.dbg_sorted_keys) A.dbg_sorted_keys[a + i] = ((uint64_t)w << 32) | key[i];
    if (A.dbg_sorted_ids) A.dbg_sorted_ids[a + i] = a + id;
    TlasNode leaf;
    for (int ax = 0; ax < 3; ++ax) {
        leaf.lo[ax] = S.box[(size_t)ax * S.capacity + id];
        leaf.hi[ax] = S.box[(size_t)(3 + ax) * S.capacity + id];
    }
    leaf.left = ~(a + id);
    leaf.right = -1;
    A.nodes[(size_t)(A.total - 1) + a + i] = leaf;
    if (i < n - 1) {
        int l, r;
        karras_node(i, n, a, key, l, r);
        S.left[i] = l;
        S.right[i] = r;
        S.parent[l] = i;
        S.parent[r] = i;
    }
}

__device__ __forceinline__ void big_child_box(const BigBuildScratch &S, const int n, const int c, const int ax, float &lo, float &hi) {
    if (c >= n - 1) {
        const int id = (int)S.val[0][c - (n - 1)];
        lo = S.box[(size_t)ax * S.capacity + id];
        hi = S.box[(size_t)(3 + ax) * S.capacity + id];
    } else {
        lo = S.node_box[(size_t)ax * S.capacity + c];
        hi = S.node_box[(size_t)(3 + ax) * S.capacity + c];
    }
}

// Bottom up: the second of a node's two children to arrive computes its box. The arrival counter's
// acquire-release at device scope orders the first child's box before the second child reads it.
__global__ __launch_bounds__(256) void big_refit_kernel(const BigBuildScratch S, const int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    int node = n - 1 + i;
    while (true) {
        const int p = S.parent[node];
        if (p < 0) break;
        const int prior = __hip_atomic_fetch_add(&S.arrivals[p], 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (prior == 0) break;
        const int l = S.left[p], r = S.right[p];
        for (int ax = 0; ax < 3; ++ax) {
            float llo, lhi, rlo, rhi;
            big_child_box(S, n, l, ax, llo, lhi);
            big_child_box(S, n, r, ax, rlo, rhi);
            S.node_box[(size_t)ax * S.capacity + p] = fminf(llo, rlo);
            S.node_box[(size_t)(3 + ax) * S.capacity + p] = fmaxf(lhi, rhi);
        }
        node = p;
    }
}

__global__ __launch_bounds__(256) void big_write_kernel(const TlasBuildArgs A, const BigBuildScratch S, const int w, const int a,
                                                        const int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int root_index = A.root_at_end[w] ? a + n - 1 : a;
    if (i == 0) A.roots[w] = root_index;
    if (i >= n - 1) return;
    auto box = [&](const int c, const int ax, float &lo, float &hi) { big_child_box(S, n, c, ax, lo, hi); };
    auto id = [&](const int k) { return (int)S.val[0][k]; };
    write_interior(A, i, n, a, root_index, S.left[i], S.right[i], S.parent[i], box, id);
}

// ------------------------------------------------------------------ ray order for ray queries --
//
// A query's rays come in the caller's order; a LiDAR's consecutive rays fan out over many degrees, so a
// wavefront of 64 of them shares little of either tree. Each world's rays are put in a stable order of a
// 12-bit direction cell (octahedral map of the direction, 64 x 64 cells, Morton order) by one workgroup per
// world. Up to 65,536 rays a world, one counting sort: each of the workgroup's 16 wavefronts takes a contiguous
// sixteenth of the world's rays, counts their cells into its own row of 16-bit counters (4096 cells x 16 wavefronts =
// 128 KB of LDS) and keeps each cell in 2 bytes of scratch; one scan of the counters in cell-major, wavefront-minor
// order gives each (cell, wavefront) its first place; each wavefront then reads its cells back and writes its rays'
// indices in ray order, a ray's rank among the equal cells of its 64 from ballots: a stable sort by cell, reading
// each direction once (16 bytes a ray of traffic where the radix sort moved 40; lidar_sweeps 1.20 -> 0.57 ms). Above
// 65,536 rays, a stable LSD radix sort over chunks of 1024 rays in three passes of 4-bit digits, through global
// memory; both give the one stable order. The tracer then gives a wavefront 64 consecutive rays of that order;
// results still go to each ray's own place. In the radix sort a ray's cell and its index travel as one 32-bit word
// (cell above bit 20, index below), so a pass moves 4 bytes a ray; the last pass writes the index alone.
constexpr int RAY_ORDER_MAX_CHUNKS = RAY_ORDER_MAX_RAYS / 1024;
constexpr int RAY_ORDER_INDEX_BITS = 20;
static_assert(RAY_ORDER_MAX_RAYS <= (1 << RAY_ORDER_INDEX_BITS), "a world's ray index fits below the cell");
constexpr int RAY_ORDER_ONE_PASS_MAX_RAYS = 65536;

// tile_digit for a BITS-bit digit: the digit of each key of a 1024-key tile, its rank among the equal digits of
// its wave, and how many of its wave share the digit (valid for the rank-0 lane).
template <int BITS>
__device__ __forceinline__ void tile_digit_bits(const bool has, const uint32_t k, const int shift, uint32_t &digit, int &rank,
                                                int &count) {
    const int lane = threadIdx.x & 63;
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    digit = (k >> shift) & ((1u << BITS) - 1u);
    uint64_t peers = ballot(has);
    for (int bit = 0; bit < BITS; ++bit) {
        const bool set = (digit >> bit) & 1u;
        const uint64_t vote = ballot(has && set);
        peers &= set ? vote : ~vote;
    }
    rank = popc64(peers & below);
    count = popc64(peers);
}

__device__ __forceinline__ uint32_t spread6(uint32_t v) {
    uint32_t r = 0;
    for (int b = 0; b < 6; ++b) r |= ((v >> b) & 1u) << (2 * b);
    return r;
}

__device__ __forceinline__ uint32_t direction_cell(const float x, const float y, const float z) {
    const float s = fabsf(x) + fabsf(y) + fabsf(z);
    if (!(s > 0.0f) || !isfinite(s)) return 4095u;
    float u = x / s, v = y / s;
    if (z < 0.0f) {
        const float uu = (1.0f - fabsf(v)) * (u >= 0.0f ? 1.0f : -1.0f);
        const float vv = (1.0f - fabsf(u)) * (v >= 0.0f ? 1.0f : -1.0f);
        u = uu;
        v = vv;
    }
    const uint32_t qu = (uint32_t)fminf(63.0f, fmaxf(0.0f, (u * 0.5f + 0.5f) * 64.0f));
    const uint32_t qv = (uint32_t)fminf(63.0f, fmaxf(0.0f, (v * 0.5f + 0.5f) * 64.0f));
    return (spread6(qu) << 1) | spread6(qv);
}

// The one-pass order of up to 65,536 rays a world (see above); cells is scratch of one 16-bit cell a ray.
__global__ __launch_bounds__(1024) void ray_order_one_pass_kernel(const float *dirs, const int R, uint16_t *cells, int *order) {
    constexpr int NW = 16, CELLS = 4096;
    // per (cell, wavefront), cell-major: first how many of the wavefront's rays fall in the cell, then where the next
    // of them goes. A count is at most 4096 and a place below R <= 65,536, so 16 bits hold either; the scan's
    // running sum runs in 32 bits (past the last occupied slot it may reach R, and those slots are never read).
    __shared__ uint16_t slot[CELLS * NW];
    __shared__ int wave_tot[NW];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const size_t base = (size_t)blockIdx.x * R;
    const int seg = (R + NW - 1) / NW;
    const int first = wave * seg, end = min(first + seg, R);
    uint32_t *slot32 = reinterpret_cast<uint32_t *>(slot);
    for (int i = tid; i < CELLS * NW / 2; i += 1024) slot32[i] = 0u;
    __syncthreads();
    // each ray's cell, kept for the second sweep, and counted: a 32-bit word holds the counters of two neighbouring
    // wavefronts, and a count stays below 2^16, so adding 1 to a wavefront's own half never carries into the other
    for (int i = first + lane; i < end; i += 64) {
        const float *p = dirs + 3 * (base + i);
        const uint32_t c = direction_cell(p[0], p[1], p[2]);
        cells[base + i] = (uint16_t)c;
        atomicAdd(&slot32[(c * NW + wave) >> 1], 1u << (16 * (wave & 1)));
    }
    __syncthreads();
    // exclusive scan of the 65,536 slots: each thread sums 64 consecutive ones, the threads' sums are scanned, then
    // each thread writes its slots' starts
    {
        constexpr int SPAN = CELLS * NW / 1024;
        const int s0 = tid * SPAN;
        int sum = 0;
        for (int k = 0; k < SPAN; ++k) sum += slot[s0 + k];
        const int incl = wave_prefix_sum(sum);
        if (lane == 63) wave_tot[wave] = incl;
        __syncthreads();
        int run = incl - sum;
        for (int wv = 0; wv < wave; ++wv) run += wave_tot[wv];
        for (int k = 0; k < SPAN; ++k) {
            const int n = slot[s0 + k];
            slot[s0 + k] = (uint16_t)run;
            run += n;
        }
    }
    __syncthreads();
    // the rays again, in the same order: each one's place is its cell's next place in this wavefront plus its rank
    // among the equal cells of its 64
    for (int at = first; at < end; at += 64) {
        const int i = at + lane;
        const bool has = i < end;
        const uint32_t c = has ? cells[base + i] : (uint32_t)CELLS;
        uint32_t digit;
        int rank, cnt;
        tile_digit_bits<12>(has, c, 0, digit, rank, cnt);
        int dest = 0;
        if (c < CELLS) dest = (int)slot[c * NW + wave] + rank;
        // every lane has read its cell's place before the cell's first lane moves it on (one wavefront's LDS
        // accesses complete in order)
        if (c < CELLS && rank == 0) slot[c * NW + wave] = (uint16_t)(dest + cnt);
        if (c < CELLS) order[base + dest] = i;
    }
}

template <int BITS, int MAX_CHUNKS>
__global__ __launch_bounds__(1024) void ray_order_kernel(const float *dirs, const int R, uint32_t *ka, uint32_t *kb, int *order) {
    constexpr int NW = 16, NB = 1 << BITS, PASSES = 12 / BITS;
    static_assert(PASSES * BITS == 12 && NB * NW <= 1024, "12-bit keys in whole digits, one histogram entry a thread");
    __shared__ int counts[NB * MAX_CHUNKS];
    __shared__ int hist[NB * NW];
    __shared__ int wave_tot[NW];
    __shared__ int carry;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const size_t base = (size_t)blockIdx.x * R;
    const int chunks = (R + 1023) / 1024;
    for (int i = tid; i < R; i += 1024) {
        const float *p = dirs + 3 * (base + i);
        ka[base + i] = (direction_cell(p[0], p[1], p[2]) << RAY_ORDER_INDEX_BITS) | (uint32_t)i;
    }
    __syncthreads();
    uint32_t *ks = ka, *kd = kb;
    for (int pass = 0; pass < PASSES; ++pass) {
        const int shift = RAY_ORDER_INDEX_BITS + BITS * pass;
        const bool last = pass == PASSES - 1;
        for (int c = 0; c < chunks; ++c) {
            const int i = c * 1024 + tid;
            const bool has = i < R;
            if (tid < NB * NW) hist[tid] = 0;
            __syncthreads();
            uint32_t digit;
            int rank, cnt;
            tile_digit_bits<BITS>(has, has ? ks[base + i] : 0u, shift, digit, rank, cnt);
            if (has && rank == 0) hist[digit * NW + wave] = cnt;
            __syncthreads();
            if (tid < NB) {
                int sum = 0;
                for (int wv = 0; wv < NW; ++wv) sum += hist[tid * NW + wv];
                counts[tid * chunks + c] = sum;
            }
            __syncthreads();
        }
        // exclusive scan of the counts in digit-major order
        const int total = NB * chunks;
        if (tid == 0) carry = 0;
        __syncthreads();
        for (int b0 = 0; b0 < total; b0 += 1024) {
            const int cin = carry;
            const int idx = b0 + tid;
            const int x = idx < total ? counts[idx] : 0;
            int incl = wave_prefix_sum(x);
            if (lane == 63) wave_tot[wave] = incl;
            __syncthreads();
            for (int wv = 0; wv < wave; ++wv) incl += wave_tot[wv];
            __syncthreads();
            if (idx < total) counts[idx] = cin + incl - x;
            if (tid == 1023) carry = cin + incl;
            __syncthreads();
        }
        for (int c = 0; c < chunks; ++c) {
            const int i = c * 1024 + tid;
            const bool has = i < R;
            const uint32_t k = has ? ks[base + i] : 0u;
            if (tid < NB * NW) hist[tid] = 0;
            __syncthreads();
            uint32_t digit;
            int rank, cnt;
            tile_digit_bits<BITS>(has, k, shift, digit, rank, cnt);
            if (has && rank == 0) hist[digit * NW + wave] = cnt;
            __syncthreads();
            block_exclusive_scan(hist, NB * NW, wave_tot);
            if (has) {
                const int dest = counts[digit * chunks + c] + hist[digit * NW + wave] - hist[digit * NW] + rank;
                if (last) order[base + dest] = (int)(k & ((1u << RAY_ORDER_INDEX_BITS) - 1u));
                else kd[base + dest] = k;
            }
            __syncthreads();
        }
        uint32_t *tk = ks; ks = kd; kd = tk;
    }
}

}  // namespace

hipError_t launch_tlas_build(const TlasBuildArgs &a, const int max_count, hipStream_t stream) {
    if (a.worlds <= 0) return hipSuccess;
    if (max_count <= 64) tlas_build_kernel<64><<<a.worlds, 64, 0, stream>>>(a);
    else if (max_count <= 128) tlas_build_kernel<128><<<a.worlds, 128, 0, stream>>>(a);
    else if (max_count <= 256) tlas_build_kernel<256><<<a.worlds, 256, 0, stream>>>(a);
    else if (max_count <= 512) tlas_build_kernel<512><<<a.worlds, 512, 0, stream>>>(a);
    else tlas_build_kernel<1024><<<a.worlds, 1024, 0, stream>>>(a);  // larger worlds return at once
    return hipGetLastError();
}

hipError_t launch_tlas_build_big(const TlasBuildArgs &a, const BigBuildScratch &s, const int world, const int first, const int n,
                                 hipStream_t stream) {
    if (n <= 1 || n > s.capacity) return hipErrorInvalidValue;
    const unsigned blocks = (unsigned)((n + 255) / 256), blocks2 = (unsigned)((2 * n - 1 + 255) / 256);
    const int tiles = (n + BIG_TILE - 1) / BIG_TILE;
    big_init_kernel<<<blocks2, 256, 0, stream>>>(s, n);
    big_place_kernel<<<blocks, 256, 0, stream>>>(a, s, world, first, n);
    big_keys_kernel<<<blocks, 256, 0, stream>>>(a, s, world, first, n);
    for (int pass = 0; pass < 8; ++pass) {
        const int src = pass & 1;
        big_sort_count_kernel<<<tiles, BIG_TILE, 0, stream>>>(s, n, src, 4 * pass, tiles);
        big_sort_scan_kernel<<<1, 1024, 0, stream>>>(s, 16 * tiles);
        big_sort_scatter_kernel<<<tiles, BIG_TILE, 0, stream>>>(s, n, src, 4 * pass, tiles);
    }
    big_karras_kernel<<<blocks, 256, 0, stream>>>(a, s, world, first, n);
    big_refit_kernel<<<blocks, 256, 0, stream>>>(s, n);
    big_write_kernel<<<blocks, 256, 0, stream>>>(a, s, world, first, n);
    return hipGetLastError();
}

hipError_t launch_ray_order(const float *dirs, const int rays_per_world, const int worlds, uint32_t *keys, int *vals, int **order,
                            hipStream_t stream) {
    if (rays_per_world <= 0 || rays_per_world > RAY_ORDER_MAX_RAYS || worlds <= 0) return hipErrorInvalidValue;
    const size_t n = (size_t)rays_per_world * worlds;
    if (rays_per_world <= RAY_ORDER_ONE_PASS_MAX_RAYS) {
        ray_order_one_pass_kernel<<<worlds, 1024, 0, stream>>>(dirs, rays_per_world, reinterpret_cast<uint16_t *>(keys), vals);
    } else {
        ray_order_kernel<4, RAY_ORDER_MAX_CHUNKS><<<worlds, 1024, 0, stream>>>(dirs, rays_per_world, keys, keys + n, vals);
    }
    *order = vals;
    return hipGetLastError();
}

}  // namespace ngx
