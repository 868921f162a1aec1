// Closest-hit tracing for batched worlds on CDNA4: depth / normal / object-ID images, ray queries, and the
// pose backward pass.
//
// The reference walks each ray alone with a private stack of 128 + 64 entries (src/mesh_dev.h:113-159),
// which on a 64-wide wavefront means 64 diverging walks and stacks spilled to scratch. Here a wavefront
// walks as a packet: it holds 64 rays of one world (an 8x8 pixel tile, or 64 consecutive rays of a
// query), fetches each node once for the whole wave through the scalar cache, and every lane tests the
// node's boxes with its own ray and its own closest distance so far. A child is entered when any lane's
// ray enters its box; the lanes that do not are masked out of it. The walk's stack is wave-uniform and
// lives one entry per lane of a vector register (written and read with writelane/readlane), so nothing
// spills. Each lane still finds exactly its own closest hit: a lane only tests triangles inside boxes its
// ray enters before its current closest distance, and keeps a hit only when strictly closer
// (mesh_dev.h:98), so the result equals the per-ray walk's except where two surfaces meet the ray at the
// same distance.
#include <mutex>
#include "ngx_kernels.h"

// Partial sums of ngx_pose_backward_depth_noalloc, which allocates nothing: NOALLOC_SCRATCH_FLOATS floats
// (32 MiB) that every GPU the library runs on holds until the process ends.
static __device__ __align__(256) float scratch[ngx::NOALLOC_SCRATCH_FLOATS];

namespace ngx {
namespace {

struct LaneHit {
    int inst;
    int tri;
    float u, v;
};

// Traversal counters of one wavefront, wave-uniform. NoStats compiles them away; WaveStats is only used by the
// instrumented kernel built with NGX_TRACE_STATS.
enum StatIndex { ST_TLAS_NODES, ST_TLAS_LANES, ST_INSTANCES, ST_INSTANCE_LANES, ST_INSTANCE_ROOT_LANES, ST_BLAS_NODES,
                 ST_BLAS_LANES, ST_TRIANGLES, ST_TRIANGLE_LANES, ST_POPS, ST_VALID, ST_COUNT };
// The same type also says whether the walk is an occlusion test (kAnyHit): then a ray's limit drops below zero
// at its first hit, so its lane leaves every box after it and the walk ends as soon as every ray is blocked or
// done; whether a ray is blocked is all that is read (mesh_dev.h:208-227 stops at the first hit too).
// kHandoff says whether the hybrid walk may hand thin subtrees to the shared per-lane walk; the RGB kernel, whose
// registers are already spent on shading, walks without it. kTwoChild says which shared per-lane walk runs: the one
// testing both children of a node a step (depth, IDs, ray queries, the backward), or the one-box-a-step walk of the
// skip layout, which needs fewer registers (the RGB kernel, whose spills the other raises: body_worlds RGB 34.7 ->
// 44.4 ms, where the depth kernel gains 10.1 -> 9.3 ms and the LiDAR query 21.1 -> 16.4 ms).
struct NoStats {
    static constexpr bool kAnyHit = false, kHandoff = true, kTwoChild = true;
    __device__ __forceinline__ void add(int, int) {}
};
struct ShadeRays {
    static constexpr bool kAnyHit = false, kHandoff = false, kTwoChild = false;
    __device__ __forceinline__ void add(int, int) {}
};
struct Occlusion {
    static constexpr bool kAnyHit = true, kHandoff = false, kTwoChild = false;
    __device__ __forceinline__ void add(int, int) {}
};
// Shadow rays of the two-pass picture, whose second pass holds no primary walk: the depth kernel's walk, stopping
// each ray at its first hit (body_worlds RGB + depth + ID 26.9 ms a step with the one-box-a-step walk here, 26.1 so).
struct OcclusionTwoChild {
    static constexpr bool kAnyHit = true, kHandoff = true, kTwoChild = true;
    __device__ __forceinline__ void add(int, int) {}
};
#ifdef NGX_TRACE_STATS
struct WaveStats {
    static constexpr bool kAnyHit = false, kHandoff = true, kTwoChild = true;
    int c[ST_COUNT];
    __device__ __forceinline__ void add(const int k, const int v) { c[k] += v; }
};
#endif

struct Wide {
    float l_lo[3], l_hi[3], r_lo[3], r_hi[3];
    int left, right, lcount, rcount;
};

__device__ __forceinline__ Wide load_wide(const WideNode *nodes, const int index) {
    const float4 *p = reinterpret_cast<const float4 *>(nodes + index);
    const float4 f0 = sload4(p), f1 = sload4(p + 1), f2 = sload4(p + 2);
    const int4 f3 = sload4i(reinterpret_cast<const int4 *>(p + 3));
    Wide w;
    w.l_lo[0] = f0.x; w.l_lo[1] = f0.y; w.l_lo[2] = f0.z; w.l_hi[0] = f0.w;
    w.l_hi[1] = f1.x; w.l_hi[2] = f1.y; w.r_lo[0] = f1.z; w.r_lo[1] = f1.w;
    w.r_lo[2] = f2.x; w.r_hi[0] = f2.y; w.r_hi[1] = f2.z; w.r_hi[2] = f2.w;
    w.left = f3.x; w.right = f3.y; w.lcount = f3.z; w.rcount = f3.w;
    return w;
}

// ---------------------------------------------------------------- batched triangle tests --
//
// Below an instance the packet thins out: a leaf is typically reached by a handful of the wave's rays, so
// testing its triangles one after the other for the whole wave leaves most lanes idle. Instead a leaf
// reached by m rays adds m x count (ray, triangle) pairs to a table in LDS, and once 64 pairs are queued,
// or the instance is finished, every lane tests one pair. A ray's pairs are numbered in the order the walk
// met its triangles, and the ray keeps the closest hit, the lowest pair number among equal distances, with
// one 64-bit LDS minimum per pair (distance bits above pair number; distances are positive, so their bits
// order as the floats do). That is what testing the triangles one by one with a strict "closer than"
// keeps (mesh_dev.h:98, :186-199). A ray's limit is not lowered until its batch is tested, so a few
// more boxes may be entered; that costs time, not correctness.
struct PairLds {
    int pair[64];                  // ray lane | triangle << 6
    unsigned long long best[64];   // per ray lane: distance bits << 32 | pair number
    float2 uv[64];                 // per pair: barycentrics of its hit
    float2 hit_uv[64];             // per ray lane: barycentrics of its closest hit so far
    int hit_inst[64], hit_tri[64]; // per ray lane: instance and triangle of its closest hit so far (-1: none)
    // the per-lane bottom-level walk's shared work: subtrees given away (ray lane | node << 6), and a ring of
    // leaves waiting for their triangle tests (first triangle; ray lane | count << 6)
    int share[128];
    int leaf_first[128], leaf_ray_count[128];
    // the packet walk's two wave-uniform stacks (top level, bottom level): node reference and lane mask
    int st_ref[2][WALK_STACK];
    uint32_t st_lo[2][WALK_STACK], st_hi[2][WALK_STACK];
};

// Orders a wavefront's LDS accesses: a wavefront's LDS operations are performed in order, so a
// compiler barrier is all that is needed between one lane's write and another lane's read.
__device__ __forceinline__ void wave_lds_order() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int rank_in(const uint64_t mask) {
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// A wave-uniform stack of (node reference, lane mask) in the wavefront's LDS: one lane writes a push, every
// lane reads a pop (the same address for all, so it is broadcast) and the value is taken as uniform.
struct LdsStack {
    int *ref;
    uint32_t *lo, *hi;
    int *overflow;
    int sp;
    __device__ __forceinline__ LdsStack(PairLds *L, const int level, int *flag)
        : ref(L->st_ref[level]), lo(L->st_lo[level]), hi(L->st_hi[level]), overflow(flag), sp(0) {}
    // Depth bounds keep sp below WALK_STACK (a top-level tree has fewer than 64 levels below a world's root, and the
    // host refuses bottom-level trees deeper than 60). A tree that breaks them anyway loses the entry that does not
    // fit, and the walk reports it through the scene's overflow flag, which the host turns into an error.
    __device__ __forceinline__ void push(const int r, const uint64_t m) {
        if (sp >= WALK_STACK) {
            if (lane_id() == 0) *overflow = 1;
            return;
        }
        if (lane_id() == 0) {
            ref[sp] = r;
            lo[sp] = (uint32_t)m;
            hi[sp] = (uint32_t)(m >> 32);
        }
        ++sp;
    }
    __device__ __forceinline__ void pop(int &r, uint64_t &m) {
        --sp;
        wave_lds_order();
        r = uniform(ref[sp]);
        m = ((uint64_t)uniform_u(hi[sp]) << 32) | uniform_u(lo[sp]);
    }
};

// Tests the queued pairs of instance inst, whose mesh-space ray is (lo, ld) in each ray's lane.
template <class Stats>
__device__ __forceinline__ void flush_pairs(const TraceScene &S, PairLds *L, int &pairs, const int inst, const float3 lo,
                                            const float3 ld, float &tmax, Stats &stats) {
    if (pairs == 0) return;
    stats.add(ST_TRIANGLES, 1);
    stats.add(ST_TRIANGLE_LANES, pairs);
    const int lane = lane_id();
    L->best[lane] = ((unsigned long long)__float_as_uint(tmax) << 32) | 0xffffffffull;
    wave_lds_order();
    const bool act = lane < pairs;
    const int e = act ? L->pair[lane] : 0;
    const int r = e & 63;
    const float3 po = mk3(__shfl(lo.x, r, 64), __shfl(lo.y, r, 64), __shfl(lo.z, r, 64));
    const float3 pd = mk3(__shfl(ld.x, r, 64), __shfl(ld.y, r, 64), __shfl(ld.z, r, 64));
    const float pt = __shfl(tmax, r, 64);
    if (act) {
        float3 v0, e1, e2;
        load_tri(S.tris, (size_t)(e >> 6), v0, e1, e2);
        float t, bu, bv;
        if (triangle_hit(v0, e1, e2, po, pd, pt, t, bu, bv)) {
            atomicMin(&L->best[r], ((unsigned long long)__float_as_uint(t) << 32) | (unsigned)lane);
            L->uv[lane] = make_float2(bu, bv);
        }
    }
    wave_lds_order();
    const unsigned long long b = L->best[lane];
    const unsigned p = (unsigned)b;
    if (p != 0xffffffffu) {
        tmax = Stats::kAnyHit ? -1.0f : __uint_as_float((unsigned)(b >> 32));
        // the hit is kept in LDS, not registers, until the walk is over
        L->hit_inst[lane] = inst;
        L->hit_tri[lane] = L->pair[p] >> 6;
        L->hit_uv[lane] = L->uv[p];
    }
    wave_lds_order();
    pairs = 0;
}

// Queues the triangles first .. first+count-1 for the rays in mask, testing the queue first whenever it
// would overflow.
template <class Stats>
__device__ __forceinline__ void add_leaf(const TraceScene &S, PairLds *L, int &pairs, const int first, const int count,
                                         const uint64_t mask, const int inst, const float3 lo, const float3 ld, float &tmax,
                                         LaneHit &h, Stats &stats) {
    const int m = popc64(mask);
    const int lane = lane_id();
    const bool in = in_mask(mask);
    const int rank = rank_in(mask);
    const int fit = 64 / m;
    for (int done = 0; done < count;) {
        const int c = min(count - done, fit);
        if (pairs + m * c > S.pair_limit) flush_pairs(S, L, pairs, inst, lo, ld, tmax, stats);
        if (in) {
            const int base = pairs + rank * c;
            for (int k = 0; k < c; ++k) L->pair[base + k] = lane | ((first + done + k) << 6);
        }
        pairs += m * c;
        done += c;
    }
}


// The bottom level of one instance walked lane by lane, for meshes whose triangles are small next to a
// wavefront's tile: there each ray of the packet needs its own leaves, so a shared walk visits the union of
// 64 paths. A lane walks one ray through one subtree of the octant-ordered skip layout (a box test per step;
// a hit descends to the first child for the ray's octant, a miss or a leaf goes to the node's escape for the
// octant), so it keeps no stack. The work is shared: a lane whose remaining range is a whole subtree (its
// node's escape is the range's end) gives the far child's subtree to a shared stack in LDS, while there is
// room, and keeps the near one; idle lanes take the newest given subtrees, with their rays' mesh-space data
// read from the rays' own lanes. So the few rays that enter an instance keep the whole wave busy. Leaves are
// queued, and their triangles are tested in batches of up to 64 (ray, triangle) pairs once 64 leaves or 64
// pairs wait, and at the end; after each batch every walking lane re-reads its ray's closest distance. Every
// node a ray enters is walked by exactly one lane, so each ray's closest hit is its own. (The RGB kernel walks so;
// the other kernels use the two-child walk further below, which shares its work the same way.)
constexpr int SHARE_CAP = SHARE_STACK, LEAF_RING = 128;
// Queued leaves that trigger a batch of triangle tests (at most 64 pairs a batch; fewer leaves wait for more).
#ifndef NGX_LEAF_BATCH
#define NGX_LEAF_BATCH 64
#endif
// In the hybrid walk, a bottom-level subtree entered by at most this many of the packet's rays is walked by the
// shared per-lane walk instead of the packet (0: never).
#ifndef NGX_HANDOFF_LANES
#define NGX_HANDOFF_LANES 16
#endif

template <bool FORCE, class Stats>
__device__ __forceinline__ void test_queued_leaves(const TraceScene &S, PairLds *L, int &head, int &count, const int g,
                                                   const float3 lo, const float3 ld, float &tmax, Stats &stats) {
    const int lane = lane_id();
    while (count > 0) {
        wave_lds_order();
        const int avail = min(count, 64);
        const bool mine = lane < avail;
        const int slot = (head + lane) & (LEAF_RING - 1);
        const int rc = mine ? L->leaf_ray_count[slot] : 0;
        const int first = mine ? L->leaf_first[slot] : 0;
        const int c = rc >> 6;
        const int c0 = __builtin_amdgcn_readlane(c, 0);
        if (c0 > 64) {
            // a leaf of more than 64 triangles (the tree's leaves hold 4, unless a mesh has many coincident
            // triangles): its next 64, the rest stays queued
            const int ray0 = __builtin_amdgcn_readlane(rc, 0) & 63, first0 = __builtin_amdgcn_readlane(first, 0);
            wave_lds_order();
            L->pair[lane] = ray0 | ((first0 + lane) << 6);
            if (lane == 0) {
                L->leaf_first[head] = first0 + 64;
                L->leaf_ray_count[head] = ray0 | ((c0 - 64) << 6);
            }
            int pairs = 64;
            flush_pairs(S, L, pairs, g, lo, ld, tmax, stats);
            continue;
        }
        const int incl = wave_prefix_sum(c);
        const bool take = mine && incl <= 64;
        const int taken = popc64(ballot(take));  // at least the first leaf, which holds 64 triangles at most
        const int pairs_total = __builtin_amdgcn_readlane(incl, taken - 1);
        if (!FORCE && pairs_total < 64 && count < 64) break;
        if (take) {
            const int ray = rc & 63;
            for (int k = 0; k < c; ++k) L->pair[incl - c + k] = ray | ((first + k) << 6);
        }
        int pairs = pairs_total;
        flush_pairs(S, L, pairs, g, lo, ld, tmax, stats);
        head = (head + taken) & (LEAF_RING - 1);
        count -= taken;
    }
}

// Gives the subtree of skip node `node` to the shared stack for every ray in mask (the caller has checked the room).
__device__ __forceinline__ void share_subtree(PairLds *L, int &top, const uint64_t mask, const int node) {
    if (in_mask(mask)) L->share[top + rank_in(mask)] = lane_id() | (node << 6);
    top += popc64(mask);
}

// Walks the `top` subtrees on the shared stack (rays of instance g, whose mesh-space rays are in their own lanes).
template <class Stats>
__device__ void walk_shared(const TraceScene &S, PairLds *L, const int g, int top, const float3 lo, const float3 ld,
                            const float3 inv_ld, float &tmax, Stats &stats) {
    const int lane = lane_id();
    int head = 0, queued = 0;
    int ray = -1, cur = -1, end = -1, oct = 0;
    float3 wlo = lo, wld = ld, winv = inv_ld;
    float wlim = tmax;
    while (true) {
        // idle lanes take the newest given subtrees
        const uint64_t idle = ballot(ray < 0);
        if (top > 0 && idle != 0) {
            const int n = min(popc64(idle), top);
            const int rk = rank_in(idle);
            const bool take = in_mask(idle) && rk < n;
            wave_lds_order();
            int start = 0;
            if (take) {
                const int e = L->share[top - 1 - rk];
                ray = e & 63;
                start = e >> 6;
            }
            top -= n;
            const int src = ray >= 0 ? ray : lane;
            wlo = mk3(__shfl(lo.x, src, 64), __shfl(lo.y, src, 64), __shfl(lo.z, src, 64));
            wld = mk3(__shfl(ld.x, src, 64), __shfl(ld.y, src, 64), __shfl(ld.z, src, 64));
            winv = mk3(__shfl(inv_ld.x, src, 64), __shfl(inv_ld.y, src, 64), __shfl(inv_ld.z, src, 64));
            wlim = __shfl(tmax, src, 64);
            if (take) {
                oct = (wld.x >= 0.0f ? 1 : 0) | (wld.y >= 0.0f ? 2 : 0) | (wld.z >= 0.0f ? 4 : 0);
                cur = start;
                end = S.skip[start].escape[oct];
            }
        }
        const uint64_t busy = ballot(ray >= 0);
        if (busy == 0) break;
        stats.add(ST_BLAS_NODES, 1);
        stats.add(ST_BLAS_LANES, popc64(busy));
        // One step in every walking lane; the decisions are taken as wave masks, and only the node loads are
        // confined to the walking lanes (idle lanes test whatever their registers hold, and their results are masked out).
        float4 f0 = make_float4(unknown_f(), unknown_f(), unknown_f(), unknown_f());
        float4 f1 = make_float4(unknown_f(), unknown_f(), unknown_f(), unknown_f());
        int esc = unknown_i();
        if (in_mask(busy)) {
            // a 32-bit byte offset from the array's base (node indices stay below 2^25)
            const char *np = reinterpret_cast<const char *>(S.skip) + (uint32_t)cur * (uint32_t)sizeof(SkipNode);
            f0 = reinterpret_cast<const float4 *>(np)[0];
            f1 = reinterpret_cast<const float4 *>(np)[1];
            esc = reinterpret_cast<const SkipNode *>(np)->escape[oct];
        }
        const int a = __float_as_int(f0.w), meta = __float_as_int(f1.w);
        float tn;
        const uint64_t hm = busy & ballot(box_hit(f0.x, f0.y, f0.z, f1.x, f1.y, f1.z, wlo, winv, box_limit(wlim), tn));
        const uint64_t lm = hm & ballot(meta > 0);  // leaves entered
        const uint64_t im = hm & ~lm;               // interior nodes entered
        const int code = -meta - 1, axis = (code >> 1) & 3, flip = code & 1;
        const int near = ((oct >> axis) & 1) != flip ? a : a + 1;
        const int far = 2 * a + 1 - near;
        // a lane whose range is this node's subtree (its escape is the range's end) may give the far child away
        const uint64_t gm = im & ballot(esc == end);
        if (in_mask(busy)) cur = in_mask(im) ? near : esc;
        if (gm != 0) {
            const uint64_t given = gm & ballot(rank_in(gm) < SHARE_CAP - top);
            if (in_mask(given)) {
                L->share[top + rank_in(given)] = ray | (far << 6);
                end = far;  // after the near subtree the walk would reach the far one: stop there
            }
            top += popc64(given);
        }
        if (lm != 0) {
            if (in_mask(lm)) {
                const int slot = (head + queued + rank_in(lm)) & (LEAF_RING - 1);
                L->leaf_first[slot] = a;
                L->leaf_ray_count[slot] = ray | (meta << 6);
            }
            queued += popc64(lm);
        }
        if (in_mask(busy) && cur == end) ray = -1;
        if (queued >= NGX_LEAF_BATCH) {
            test_queued_leaves<false>(S, L, head, queued, g, lo, ld, tmax, stats);
            wlim = __shfl(tmax, ray >= 0 ? ray : lane, 64);
        }
    }
    test_queued_leaves<true>(S, L, head, queued, g, lo, ld, tmax, stats);
}

// The shared per-lane walk over the two-child layout the packet walk uses: a step loads one node and tests both of
// its children, as the reference does (mesh_dev.h:133-150), so a ray's chain of dependent node loads is half as long
// as in the walk above, whose arrival at a node is its box test. A work item is a subtree whose box its ray is known to
// enter. A lane entering both children of its node continues with the nearer one and gives the farther to the shared
// stack in LDS, or, when that is full, keeps it on its own stack; an idle lane first resumes its own entries, then
// takes the newest given subtree (with its ray's mesh-space data read from the ray's lane). A lane's own entries lie on
// one path down a tree, one a level at most, and the host refuses bottom-level trees of WALK_STACK levels or more, so
// WALK_STACK entries hold them. Leaves are queued and tested in batches as in the walk above.
template <class Stats>
__device__ void walk_shared_two_child(const TraceScene &S, PairLds *L, const int g, int top, const float3 lo,
                                      const float3 ld, const float3 inv_ld, float &tmax, Stats &stats) {
    const int lane = lane_id();
    int head = 0, queued = 0;
    int ray = -1, cur = -1;
    float3 wlo = lo, winv = inv_ld;
    float wlim = tmax;
    int own[WALK_STACK];
    int own_n = 0;
    while (true) {
        const uint64_t idle = ballot(ray < 0);
        const uint64_t resume = idle & ballot(own_n > 0);
        if (idle != 0 && (top > 0 || resume != 0)) {
            const uint64_t takers = idle & ~resume;
            const int n = min(popc64(takers), top);
            const int rk = rank_in(takers);
            wave_lds_order();
            int e = -1;
            if (in_mask(takers) && rk < n) e = L->share[top - 1 - rk];
            if (in_mask(resume)) e = own[--own_n];
            top -= n;
            if (e >= 0) {
                ray = e & 63;
                cur = e >> 6;
            }
            const int src = ray >= 0 ? ray : lane;
            wlo = mk3(__shfl(lo.x, src, 64), __shfl(lo.y, src, 64), __shfl(lo.z, src, 64));
            winv = mk3(__shfl(inv_ld.x, src, 64), __shfl(inv_ld.y, src, 64), __shfl(inv_ld.z, src, 64));
            wlim = __shfl(tmax, src, 64);
        }
        const uint64_t busy = ballot(ray >= 0);
        if (busy == 0) break;
        stats.add(ST_BLAS_NODES, 1);
        stats.add(ST_BLAS_LANES, popc64(busy));
        // idle lanes test whatever their registers hold, and their results are masked out
        float4 f0 = make_float4(unknown_f(), unknown_f(), unknown_f(), unknown_f());
        float4 f1 = make_float4(unknown_f(), unknown_f(), unknown_f(), unknown_f());
        float4 f2 = make_float4(unknown_f(), unknown_f(), unknown_f(), unknown_f());
        int4 f3 = make_int4(unknown_i(), unknown_i(), unknown_i(), unknown_i());
        if (in_mask(busy)) {
            // a 32-bit byte offset from the array's base (node indices stay below 2^25)
            const float4 *np = reinterpret_cast<const float4 *>(reinterpret_cast<const char *>(S.blas_wide) +
                                                                (uint32_t)cur * (uint32_t)sizeof(WideNode));
            f0 = np[0];
            f1 = np[1];
            f2 = np[2];
            f3 = *reinterpret_cast<const int4 *>(np + 3);
        }
        const float lim = box_limit(wlim);
        float dl, dr;
        const uint64_t hl = busy & ballot(box_hit(f0.x, f0.y, f0.z, f0.w, f1.x, f1.y, wlo, winv, lim, dl));
        const uint64_t hr = busy & ballot(box_hit(f1.z, f1.w, f2.x, f2.y, f2.z, f2.w, wlo, winv, lim, dr));
        const uint64_t lleaf = ballot(f3.x < 0), rleaf = ballot(f3.y < 0);
        const uint64_t ll = hl & lleaf, rl = hr & rleaf;    // leaves entered
        const uint64_t li = hl & ~lleaf, ri = hr & ~rleaf;  // interior children entered
        const int nl = popc64(ll), nr = popc64(rl);
        if (nl + nr != 0) {
            // a step queues at most 128 leaves, so the ring holds them once what waits is tested
            if (queued + nl + nr > LEAF_RING) {
                test_queued_leaves<true>(S, L, head, queued, g, lo, ld, tmax, stats);
                wlim = __shfl(tmax, ray >= 0 ? ray : lane, 64);
            }
            if (in_mask(ll)) {
                const int slot = (head + queued + rank_in(ll)) & (LEAF_RING - 1);
                L->leaf_first[slot] = ~f3.x;
                L->leaf_ray_count[slot] = ray | (f3.z << 6);
            }
            if (in_mask(rl)) {
                const int slot = (head + queued + nl + rank_in(rl)) & (LEAF_RING - 1);
                L->leaf_first[slot] = ~f3.y;
                L->leaf_ray_count[slot] = ray | (f3.w << 6);
            }
            queued += nl + nr;
        }
        const uint64_t both = li & ri;
        const bool left_near = dl <= dr;
        const int near = in_mask(both) ? (left_near ? f3.x : f3.y) : in_mask(li) ? f3.x : in_mask(ri) ? f3.y : -1;
        if (both != 0) {
            const int far = left_near ? f3.y : f3.x;
            // the lowest-ranked lanes fill the shared stack's room; the others keep their far child
            const uint64_t given = both & ballot(rank_in(both) < S.share_cap - top);
            if (in_mask(given)) L->share[top + rank_in(given)] = ray | (far << 6);
            top += popc64(given);
            if (in_mask(both & ~given)) {
                if (own_n < WALK_STACK) own[own_n++] = ray | (far << 6);
                else *S.overflow = 1;
            }
        }
        if (in_mask(busy)) {
            cur = near;
            if (near < 0) ray = -1;
        }
        if (queued >= NGX_LEAF_BATCH) {
            test_queued_leaves<false>(S, L, head, queued, g, lo, ld, tmax, stats);
            wlim = __shfl(tmax, ray >= 0 ? ray : lane, 64);
        }
    }
    test_queued_leaves<true>(S, L, head, queued, g, lo, ld, tmax, stats);
}

// One instance for the lanes in mask: the ray moved into mesh space (mesh_dev.h:164-171, direction not
// renormalized so t is shared), then the mesh's own tree.
// BLAS: how bottom levels are walked: 0 as a packet; 1 lane by lane for meshes of lane_min_tris triangles or more,
// as a packet otherwise; 2 lane by lane for every mesh (the packet code is then not built, which frees registers; a
// mesh of a single leaf is still tested as one batch).
template <int BLAS, class Stats>
__device__ void visit_instance(const TraceScene &S, PairLds *L, const int g, const uint64_t mask, const float3 o,
                               const float3 d, float &tmax, LaneHit &h, Stats &stats) {
    stats.add(ST_INSTANCES, 1);
    stats.add(ST_INSTANCE_LANES, popc64(mask));
    const float4 *rp = reinterpret_cast<const float4 *>(S.inst + g);
    const float4 r0 = sload4(rp), r1 = sload4(rp + 1), r2 = sload4(rp + 2);
    const int4 meta = sload4i(reinterpret_cast<const int4 *>(rp + 3));
    const float3 lo = mk3(dot3(xyz(r0), o) + r0.w, dot3(xyz(r1), o) + r1.w, dot3(xyz(r2), o) + r2.w);
    const float3 ld = mk3(dot3(xyz(r0), d), dot3(xyz(r1), d), dot3(xyz(r2), d));
    const float3 inv_d = safe_inverse(ld);
    const float4 *mp = reinterpret_cast<const float4 *>(S.mesh_roots + meta.x);
    const float4 m0 = sload4(mp), m1 = sload4(mp + 1);
    float root_t;
    const uint64_t root_mask = mask & ballot(box_hit(m0.x, m0.y, m0.z, m1.x, m1.y, m1.z, lo, inv_d, box_limit(tmax), root_t));
    if (root_mask == 0) return;
    stats.add(ST_INSTANCE_ROOT_LANES, popc64(root_mask));
    int shared = 0;  // subtrees given to the shared per-lane walk
    constexpr bool LANES = BLAS != 0;
    const int root_ref = __float_as_int(m0.w);
    if (BLAS == 2 || (BLAS == 1 && meta.z >= S.lane_min_tris)) {
        if constexpr (Stats::kTwoChild) {
            // the two-child layout starts below the root's box; a mesh of one leaf is tested as a batch below
            if (root_ref >= 0) {
                share_subtree(L, shared, root_mask, root_ref);
                walk_shared_two_child(S, L, g, shared, lo, ld, inv_d, tmax, stats);
                return;
            }
        } else {
            share_subtree(L, shared, root_mask, meta.w);
            walk_shared(S, L, g, shared, lo, ld, inv_d, tmax, stats);
            return;
        }
    }
    int pairs = 0;
    if (root_ref < 0) {
        add_leaf(S, L, pairs, ~root_ref, __float_as_int(m1.w), root_mask, g, lo, ld, tmax, h, stats);
        flush_pairs(S, L, pairs, g, lo, ld, tmax, stats);
        return;
    }
    LdsStack st(L, 1, S.overflow);
    int cur = root_ref;
    uint64_t cm = root_mask;
    while (true) {
        const Wide w = load_wide(S.blas_wide, cur);
        stats.add(ST_BLAS_NODES, 1);
        stats.add(ST_BLAS_LANES, popc64(cm));
        // every lane runs the tests (a vector instruction costs the same for 1 lane or 64); the masks keep
        // the lanes of this node
        const float lim = box_limit(tmax);
        float dl, dr;
        const uint64_t ml = cm & ballot(box_hit(w.l_lo[0], w.l_lo[1], w.l_lo[2], w.l_hi[0], w.l_hi[1], w.l_hi[2], lo, inv_d, lim, dl));
        const uint64_t mr = cm & ballot(box_hit(w.r_lo[0], w.r_lo[1], w.r_lo[2], w.r_hi[0], w.r_hi[1], w.r_hi[2], lo, inv_d, lim, dr));
        int next_ref, next_count;
        uint64_t next_mask;
        if (ml != 0 && mr != 0) {
            // a lane is nearer the left child when it enters only the left, or the left no later than the right
            const uint64_t near_left = ml & (~mr | ballot(dl <= dr));
            const bool left_first = 2 * popc64(near_left) >= popc64(ml | mr);
            // a leaf child is pushed as its parent and side, and its run read back when popped
            if (left_first) {
                st.push(w.right >= 0 ? w.right : ~(2 * cur + 1), mr);
                next_ref = w.left; next_count = w.lcount; next_mask = ml;
            } else {
                st.push(w.left >= 0 ? w.left : ~(2 * cur), ml);
                next_ref = w.right; next_count = w.rcount; next_mask = mr;
            }
        } else if (ml != 0) {
            next_ref = w.left; next_count = w.lcount; next_mask = ml;
        } else if (mr != 0) {
            next_ref = w.right; next_count = w.rcount; next_mask = mr;
        } else {
            next_ref = 0; next_count = 0; next_mask = 0;
        }
        // Leaves are queued where they are reached; a subtree entered by few rays is given to the shared per-lane
        // walk (hybrid walk only); then the walk resumes from the stack.
        bool finished = false;
        while (true) {
            if (next_mask != 0 && next_ref >= 0) {
                const int n = popc64(next_mask);
                if (!LANES || !Stats::kHandoff || n > NGX_HANDOFF_LANES || shared + n > SHARE_CAP) break;
                share_subtree(L, shared, next_mask, Stats::kTwoChild ? next_ref : S.blas_wide_node[next_ref]);
            } else if (next_mask != 0) {
                add_leaf(S, L, pairs, ~next_ref, next_count, next_mask, g, lo, ld, tmax, h, stats);
            }
            if (st.sp == 0) {
                finished = true;
                break;
            }
            int r;
            st.pop(r, next_mask);
            if (r >= 0) {
                next_ref = r;
            } else {
                const int parent = (~r) >> 1, side = (~r) & 1;
                const Wide pw = load_wide(S.blas_wide, parent);
                next_ref = side ? pw.right : pw.left;
                next_count = side ? pw.rcount : pw.lcount;
            }
        }
        if (finished) break;
        cur = next_ref;
        cm = next_mask;
    }
    flush_pairs(S, L, pairs, g, lo, ld, tmax, stats);
    if (LANES && shared > 0) {
        if constexpr (Stats::kTwoChild) walk_shared_two_child(S, L, g, shared, lo, ld, inv_d, tmax, stats);
        else walk_shared(S, L, g, shared, lo, ld, inv_d, tmax, stats);
    }
}

// The closest instance hit in one world for the lanes in live (mesh_dev.h:177-205). tmax is each lane's
// distance to beat on entry and its closest distance on return.
template <int BLAS, class Stats>
__device__ void trace_world(const TraceScene &S, PairLds *L, const int world, const float3 o, const float3 d,
                            const uint64_t live, float &tmax, LaneHit &h, Stats &stats) {
    if (live == 0) return;
    const int root = uniform(S.roots[world]);
    if (root < 0) return;
    // the origin in vector registers: for camera rays it is the same in every lane
    const float3 ov = in_vgpr3(o);
    if (root >= S.total - 1) {
        visit_instance<BLAS>(S, L, root - (S.total - 1), live, ov, d, tmax, h, stats);
        return;
    }
    const float3 inv_d = safe_inverse(d);
    // One entry is pushed per level at most. Below a world's root every level fixes at least one more bit
    // of a Morton code (30 bits) or, among equal codes, of the leaves' positions (at most ceil(log2 n) + 1
    // bits over n consecutive positions): fewer than 64 levels for any n < 2^31, so 64 entries do.
    LdsStack st(L, 0, S.overflow);
    int cur = root;
    uint64_t cm = live;
    while (true) {
        const Wide w = load_wide(S.tlas_wide, cur);
        stats.add(ST_TLAS_NODES, 1);
        stats.add(ST_TLAS_LANES, popc64(cm));
        const float lim = box_limit(tmax);
        float dl, dr;
        const uint64_t ml = cm & ballot(box_hit(w.l_lo[0], w.l_lo[1], w.l_lo[2], w.l_hi[0], w.l_hi[1], w.l_hi[2], ov, inv_d, lim, dl));
        const uint64_t mr = cm & ballot(box_hit(w.r_lo[0], w.r_lo[1], w.r_lo[2], w.r_hi[0], w.r_hi[1], w.r_hi[2], ov, inv_d, lim, dr));
        int next_ref;
        uint64_t next_mask;
        if (ml != 0 && mr != 0) {
            const uint64_t near_left = ml & (~mr | ballot(dl <= dr));
            const bool left_first = 2 * popc64(near_left) >= popc64(ml | mr);
            if (left_first) {
                st.push(w.right, mr);
                next_ref = w.left; next_mask = ml;
            } else {
                st.push(w.left, ml);
                next_ref = w.right; next_mask = mr;
            }
        } else if (ml != 0) {
            next_ref = w.left; next_mask = ml;
        } else if (mr != 0) {
            next_ref = w.right; next_mask = mr;
        } else {
            next_ref = 0; next_mask = 0;
        }
        while (next_mask == 0 || next_ref < 0) {
            if (next_mask != 0) visit_instance<BLAS>(S, L, ~next_ref, next_mask, ov, d, tmax, h, stats);
            if (st.sp == 0) return;
            st.pop(next_ref, next_mask);
        }
        cur = next_ref;
        cm = next_mask;
    }
}

// ------------------------------------------------------------------------------ per-lane walk --
//
// Each lane walks its own ray through both levels, in exactly the reference's order (near child first,
// mesh_dev.h:138-150, the far child visited only while its box is entered before the closest hit so far,
// :155), so ties break as the reference breaks them. No stack: a trail of bits per level remembers, for
// every node on the current path, whether its far child is still to visit and which side was near, and the
// walk climbs back through parent pointers. The top and bottom levels share one node step (the operands
// are selected per lane), so lanes working on different instances still advance together; leaves are
// postponed until every lane of the wave has one, then instance entries and triangle runs are done
// together (the "while-while" organisation of Aila and Laine, 2009).
template <class Stats>
__device__ void trace_world_lanes(const TraceScene &S, const int world, const float3 o, const float3 d, const bool live,
                                  float &tmax, LaneHit &h, Stats &stats) {
    constexpr int NODE = 0, LEAF = 1, BACK = 2, DONE = 3;
    const float3 inv_d = safe_inverse(d);
    int state = DONE, level = 0;
    int cur_t = -1, cur_b = -1;
    uint64_t bits_t = 0, side_t = 0, bits_b = 0, side_b = 0;
    int depth_t = 0, depth_b = 0;  // levels below the level's root: the trails hold 64 levels
    int leaf_ref = 0, leaf_count = 0, g = 0;
    float3 lo = o, ld = d, inv_ld = inv_d;
    if (live) {
        // an empty world's root is -1; test that first, since with no instances at all total - 1 is -1 too
        const int root = S.roots[world];
        if (root < 0) {
            state = DONE;
        } else if (root >= S.total - 1) {
            state = LEAF;
            leaf_ref = ~(root - (S.total - 1));
        } else {
            state = NODE;
            cur_t = root;
        }
    }
    while (ballot(state != DONE)) {
        // ---- node steps until every lane waits at a leaf or is done
        while (true) {
            const bool on_node = state == NODE, on_back = state == BACK;
            const uint64_t active = ballot(on_node || on_back);
            if (active == 0) break;
            stats.add(ST_TLAS_NODES, 1);
            stats.add(ST_TLAS_LANES, popc64(active));
            if (on_node || on_back) {
                const bool bottom = level == 1;
                int cur = bottom ? cur_b : cur_t;
                uint64_t bits = bottom ? bits_b : bits_t;
                uint64_t side = bottom ? side_b : side_t;
                int depth = bottom ? depth_b : depth_t;
                const bool pending = on_back && (bits & 1ull);
                if (on_node || pending) {
                    const float4 *np = reinterpret_cast<const float4 *>((bottom ? S.blas_wide : S.tlas_wide) + cur);
                    const float4 f0 = np[0], f1 = np[1], f2 = np[2];
                    const int4 f3 = *reinterpret_cast<const int4 *>(np + 3);
                    const float3 ro = bottom ? lo : o, ri = bottom ? inv_ld : inv_d;
                    const float dl = box_entry(f0.x, f0.y, f0.z, f0.w, f1.x, f1.y, ro, ri, tmax);
                    const float dr = box_entry(f1.z, f1.w, f2.x, f2.y, f2.z, f2.w, ro, ri, tmax);
                    const bool hl = dl < NO_HIT, hr = dr < NO_HIT;
                    int child = 0, child_count = 0;
                    bool go = false;
                    if (on_node) {
                        if (hl && hr) {
                            const bool left_first = dl <= dr;
                            bits |= 1ull;
                            side = left_first ? (side & ~1ull) : (side | 1ull);
                            child = left_first ? f3.x : f3.y;
                            child_count = left_first ? f3.z : f3.w;
                            go = true;
                        } else if (hl || hr) {
                            child = hl ? f3.x : f3.y;
                            child_count = hl ? f3.z : f3.w;
                            go = true;
                        } else {
                            state = BACK;
                        }
                    } else {
                        bits &= ~1ull;
                        const bool far_left = (side & 1ull) != 0;  // the near child was the right one
                        if (far_left ? hl : hr) {
                            child = far_left ? f3.x : f3.y;
                            child_count = far_left ? f3.z : f3.w;
                            go = true;
                        }
                    }
                    if (go) {
                        if (child >= 0) {
                            cur = child;
                            bits <<= 1;
                            side <<= 1;
                            // a 65th level would push the root's bit out of the trail
                            if (++depth >= 64) *S.overflow = 1;
                            state = NODE;
                        } else {
                            state = LEAF;
                            leaf_ref = child;
                            leaf_count = child_count;
                        }
                    }
                } else {
                    // climb one level; a level whose root is finished hands back to the level above
                    const int parent = cur < 0 ? -1 : (bottom ? S.blas_parent : S.tlas_parent)[cur];
                    if (parent < 0) {
                        if (bottom) level = 0;  // the instance is done: resume the top level where it left off
                        else state = DONE;
                    } else {
                        cur = parent;
                        bits >>= 1;
                        side >>= 1;
                        --depth;
                    }
                }
                if (bottom) {
                    cur_b = cur;
                    bits_b = bits;
                    side_b = side;
                    depth_b = depth;
                } else {
                    cur_t = cur;
                    bits_t = bits;
                    side_t = side;
                    depth_t = depth;
                }
            }
        }
        // ---- instance entries: the ray moved into mesh space, then the mesh's root box
        const bool enter = state == LEAF && level == 0;
        if (ballot(enter)) {
            stats.add(ST_INSTANCES, 1);
            stats.add(ST_INSTANCE_LANES, popc64(ballot(enter)));
            if (enter) {
                g = ~leaf_ref;
                const float4 *rp = reinterpret_cast<const float4 *>(S.inst + g);
                const float4 r0 = rp[0], r1 = rp[1], r2 = rp[2];
                const int mesh = reinterpret_cast<const int4 *>(rp + 3)->x;
                lo = mk3(dot3(xyz(r0), o) + r0.w, dot3(xyz(r1), o) + r1.w, dot3(xyz(r2), o) + r2.w);
                ld = mk3(dot3(xyz(r0), d), dot3(xyz(r1), d), dot3(xyz(r2), d));
                inv_ld = safe_inverse(ld);
                const float4 *mp = reinterpret_cast<const float4 *>(S.mesh_roots + mesh);
                const float4 m0 = mp[0], m1 = mp[1];
                if (box_entry(m0.x, m0.y, m0.z, m1.x, m1.y, m1.z, lo, inv_ld, tmax) < NO_HIT) {
                    level = 1;
                    bits_b = side_b = 0;
                    depth_b = 0;
                    const int root_ref = __float_as_int(m0.w);
                    if (root_ref >= 0) {
                        cur_b = root_ref;
                        state = NODE;
                    } else {
                        cur_b = -1;
                        leaf_ref = root_ref;
                        leaf_count = __float_as_int(m1.w);
                    }
                } else {
                    state = BACK;
                }
            }
        }
        // ---- triangle runs
        const bool tris = state == LEAF && level == 1;
        if (ballot(tris)) {
            stats.add(ST_TRIANGLES, 1);
            stats.add(ST_TRIANGLE_LANES, popc64(ballot(tris)));
            if (tris) {
                const int first = ~leaf_ref;
                for (int j = first; j < first + leaf_count; ++j) {
                    float3 v0, e1, e2;
                    load_tri(S.tris, (size_t)j, v0, e1, e2);
                    float t, bu, bv;
                    if (triangle_hit(v0, e1, e2, lo, ld, tmax, t, bu, bv)) {
                        tmax = Stats::kAnyHit ? -1.0f : t;
                        h.inst = g;
                        h.tri = j;
                        h.u = bu;
                        h.v = bv;
                    }
                }
                state = BACK;
            }
        }
    }
}

// Spheres, then the floor, then the meshes with the best so far as their limit (render_core.h:78-112).
// Returns the object, or OBJECT_NONE; t and h describe the hit.
// The walks: per-lane (both levels lane by lane), packet (both levels as a packet), and hybrid (the packet
// walk with fine meshes' bottom levels walked lane by lane).
// The walks: per-lane (both levels lane by lane), packet (both levels as a packet), hybrid (the packet walk with fine
// meshes' bottom levels walked lane by lane), and shared (the packet top level, every bottom level lane by lane: what
// the hybrid walk does when every mesh is fine, built without the packet bottom level; ray queries only).
enum Walk { WALK_LANES = 0, WALK_PACKET = 1, WALK_HYBRID = 2, WALK_SHARED = 3 };
template <int WALK>
constexpr int blas_mode() { return WALK == WALK_SHARED ? 2 : WALK == WALK_HYBRID ? 1 : 0; }

template <int WALK, class Stats>
__device__ __forceinline__ int scene_hit(const TraceScene &S, PairLds *L, const int world, const float3 o, const float3 d,
                                         const uint64_t live, float &best, LaneHit &h, Stats &stats, const float start = 1e30f) {
    best = start;
    int best_sphere = -1;
    for (int i = 0; i < S.sphere_count; ++i) {
        const float4 sp = sload4(S.spheres + i);
        const float3 L = sub3(xyz(sp), o);
        const float tca = dot3(L, d);
        const float r2 = sp.w * sp.w;
        const float d2 = dot3(L, L) - tca * tca;
        if (d2 > r2) continue;
        const float thc = sqrtf(r2 - d2);
        float t = tca - thc;
        if (t < 0.0f) t = tca + thc;
        if (t >= 0.0f && t < best) {
            best = t;
            best_sphere = i;
        }
    }
    bool hit_floor = false;
    if (fabsf(d.y) > 1e-3f) {
        const float dist = (S.floor_y - o.y) / d.y;
        if (dist > 0.0f && dist < best) {
            const float3 pt = add3(o, mul3(d, dist));
            if (fabsf(pt.x) < S.floor_hx && pt.z > S.floor_zmin && pt.z < S.floor_zmax) {
                best = dist;
                hit_floor = true;
            }
        }
    }
    h.inst = -1;
    h.tri = -1;
    if (WALK != WALK_LANES) {
        const int lane = lane_id();
        L->hit_inst[lane] = -1;
        wave_lds_order();
        trace_world<blas_mode<WALK>()>(S, L, world, o, d, live, best, h, stats);
        wave_lds_order();
        h.inst = L->hit_inst[lane];
        if (h.inst >= 0) {
            h.tri = L->hit_tri[lane];
            const float2 uv = L->hit_uv[lane];
            h.u = uv.x;
            h.v = uv.y;
        }
    } else trace_world_lanes(S, world, o, d, ((live >> lane_id()) & 1ull) != 0, best, h, stats);
    if (h.inst >= 0) return OBJECT_INSTANCE;
    if (hit_floor) return OBJECT_FLOOR;
    if (best_sphere >= 0) return OBJECT_SPHERE + best_sphere;
    return OBJECT_NONE;
}

// Normal and object ID of a hit (render_core.h:119-141, mesh_dev.h:238-262).
__device__ __forceinline__ void surface(const TraceScene &S, const int kind, const float3 o, const float3 d, const float t,
                                        const LaneHit &h, float3 &normal, int &object) {
    if (kind == OBJECT_INSTANCE) {
        const InstanceRec &rec = S.inst[h.inst];
        const float3 r0 = xyz(rec.inv[0]), r1 = xyz(rec.inv[1]), r2 = xyz(rec.inv[2]);
        float3 v0, e1, e2;
        load_tri(S.tris, (size_t)h.tri, v0, e1, e2);
        const float3 nl = cross3(e1, e2);
        float3 geometric = normalize3(add3(add3(mul3(r0, nl.x), mul3(r1, nl.y)), mul3(r2, nl.z)));
        if (dot3(geometric, d) > 0.0f) geometric = neg3(geometric);
        normal = geometric;
        if (S.attrs != nullptr) {
            const float4 *ap = S.attrs + 4 * (size_t)h.tri;
            const float4 a0 = ap[0], a1 = ap[1], a2 = ap[2], a3 = ap[3];
            if (__float_as_int(a3.w) & 1) {
                const float w0 = 1.0f - h.u - h.v;
                const float3 n = add3(add3(mul3(xyz(a0), w0), mul3(xyz(a1), h.u)), mul3(xyz(a2), h.v));
                if (dot3(n, n) > 1e-12f) {
                    const float3 smooth = normalize3(add3(add3(mul3(r0, n.x), mul3(r1, n.y)), mul3(r2, n.z)));
                    normal = dot3(smooth, geometric) < 0.0f ? neg3(smooth) : smooth;
                }
            }
        }
        object = OBJECT_INSTANCE + rec.tmpl;
    } else if (kind == OBJECT_FLOOR) {
        normal = mk3(0.0f, 1.0f, 0.0f);
        object = OBJECT_FLOOR;
    } else {
        const float4 sp = S.spheres[kind - OBJECT_SPHERE];
        normal = normalize3(sub3(add3(o, mul3(d, t)), xyz(sp)));
        object = kind;
    }
}

// The object ID surface() gives, without the normal: what a frame of depth and IDs alone needs.
__device__ __forceinline__ int object_of(const TraceScene &S, const int kind, const LaneHit &h) {
    return kind == OBJECT_INSTANCE ? OBJECT_INSTANCE + S.inst[h.inst].tmpl : kind;
}

struct Cam {
    float3 pos, right, up, fwd;
};

// camera_from_matrix (render_core.h:264-273): the matrix's columns, renormalized.
__device__ __forceinline__ Cam load_camera(const float *m12) {
    const float4 *p = reinterpret_cast<const float4 *>(m12);
    const float4 a = sload4(p), b = sload4(p + 1), c = sload4(p + 2);
    Cam cam;
    cam.pos = mk3(a.w, b.w, c.w);
    cam.right = normalize3(mk3(a.x, b.x, c.x));
    cam.up = normalize3(mk3(a.y, b.y, c.y));
    cam.fwd = normalize3(mk3(-a.z, -b.z, -c.z));
    return cam;
}

// camera_ray (device_scene.h:114-119) through a pixel's centre.
__device__ __forceinline__ float3 pixel_ray(const Cam &c, const float fov_scale, const float aspect, const int x, const int y,
                                            const int width, const int height) {
    const float px = (float)x + 0.5f, py = (float)y + 0.5f;
    const float sx = (2.0f * px / (float)width - 1.0f) * fov_scale * aspect;
    const float sy = -(2.0f * py / (float)height - 1.0f) * fov_scale;
    return normalize3(add3(add3(mul3(c.right, sx), mul3(c.up, sy)), c.fwd));
}

// The tile of a wave: tiles are numbered view-major, and with xcd_remap consecutive workgroups, which
// the dispatcher deals round-robin over the eight XCDs, are renumbered so each XCD works through one
// contiguous range of tiles (the same views, so the same worlds' trees in its L2).
__device__ __forceinline__ long long wave_tile(const int xcd_remap, const long long waves_total) {
    const long long blocks = (long long)gridDim.x;
    long long b = blockIdx.x;
    if (xcd_remap && blocks >= 8) {
        const long long per = blocks / 8, rem = blocks % 8, xcd = b % 8, k = b / 8;
        // XCD x receives blocks x, x+8, ...: give it the x-th contiguous range
        b = xcd < rem ? xcd * (per + 1) + k : rem * (per + 1) + (xcd - rem) * per + k;
    }
    (void)waves_total;
    return b * (blockDim.x >> 6) + uniform(threadIdx.x >> 6);
}

// LDS of the packet walk's triangle batches: one PairLds per wavefront of a 256-thread workgroup.
#define NGX_PAIR_LDS(name) __shared__ PairLds name##_all[4]; PairLds *name = WALK != WALK_LANES ? &name##_all[threadIdx.x >> 6] : nullptr

// One 64-pixel tile of one view: camera ray per lane, closest hit, depth / normal / object ID.
template <int WALK>
__device__ __forceinline__ void render_tile(const RenderArgs &A, PairLds *L, const long long tile) {
    const int lane = lane_id();
    const int view = (int)(tile / A.tiles_per_view);
    const int t = (int)(tile % A.tiles_per_view);
    const int tx = t % A.tiles_x, ty = t / A.tiles_x;
    const int x = tx * A.tile_w + lane % A.tile_w;
    const int y = ty * A.tile_h + lane / A.tile_w;
    const bool valid = x < A.width && y < A.height;
    const int world = view / A.cameras_per_world;
    const Cam cam = load_camera(A.cams + 12 * (size_t)view);
    const float3 d = pixel_ray(cam, A.fov_scale, A.aspect, x, y, A.width, A.height);
    float best;
    LaneHit h;
    NoStats stats;
    const int kind = scene_hit<WALK>(A.s, L, world, cam.pos, d, ballot(valid), best, h, stats);
    if (!valid) return;
    const size_t idx = ((size_t)view * A.height + y) * A.width + x;
    const bool found = kind != OBJECT_NONE;
    if (A.depth) A.depth[idx] = found ? best * dot3(d, cam.fwd) : __int_as_float(0x7f800000);
    if (A.normal) {
        float3 normal = mk3(0.0f, 0.0f, 0.0f);
        int object = OBJECT_NONE;
        if (found) surface(A.s, kind, cam.pos, d, best, h, normal, object);
        A.normal[3 * idx] = normal.x;
        A.normal[3 * idx + 1] = normal.y;
        A.normal[3 * idx + 2] = normal.z;
        if (A.object_id) A.object_id[idx] = object;
    } else if (A.object_id) {
        A.object_id[idx] = found ? object_of(A.s, kind, h) : OBJECT_NONE;
    }
}

// BLOCKS: workgroups per compute unit the kernel is built for (their registers follow: 7 gives at most 72 VGPRs,
// with more spilled; body_worlds 10.49 ms at 6, 10.20 ms at 7). The packet and hybrid walks are built at 5, 6 and 7
// and the autotuner picks one; the per-lane walk at 1.

// One tile per wavefront, tiles dealt by the dispatcher.
template <int WALK, int BLOCKS>
__global__ __launch_bounds__(256, BLOCKS) void render_kernel(const RenderArgs A) {
    NGX_PAIR_LDS(L);
    const long long tile = wave_tile(A.xcd_remap, A.total_tiles);
    if (tile >= A.total_tiles) return;
    render_tile<WALK>(A, L, tile);
}

// Persistent: a fixed grid whose wavefronts take tiles until none are left. The tiles are cut into eight
// contiguous ranges, one per XCD (the dispatcher deals workgroups round-robin over the XCDs, so workgroup b
// runs on XCD b mod 8), so the worlds of a range stay in one XCD's L2; each range is cut into COUNTER_SLICES
// slices with a counter each on its own cache line. A wavefront starts at the slice its index within the XCD
// picks, takes PERSISTENT_CHUNK (1) tile per counter increment until the slice is used up, then
// moves on to the XCD's next slice; it leaves once it has found every slice of its range used up, so every
// wavefront ends after at most COUNTER_SLICES empty increments. The grid is a multiple of 8 workgroups, so every
// XCD's range has wavefronts to take it.
#ifndef NGX_PERSISTENT_CHUNK
#define NGX_PERSISTENT_CHUNK 1
#endif
constexpr int PERSISTENT_CHUNK = NGX_PERSISTENT_CHUNK;

template <class Work>
__device__ __forceinline__ void take_tiles(int *counters, const long long total, Work &&work) {
    const int xcd = blockIdx.x & 7;
    const long long first = total * xcd / 8, count = total * (xcd + 1) / 8 - first;
    const int home = uniform((int)((blockIdx.x >> 3) * (blockDim.x >> 6) + (threadIdx.x >> 6)) % COUNTER_SLICES);
    for (int i = 0; i < COUNTER_SLICES; ++i) {
        const int s = (home + i) % COUNTER_SLICES;
        const long long lo = count * s / COUNTER_SLICES, n = count * (s + 1) / COUNTER_SLICES - lo;
        int *c = counters + (xcd * COUNTER_SLICES + s) * COUNTER_STRIDE;
        if (i > 0) {
            // a slice another wavefront has used up is passed over with a load rather than an increment, so the
            // end of the launch costs each wavefront one increment instead of one per slice
            int seen = 0;
            if (lane_id() == 0) seen = __hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (__builtin_amdgcn_readlane(seen, 0) >= n) continue;
        }
        while (true) {
            int t = 0;
            if (lane_id() == 0) t = atomicAdd(c, PERSISTENT_CHUNK);
            t = __builtin_amdgcn_readlane(t, 0);
            if (t >= n) break;
            const int end = (int)min((long long)t + PERSISTENT_CHUNK, n);
            for (int k = t; k < end; ++k) work(first + lo + k);
        }
    }
}

template <int WALK, int BLOCKS>
__global__ __launch_bounds__(256, BLOCKS) void render_persistent_kernel(const RenderArgs A) {
    NGX_PAIR_LDS(L);
    take_tiles(A.tile_counters, A.total_tiles, [&](const long long tile) { render_tile<WALK>(A, L, tile); });
}

__global__ void zero_counters_kernel(int *c) {
    if (threadIdx.x < 8 * COUNTER_SLICES) c[threadIdx.x * COUNTER_STRIDE] = 0;
}

#ifdef NGX_TRACE_STATS
// The render kernel's traversal with its counters: each wave writes ST_COUNT ints to its own slot, lane k
// storing counter k, so no atomics are involved.
static_assert(ST_COUNT == TRACE_STAT_COUNT && ST_COUNT <= 64, "one counter a lane, as many as the host reads");
template <int WALK>
__global__ __launch_bounds__(256) void render_stats_kernel(const RenderArgs A, int *out) {
    NGX_PAIR_LDS(L);
    const long long tile = wave_tile(A.xcd_remap, A.total_tiles);
    if (tile >= A.total_tiles) return;
    const int lane = lane_id();
    const int view = (int)(tile / A.tiles_per_view);
    const int t = (int)(tile % A.tiles_per_view);
    const int tx = t % A.tiles_x, ty = t / A.tiles_x;
    const int x = tx * A.tile_w + lane % A.tile_w;
    const int y = ty * A.tile_h + lane / A.tile_w;
    const bool valid = x < A.width && y < A.height;
    const int world = view / A.cameras_per_world;
    const Cam cam = load_camera(A.cams + 12 * (size_t)view);
    const float3 d = pixel_ray(cam, A.fov_scale, A.aspect, x, y, A.width, A.height);
    float best;
    LaneHit h;
    WaveStats stats;
    for (int k = 0; k < ST_COUNT; ++k) stats.c[k] = 0;
    stats.add(ST_VALID, popc64(ballot(valid)));
    scene_hit<WALK>(A.s, L, world, cam.pos, d, ballot(valid), best, h, stats);
    int mine = 0;
    for (int k = 0; k < ST_COUNT; ++k) mine = lane == k ? stats.c[k] : mine;
    if (lane < ST_COUNT) out[(size_t)tile * ST_COUNT + lane] = mine;
}
#endif

// ------------------------------------------------------------------------------ RGB picture --
//
// The classic picture (render_core.h:146-283): per pixel, spp jittered camera rays; per ray, up to
// MAX_BOUNCES mirror bounces; at each hit, direct light from every sphere light sampled at a random point of
// the light (so shadows are soft), tested with a shadow ray, plus an ambient term; rays that leave the scene
// see the sky gradient. The random stream is the reference's: a PCG hash seeded by the pixel and the frame,
// drawn in the same order. A wavefront traces its 64 rays together with the walk of the depth image, so its
// primary rays are a packet; after a bounce the lanes' rays go their own ways and the packet walk still
// serves them (each lane tests with its own ray).
constexpr int MAX_BOUNCES = 4;
constexpr float SURFACE_EPS = 1e-3f;

__device__ __forceinline__ uint32_t pcg_hash(const uint32_t x) {
    const uint32_t state = x * 747796405u + 2891336453u;
    const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
__device__ __forceinline__ float rand01(uint32_t &state) {
    state = pcg_hash(state);
    return (float)state * 2.3283064365386963e-10f;
}
__device__ __forceinline__ float saturate(const float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }
__device__ __forceinline__ float3 mul3v(const float3 a, const float3 b) { return mk3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ float3 reflect3(const float3 i, const float3 n) { return sub3(i, mul3(n, 2.0f * dot3(i, n))); }
__device__ __forceinline__ float3 offset_origin(const float3 p, const float3 n, const float3 dir) {
    return add3(p, mul3(n, dot3(dir, n) < 0.0f ? -SURFACE_EPS : SURFACE_EPS));
}
__device__ __forceinline__ float3 sky(const ShadeScene &E, const float3 dir) {
    const float t = saturate(dot3(dir, E.sky_up) * 0.5f + 0.5f);
    return mul3(add3(E.sky_color_down, mul3(sub3(E.sky_color_up, E.sky_color_down), t)), E.sky_intensity);
}
// geom_dev.h:141-152: a uniform direction, then a point on the light's sphere. The two random numbers are drawn
// whatever the radius; a point light (radius 0, the same for every lane) is its centre, which is what adding the
// finite unit vector times 0 gives, so its sine and cosine are not evaluated.
__device__ __forceinline__ float3 sample_sphere_light(const float3 c, const float radius, uint32_t &rng) {
    const float u = rand01(rng), v = rand01(rng);
    if (radius == 0.0f) return c;
    const float z = 1.0f - 2.0f * u;
    const float rad = sqrtf(fmaxf(0.0f, 1.0f - z * z));
    const float phi = 6.28318530718f * v;
    return add3(c, mul3(mk3(rad * cosf(phi), rad * sinf(phi), z), radius));
}
// device_scene.h:98-111
__device__ __forceinline__ float light_factor(const float4 q, const float4 dir, const float3 towards, const float d) {
    float f = 1.0f;
    if (q.y > 0.0f) {
        const float r = d / q.y;
        f = 1.0f / (1.0f + r * r);
    }
    if (q.z > -1.5f) {
        const float c = -(dir.x * towards.x + dir.y * towards.y + dir.z * towards.z);
        const float t = fminf(fmaxf((c - q.z) / fmaxf(q.w - q.z, 1e-6f), 0.0f), 1.0f);
        f *= t * t * (3.0f - 2.0f * t);
    }
    return f;
}

// Whether anything lies between o and o + max_dist d, for the lanes in live (render_core.h:146-160).
template <int WALK, class Walk = Occlusion>
__device__ __forceinline__ bool occluded(const TraceScene &S, PairLds *L, const int world, const float3 o, const float3 d,
                                         const uint64_t live, const float max_dist) {
    const bool mine = in_mask(live);
    bool blocked = false;
    for (int i = 0; i < S.sphere_count; ++i) {
        const float4 sp = sload4(S.spheres + i);
        const float3 Lc = sub3(xyz(sp), o);
        const float tca = dot3(Lc, d);
        const float r2 = sp.w * sp.w;
        const float d2 = dot3(Lc, Lc) - tca * tca;
        if (d2 > r2) continue;
        const float thc = sqrtf(r2 - d2);
        float t = tca - thc;
        if (t < 0.0f) t = tca + thc;
        if (t >= 0.0f && t < max_dist) blocked = true;
    }
    if (fabsf(d.y) > 1e-3f) {
        const float dist = (S.floor_y - o.y) / d.y;
        if (dist > 0.0f && dist < max_dist) {
            const float3 pt = add3(o, mul3(d, dist));
            if (fabsf(pt.x) < S.floor_hx && pt.z > S.floor_zmin && pt.z < S.floor_zmax) blocked = true;
        }
    }
    // the walk limited to max_dist, stopping each ray at its first hit: it finds a hit exactly when the
    // reference's any-hit walk does (mesh_dev.h:208-227)
    const uint64_t rest = ballot(mine && !blocked);
    if (rest != 0) {
        float best = max_dist;
        LaneHit h;
        Walk stats;
        if (WALK != WALK_LANES) {
            const int lane = lane_id();
            L->hit_inst[lane] = -1;
            wave_lds_order();
            trace_world<blas_mode<WALK>()>(S, L, world, o, d, rest, best, h, stats);
            wave_lds_order();
            if (in_mask(rest) && L->hit_inst[lane] >= 0) blocked = true;
        } else {
            h.inst = -1;
            trace_world_lanes(S, world, o, d, in_mask(rest), best, h, stats);
            if (in_mask(rest) && h.inst >= 0) blocked = true;
        }
    }
    return mine && blocked;
}

__device__ __forceinline__ float3 camera_ray_at(const Cam &c, const float fov_scale, const float aspect, const float px, const float py,
                                                const int width, const int height) {
    const float sx = (2.0f * px / (float)width - 1.0f) * fov_scale * aspect;
    const float sy = -(2.0f * py / (float)height - 1.0f) * fov_scale;
    return normalize3(add3(add3(mul3(c.right, sx), mul3(c.up, sy)), c.fwd));
}

// srgb_channel (geom_dev.h:169-172). The reference's GPU path takes x^(1/2.4) with a reduced-precision pow
// (geom_dev.h:87-94); here it is 2^(log2(x) / 2.4) through the hardware base-2 logarithm and exponential, a few ulp
// from powf, so the byte it rounds to differs from powf's only for values within about 2e-4 of a rounding boundary
// (the 8-body falling-bodies scene's RGB kernel: 3.45 ms with powf, 3.08 ms so)
__device__ __forceinline__ float srgb_channel(const float v) {
    const float x = saturate(v);
    return x <= 0.0031308f ? 12.92f * x : 1.055f * exp2f(__log2f(x) * (1.0f / 2.4f)) - 0.055f;
}

// A wavefront's pixels: an 8x8 (or tile_w x 64 / tile_w) tile of one view.
struct TilePixel {
    int view, x, y, world;
    bool valid;
    size_t index;  // into [views][H][W]
};
__device__ __forceinline__ TilePixel tile_pixel(const RenderArgs &A, const long long tile) {
    const int lane = lane_id();
    TilePixel p;
    p.view = (int)(tile / A.tiles_per_view);
    const int tt = (int)(tile % A.tiles_per_view);
    p.x = (tt % A.tiles_x) * A.tile_w + lane % A.tile_w;
    p.y = (tt / A.tiles_x) * A.tile_h + lane / A.tile_w;
    p.valid = p.x < A.width && p.y < A.height;
    p.world = p.view / A.cameras_per_world;
    p.index = ((size_t)p.view * A.height + p.y) * A.width + p.x;
    return p;
}
// The pixel's random stream, seeded by the pixel and the frame (geom_dev.h:122-126).
__device__ __forceinline__ uint32_t pixel_stream(const RenderArgs &A, const TilePixel &p, const int frame) {
    return pcg_hash((uint32_t)(p.y * A.width + p.x) ^ pcg_hash((uint32_t)frame + 1u));
}
// The pixel's bytes from the sum of its samples (render_core.h:276-283), red in the low byte.
__device__ __forceinline__ uint32_t rgb_bytes(const RgbArgs &G, const float3 sum) {
    const float3 c = mul3(sum, 1.0f / (float)G.spp);
    const uint32_t r = (uint32_t)(saturate(srgb_channel(c.x)) * 255.0f + 0.5f);
    const uint32_t g = (uint32_t)(saturate(srgb_channel(c.y)) * 255.0f + 0.5f);
    const uint32_t b = (uint32_t)(saturate(srgb_channel(c.z)) * 255.0f + 0.5f);
    return r | g << 8 | b << 16;
}
__device__ __forceinline__ void store_rgb(const RgbArgs &G, const size_t index, const uint32_t bytes) {
    uint8_t *out = G.rgb + 3 * index;
    out[0] = (uint8_t)bytes;
    out[1] = (uint8_t)(bytes >> 8);
    out[2] = (uint8_t)(bytes >> 16);
}
__device__ __forceinline__ void write_rgb(const RgbArgs &G, const size_t index, const float3 sum) {
    store_rgb(G, index, rgb_bytes(G, sum));
}

// A hit's point, surface normal and material records (render_core.h:206-214): the instance's template material,
// the floor's checker tint, or the sphere's material.
__device__ __forceinline__ void hit_material(const RgbArgs &G, const int kind, const float3 o, const float3 d, const float best,
                                             const LaneHit &h, float3 &point, float3 &normal, float4 &m0, float4 &m1) {
    const RenderArgs &A = G.r;
    const ShadeScene &E = G.shade;
    point = add3(o, mul3(d, best));
    int object;
    surface(A.s, kind, o, d, best, h, normal, object);
    if (kind == OBJECT_INSTANCE) {
        const float4 *mp = E.instance_materials + 2 * (size_t)A.s.inst[h.inst].tmpl;
        m0 = mp[0];
        m1 = mp[1];
    } else if (kind == OBJECT_FLOOR) {
        m0 = E.floor_material[0];
        m1 = E.floor_material[1];
        const int cx = (int)floorf(point.x * E.floor_checker_scale), cz = (int)floorf(point.z * E.floor_checker_scale);
        const float3 tint = ((cx + cz) & 1) ? mk3(1.0f, 1.0f, 1.0f) : mk3(1.0f, 0.7f, 0.3f);
        m0.x = tint.x * 0.3f;
        m0.y = tint.y * 0.3f;
        m0.z = tint.z * 0.3f;
    } else {
        const float4 *mp = E.sphere_materials + 2 * (size_t)(kind - OBJECT_SPHERE);
        m0 = mp[0];
        m1 = mp[1];
    }
}

// Light i at a lane's hit (render_core.h:215-233): its sample point drawn from rng, the direction and distance to
// it, and the diffuse (cd) and specular (cs) light it adds if nothing blocks it. The light's terms come first:
// where both are zero (a surface facing away from the light, and no highlight) adding them changes nothing whether
// or not the light is blocked (the sums start at +0 and only grow), so the result says whether the lane needs a
// shadow ray at all.
__device__ __forceinline__ bool light_terms(const ShadeScene &E, const int i, const bool hit, const float3 point, const float3 normal,
                                            const float3 d, const float4 m1, uint32_t &rng, float3 &cd, float3 &cs,
                                            float3 &light_dir, float &light_distance) {
    const float4 lp0 = E.lights[4 * i], lq = E.lights[4 * i + 1], ldir = E.lights[4 * i + 2], lcol = E.lights[4 * i + 3];
    light_dir = mk3(0.0f, 1.0f, 0.0f);
    light_distance = 0.0f;
    float strength = 0.0f;
    if (hit) {
        const float3 lp = sample_sphere_light(xyz(lp0), lq.x, rng);
        const float3 to_light = sub3(lp, point);
        light_distance = sqrtf(dot3(to_light, to_light));
        light_dir = mul3(to_light, rcp_rn(light_distance));
        strength = lp0.w * light_factor(lq, ldir, light_dir, light_distance);
    }
    const bool test = hit && strength > 0.0f;
    cd = mk3(0.0f, 0.0f, 0.0f);
    cs = cd;
    if (test) {
        const float3 lit = mul3(xyz(lcol), strength);
        cd = mul3(lit, fmaxf(0.0f, dot3(light_dir, normal)));
        cs = mul3(lit, powf(fmaxf(0.0f, dot3(reflect3(light_dir, normal), d)), m1.z));
    }
    return test && (cd.x != 0.0f || cd.y != 0.0f || cd.z != 0.0f || cs.x != 0.0f || cs.y != 0.0f || cs.z != 0.0f);
}

// A hit's light from the summed unblocked terms and the ambient term (render_core.h:234-237), added to radiance.
__device__ __forceinline__ void add_direct(const ShadeScene &E, const float4 m0, const float4 m1, const float3 diffuse,
                                           const float3 specular, const float throughput, float3 &radiance) {
    const float3 dc = xyz(m0);
    const float3 direct = add3(mul3v(dc, mul3(diffuse, m0.w)), mul3(specular, m1.x));
    radiance = add3(radiance, mul3(add3(direct, mul3(dc, E.ambient * m0.w)), throughput));
}

// What one traced segment of a path adds (render_core.h:206-240): the lanes in active traced o + t d, and kind
// says what each hit (OBJECT_NONE: nothing, so the sky), best and h where. radiance and throughput accumulate as
// in the reference; active is cleared where the path ends, and elsewhere o and d become the mirror bounce. Every
// lane of the wavefront calls it (the shadow walk is a wavefront's); Shadow is the shadow rays' walk.
template <int WALK, class Shadow>
__device__ __forceinline__ void shade_hit(const RgbArgs &G, PairLds *L, const int world, const int kind, const float best,
                                          const LaneHit &h, bool &active, float3 &o, float3 &d, uint32_t &rng,
                                          float3 &radiance, float &throughput) {
    const RenderArgs &A = G.r;
    const ShadeScene &E = G.shade;
    const bool hit = active && kind != OBJECT_NONE;
    if (active && !hit) {
        radiance = add3(radiance, mul3(sky(E, d), throughput));
        active = false;
    }
    float3 normal = mk3(0.0f, 1.0f, 0.0f), point = o;
    float4 m0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), m1 = m0;
    if (hit) hit_material(G, kind, o, d, best, h, point, normal, m0, m1);
    float3 diffuse = mk3(0.0f, 0.0f, 0.0f), specular = mk3(0.0f, 0.0f, 0.0f);
    for (int i = 0; i < E.light_count; ++i) {
        float3 cd, cs, light_dir;
        float light_distance;
        const bool need = light_terms(E, i, hit, point, normal, d, m1, rng, cd, cs, light_dir, light_distance);
        // a wave with no lane needing a shadow ray skips the walk
        const uint64_t shadow = ballot(need);
        if (!shadow) continue;
        const bool blocked = occluded<WALK, Shadow>(A.s, L, world, offset_origin(point, normal, light_dir), light_dir, shadow,
                                                    light_distance);
        if (need && !blocked) {
            diffuse = add3(diffuse, cd);
            specular = add3(specular, cs);
        }
    }
    if (hit) {
        add_direct(E, m0, m1, diffuse, specular, throughput, radiance);
        throughput *= m1.y;
        if (throughput < MIN_THROUGHPUT) {
            active = false;
        } else {
            d = reflect3(d, normal);
            o = offset_origin(point, normal, d);
        }
    }
}

// Workgroups per compute unit the RGB kernels are compiled for: 4 (16 wavefronts, at most 128 VGPRs) for the
// packet and per-lane walks, 7 (at most 72 VGPRs, many more spills) for the hybrid walk, whose scenes have more
// latency to hide (body_worlds RGB: 25.7 ms at 5, 24.3 at 6, 23.5 at 7, 33.0 at 8; the 8-body falling-bodies scene,
// packet walk, is 3.1 ms at 4 and 4.2 at 6).
// The packet and hybrid walks' RGB kernels are built at 4, 5, 6 and 7 workgroups per compute unit, the per-lane
// walk's at 4; the autotuner picks.
template <int WALK>
__device__ __forceinline__ void rgb_tile(const RgbArgs &G, PairLds *L, const long long tile) {
    const RenderArgs &A = G.r;
    const TilePixel p = tile_pixel(A, tile);
    const Cam cam = load_camera(A.cams + 12 * (size_t)p.view);
    uint32_t rng = pixel_stream(A, p, G.frame);
    float3 sum = mk3(0.0f, 0.0f, 0.0f);
    for (int sample = 0; sample < G.spp; ++sample) {
        const float jx = rand01(rng), jy = rand01(rng);
        float3 o = cam.pos;
        float3 d = camera_ray_at(cam, A.fov_scale, A.aspect, (float)p.x + jx, (float)p.y + jy, A.width, A.height);
        float3 radiance = mk3(0.0f, 0.0f, 0.0f);
        float throughput = 1.0f;
        bool active = p.valid;
        for (int depth = 0; ballot(active); ++depth) {
            float best = 0.0f;
            LaneHit h;
            ShadeRays stats;
            int kind = OBJECT_NONE;
            if (depth <= MAX_BOUNCES) kind = scene_hit<WALK>(A.s, L, p.world, o, d, ballot(active), best, h, stats);
            shade_hit<WALK, Occlusion>(G, L, p.world, kind, best, h, active, o, d, rng, radiance, throughput);
        }
        sum = add3(sum, radiance);
    }
    if (p.valid) write_rgb(G, p.index, sum);
}

template <int WALK, int BLOCKS>
__global__ __launch_bounds__(256, BLOCKS) void rgb_kernel(const RgbArgs G) {
    NGX_PAIR_LDS(L);
    const long long tile = wave_tile(G.r.xcd_remap, G.r.total_tiles);
    if (tile >= G.r.total_tiles) return;
    rgb_tile<WALK>(G, L, tile);
}

// The same tiles from a fixed grid taking them from the per-XCD slice counters, as the persistent depth
// kernel does.
template <int WALK, int BLOCKS>
__global__ __launch_bounds__(256, BLOCKS) void rgb_persistent_kernel(const RgbArgs G) {
    NGX_PAIR_LDS(L);
    take_tiles(G.r.tile_counters, G.r.total_tiles, [&](const long long tile) { rgb_tile<WALK>(G, L, tile); });
}

// The picture in two passes, for one sample a pixel where no surface reflects (every reflectivity below
// MIN_THROUGHPUT, so every path ends at its first segment). The first pass traces the jittered camera rays alone,
// with the depth kernel's walk and register budget, and keeps each pixel's hit (16 bytes, and 8 more of
// barycentrics when a mesh has vertex normals); the second draws the same ray and random numbers again and shades
// the hit as the one-pass kernel does, shadow rays and all, so the bytes are the same.
template <int WALK>
__device__ __forceinline__ void rgb_hit_tile(const RgbArgs &G, PairLds *L, const long long tile) {
    const RenderArgs &A = G.r;
    const TilePixel p = tile_pixel(A, tile);
    const Cam cam = load_camera(A.cams + 12 * (size_t)p.view);
    uint32_t rng = pixel_stream(A, p, G.frame);
    const float jx = rand01(rng), jy = rand01(rng);
    const float3 d = camera_ray_at(cam, A.fov_scale, A.aspect, (float)p.x + jx, (float)p.y + jy, A.width, A.height);
    float best;
    LaneHit h;
    NoStats stats;
    const int kind = scene_hit<WALK>(A.s, L, p.world, cam.pos, d, ballot(p.valid), best, h, stats);
    if (!p.valid) return;
    G.hits[p.index] = make_int4(__float_as_int(best), kind, h.inst, h.tri);
    if (G.hit_uv != nullptr) G.hit_uv[p.index] = kind == OBJECT_INSTANCE ? make_float2(h.u, h.v) : make_float2(0.0f, 0.0f);
}

// The ray, its random numbers and its hit as the first pass traced them.
struct ShadeStart {
    TilePixel p;
    uint32_t rng;
    float3 o, d;
    int kind;
    float best;
    LaneHit h;
};
__device__ __forceinline__ ShadeStart shade_start(const RgbArgs &G, const long long tile) {
    const RenderArgs &A = G.r;
    ShadeStart q;
    q.p = tile_pixel(A, tile);
    const Cam cam = load_camera(A.cams + 12 * (size_t)q.p.view);
    q.rng = pixel_stream(A, q.p, G.frame);
    const float jx = rand01(q.rng), jy = rand01(q.rng);
    q.o = cam.pos;
    q.d = camera_ray_at(cam, A.fov_scale, A.aspect, (float)q.p.x + jx, (float)q.p.y + jy, A.width, A.height);
    q.kind = OBJECT_NONE;
    q.best = 0.0f;
    q.h.inst = -1;
    q.h.tri = -1;
    q.h.u = 0.0f;
    q.h.v = 0.0f;
    if (q.p.valid) {
        const int4 r = G.hits[q.p.index];
        q.best = __int_as_float(r.x);
        q.kind = r.y;
        q.h.inst = r.z;
        q.h.tri = r.w;
        if (G.hit_uv != nullptr) {
            const float2 uv = G.hit_uv[q.p.index];
            q.h.u = uv.x;
            q.h.v = uv.y;
        }
    }
    return q;
}

template <int WALK>
__device__ __forceinline__ void rgb_shade_tile(const RgbArgs &G, PairLds *L, const long long tile) {
    ShadeStart q = shade_start(G, tile);
    float3 radiance = mk3(0.0f, 0.0f, 0.0f);
    float throughput = 1.0f;
    bool active = q.p.valid;
    shade_hit<WALK, OcclusionTwoChild>(G, L, q.p.world, q.kind, q.best, q.h, active, q.o, q.d, q.rng, radiance, throughput);
    if (q.p.valid) write_rgb(G, q.p.index, add3(mk3(0.0f, 0.0f, 0.0f), radiance));
}

// The same pass for a scene of one light. With one light and one segment a pixel takes one of two colours: the
// one with the light's terms added, if its shadow ray is not blocked, and the one without. Both are worked out
// (as shade_hit sums them) and encoded before the shadow walk, so two words, not the shading state, are what a
// lane holds across it, and the kernel fits in fewer registers.
template <int WALK>
__device__ __forceinline__ void rgb_shade_one_light_tile(const RgbArgs &G, PairLds *L, const long long tile) {
    const ShadeScene &E = G.shade;
    const ShadeStart q = shade_start(G, tile);
    const float throughput = 1.0f;
    const bool hit = q.p.valid && q.kind != OBJECT_NONE;
    float3 radiance = mk3(0.0f, 0.0f, 0.0f);
    if (q.p.valid && !hit) radiance = add3(radiance, mul3(sky(E, q.d), throughput));
    float3 normal = mk3(0.0f, 1.0f, 0.0f), point = q.o;
    float4 m0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), m1 = m0;
    if (hit) hit_material(G, q.kind, q.o, q.d, q.best, q.h, point, normal, m0, m1);
    uint32_t rng = q.rng;
    float3 cd, cs, light_dir;
    float light_distance;
    const bool need = light_terms(E, 0, hit, point, normal, q.d, m1, rng, cd, cs, light_dir, light_distance);
    const float3 zero = mk3(0.0f, 0.0f, 0.0f);
    float3 dark = radiance, lit = radiance;
    if (hit) {
        add_direct(E, m0, m1, zero, zero, throughput, dark);
        add_direct(E, m0, m1, add3(zero, cd), add3(zero, cs), throughput, lit);
    }
    const uint32_t dark_bytes = rgb_bytes(G, add3(zero, dark));
    const uint32_t lit_bytes = rgb_bytes(G, add3(zero, lit));
    const uint64_t shadow = ballot(need);
    bool blocked = false;
    if (shadow)
        blocked = occluded<WALK, OcclusionTwoChild>(G.r.s, L, q.p.world, offset_origin(point, normal, light_dir), light_dir, shadow,
                                                    light_distance);
    if (q.p.valid) store_rgb(G, q.p.index, need && !blocked ? lit_bytes : dark_bytes);
}

template <int WALK, int BLOCKS>
__global__ __launch_bounds__(256, BLOCKS) void rgb_hit_kernel(const RgbArgs G) {
    NGX_PAIR_LDS(L);
    take_tiles(G.r.tile_counters, G.r.total_tiles, [&](const long long tile) { rgb_hit_tile<WALK>(G, L, tile); });
}

// ONE_LIGHT: the scene has exactly one light (the kernel built for it holds less across the shadow walk).
template <int WALK, int BLOCKS, bool ONE_LIGHT>
__global__ __launch_bounds__(256, BLOCKS) void rgb_shade_kernel(const RgbArgs G) {
    NGX_PAIR_LDS(L);
    take_tiles(G.r.tile_counters, G.r.total_tiles, [&](const long long tile) {
        if constexpr (ONE_LIGHT) rgb_shade_one_light_tile<WALK>(G, L, tile);
        else rgb_shade_tile<WALK>(G, L, tile);
    });
}

// Ray queries (render_core.h:331-349): 64 consecutive rays a wave (with xcd_remap, each XCD takes a
// contiguous range of the waves, so the rays an XCD holds at once are neighbours in the direction order); a
// wave whose rays fall in several worlds walks each world's rays as its own packet.
// 6 workgroups (24 wavefronts) per compute unit, at most 84 VGPRs: the ray queries wait on memory more than the
// render does, and a sixth wavefront per SIMD hides more of it (lidar_sweeps 22.7 -> 21.5 ms; 7 gives 22.6)
// (built at 5, 6 and 7 for the packet and hybrid walks, the autotuner picking; the per-lane walk's at 6)
template <int WALK, int BLOCKS>
__global__ __launch_bounds__(256, BLOCKS) void ray_kernel(const RayArgs A) {
    NGX_PAIR_LDS(L);
    const long long base = wave_tile(A.xcd_remap, 0) * 64;
    if (base >= A.count) return;
    const int lane = lane_id();
    const long long slot = base + lane;
    const bool valid = slot < A.count;
    const int world = valid ? (int)(slot / A.rays_per_world) : -1;
    // the ray this lane traces: the slot itself, or the world's ray at that place in the direction order
    const long long i = valid && A.order ? (long long)world * A.rays_per_world + A.order[slot] : slot;
    float3 o = mk3(0.0f, 0.0f, 0.0f), d = mk3(0.0f, 0.0f, 1.0f);
    bool ok = false;
    if (valid) {
        o = mk3(A.origins[3 * i], A.origins[3 * i + 1], A.origins[3 * i + 2]);
        d = mk3(A.dirs[3 * i], A.dirs[3 * i + 1], A.dirs[3 * i + 2]);
        const float len2 = dot3(d, d);
        ok = len2 > 1e-30f && len2 < 1e30f;
        if (ok) d = mul3(d, rcp_rn(sqrtf(len2)));
    }
    float best = 1e30f;
    LaneHit h;
    h.inst = -1;
    h.tri = -1;
    int kind = OBJECT_NONE;
    if (WALK != WALK_LANES) {
        uint64_t pending = ballot(ok);
        while (pending != 0) {
            const int w = __builtin_amdgcn_readlane(world, first_lane(pending));
            const uint64_t mine = ballot(ok && world == w);
            float b;
            LaneHit hh;
            NoStats stats;
            const int k = scene_hit<WALK>(A.s, L, w, o, d, mine, b, hh, stats);
            if (in_mask(mine)) {
                best = b;
                h = hh;
                kind = k;
            }
            pending &= ~mine;
        }
    } else {
        // each lane walks its own world
        NoStats stats;
        kind = scene_hit<WALK_LANES>(A.s, L, world, o, d, ballot(ok), best, h, stats);
    }
    if (!valid) return;
    const bool found = ok && kind != OBJECT_NONE && best <= A.max_distance;
    float3 normal = mk3(0.0f, 0.0f, 0.0f);
    int object = OBJECT_NONE;
    if (found) surface(A.s, kind, o, d, best, h, normal, object);
    A.distance[i] = found ? best : __int_as_float(0x7f800000);
    A.normal[3 * i] = normal.x;
    A.normal[3 * i + 1] = normal.y;
    A.normal[3 * i + 2] = normal.z;
    A.object_id[i] = object;
}

// ------------------------------------------------------------------------------ backward --
//
// d(depth)/d(pose) through the fixed hit triangle: a pose change that moves the material point the ray
// hit by dX moves the hit distance by n.dX / n.d, n being the triangle's world normal (any scale). With
// X = Rb y + pb (y the point in its body's frame), dL/dpb = a and dL/dq_unit = a^T (dR/dq_unit) y, where
// a = dL/dt * n / n.d; the projection onto the raw quaternion is applied once per body afterwards.

// The 7 gradient terms of one lane's hit: returns false when the hit does not depend on a body pose.
__device__ __forceinline__ bool hit_pose_terms(const TraceScene &S, const int world, const int bodies_per_world,
                                               const int *body_of_instance, const float *local_rot_scale, const float *body_poses,
                                               const float3 o, const float3 d, const LaneHit &h, const float dl_dt,
                                               int &body, float g[7]) {
    const InstanceRec &rec = S.inst[h.inst];
    body = body_of_instance[rec.tmpl];
    if (body < 0) return false;
    const float3 r0 = xyz(rec.inv[0]), r1 = xyz(rec.inv[1]), r2 = xyz(rec.inv[2]);
    float3 v0, e1, e2;
    load_tri(S.tris, (size_t)h.tri, v0, e1, e2);
    const float3 nl = cross3(e1, e2);
    const float3 n = add3(add3(mul3(r0, nl.x), mul3(r1, nl.y)), mul3(r2, nl.z));
    const float nd = dot3(n, d);
    if (nd == 0.0f) return false;
    const float s = dl_dt / nd;
    const float3 a = mul3(n, s);
    // the hit's point in mesh space, then in the body's frame: y = Rl diag(s) x + pl
    const float3 x = add3(add3(v0, mul3(e1, h.u)), mul3(e2, h.v));
    const float *L = local_rot_scale + 12 * (size_t)rec.tmpl;
    const float3 y = mk3(L[0] * x.x + L[1] * x.y + L[2] * x.z + L[3], L[4] * x.x + L[5] * x.y + L[6] * x.z + L[7],
                         L[8] * x.x + L[9] * x.y + L[10] * x.z + L[11]);
    const float *bp = body_poses + 7 * ((size_t)world * bodies_per_world + body);
    const float q0 = bp[3], q1 = bp[4], q2 = bp[5], q3 = bp[6];
    const float qn = sqrtf(q0 * q0 + q1 * q1 + q2 * q2 + q3 * q3);
    const float qx = q0 / qn, qy = q1 / qn, qz = q2 / qn, qw = q3 / qn;
    g[0] = a.x;
    g[1] = a.y;
    g[2] = a.z;
    // a^T (dR/dq_k) y for the unit quaternion's components, R as in instances.cpp:130-132
    const float ax = a.x, ay = a.y, az = a.z, yx = y.x, yy = y.y, yz = y.z;
    g[3] = ax * (2 * qy * yy + 2 * qz * yz) + ay * (2 * qy * yx - 4 * qx * yy - 2 * qw * yz) +
           az * (2 * qz * yx + 2 * qw * yy - 4 * qx * yz);
    g[4] = ax * (-4 * qy * yx + 2 * qx * yy + 2 * qw * yz) + ay * (2 * qx * yx + 2 * qz * yz) +
           az * (-2 * qw * yx + 2 * qz * yy - 4 * qy * yz);
    g[5] = ax * (-4 * qz * yx - 2 * qw * yy + 2 * qx * yz) + ay * (2 * qw * yx - 4 * qz * yy + 2 * qy * yz) +
           az * (2 * qx * yx + 2 * qy * yy);
    g[6] = ax * (-2 * qz * yy + 2 * qy * yz) + ay * (2 * qz * yx - 2 * qx * yz) + az * (-2 * qy * yx + 2 * qx * yy);
    return true;
}

// Adds each lane's 7 terms into acc[body][7] (this wave's own slice of LDS), one body at a time, with a
// fixed-order wave sum, so the result does not depend on scheduling.
__device__ __forceinline__ void wave_accumulate(const bool has, const int body, const float g[7], float *acc) {
    const int lane = lane_id();
    uint64_t pending = ballot(has);
    while (pending != 0) {
        const int b0 = __builtin_amdgcn_readlane(body, first_lane(pending));
        const bool mine = has && body == b0;
        const uint64_t m = ballot(mine);
        // the seven sums side by side, then one lane adds them in
        float s[7];
        for (int k = 0; k < 7; ++k) s[k] = wave_total(mine ? g[k] : 0.0f);
        if (lane == 0)
            for (int k = 0; k < 7; ++k) acc[7 * b0 + k] += s[k];
        pending &= ~m;
    }
}

constexpr int BWD_WAVES = 4;

// 5 workgroups of 4 wavefronts per compute unit (at most 100 VGPRs; posed_bodies forward + backward 746k -> 794k
// images/s; 6 gives 781k)
// (built at 4, 5 and 6 for the packet and hybrid walks, the autotuner picking; the per-lane walk's at 5)
template <int WALK, int BLOCKS>
__global__ __launch_bounds__(64 * BWD_WAVES, BLOCKS) void depth_backward_kernel(const DepthBackwardArgs A) {
    extern __shared__ float acc_all[];  // [BWD_WAVES][bodies][7]
    NGX_PAIR_LDS(L);
    const int view = blockIdx.x / A.parts, part = blockIdx.x % A.parts;
    const int wave = uniform(threadIdx.x >> 6), lane = lane_id();
    const int B = A.bodies_per_world;
    float *acc = acc_all + (size_t)wave * B * 7;
    for (int i = lane; i < B * 7; i += 64) acc[i] = 0.0f;
    const int world = view / A.cameras_per_world;
    const Cam cam = load_camera(A.cams + 12 * (size_t)view);
    const int tiles_x = (A.width + 7) / 8, tiles_y = (A.height + 7) / 8;
    for (int t = part * BWD_WAVES + wave; t < tiles_x * tiles_y; t += A.parts * BWD_WAVES) {
        const int x = (t % tiles_x) * 8 + (lane & 7), y = (t / tiles_x) * 8 + (lane >> 3);
        const bool valid = x < A.width && y < A.height;
        const size_t idx = ((size_t)view * A.height + (valid ? y : 0)) * A.width + (valid ? x : 0);
        const float3 d = pixel_ray(cam, A.fov_scale, A.aspect, x, y, A.width, A.height);
        const float depth = valid ? A.depth[idx] : 0.0f;
        const float gd = valid ? A.grad_depth[idx] : 0.0f;
        const float cosf_ = dot3(d, cam.fwd);
        // re-trace only rays whose forward hit something, bounded just beyond the forward's distance
        const bool want = valid && gd != 0.0f && depth < __int_as_float(0x7f800000);
        const float bound = want ? (depth / cosf_) * (1.0f + 1e-5f) + 1e-6f : 0.0f;
        float best;
        LaneHit h;
        NoStats stats;
        const int kind = scene_hit<WALK>(A.s, L, world, cam.pos, d, ballot(want), best, h, stats, bound);
        int body = -1;
        float g[7] = {0, 0, 0, 0, 0, 0, 0};
        bool has = false;
        if (want && kind == OBJECT_INSTANCE)
            has = hit_pose_terms(A.s, world, B, A.body_of_instance, A.local_rot_scale, A.body_poses, cam.pos, d, h, gd * cosf_, body, g);
        wave_accumulate(has, body, g, acc);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < B * 7; i += blockDim.x) {
        float s = 0.0f;
        for (int wv = 0; wv < BWD_WAVES; ++wv) s += acc_all[(size_t)wv * B * 7 + i];
        A.partial[((size_t)view * A.parts + part) * B * 7 + i] = s;
    }
}

template <int WALK>
__global__ __launch_bounds__(64 * BWD_WAVES) void range_backward_kernel(const RangeBackwardArgs A) {
    extern __shared__ float acc_all[];
    NGX_PAIR_LDS(L);
    const int world = blockIdx.x / A.chunks, chunk = blockIdx.x % A.chunks;
    const int wave = uniform(threadIdx.x >> 6), lane = lane_id();
    const int B = A.bodies_per_world;
    float *acc = acc_all + (size_t)wave * B * 7;
    for (int i = lane; i < B * 7; i += 64) acc[i] = 0.0f;
    const int R = A.rays_per_world;
    const int per = (R + A.chunks - 1) / A.chunks;
    const int r0 = chunk * per, r1 = min(R, r0 + per);
    for (int r = r0 + wave * 64; r < r1; r += 64 * BWD_WAVES) {
        const int ri = r + lane;
        const bool valid = ri < r1;
        const size_t i = (size_t)world * R + (valid ? ri : r0);
        float3 o = mk3(A.origins[3 * i], A.origins[3 * i + 1], A.origins[3 * i + 2]);
        float3 d = mk3(A.dirs[3 * i], A.dirs[3 * i + 1], A.dirs[3 * i + 2]);
        const float len2 = dot3(d, d);
        const bool ok = valid && len2 > 1e-30f && len2 < 1e30f;
        if (ok) d = mul3(d, rcp_rn(sqrtf(len2)));
        const float dist = valid ? A.distance[i] : 0.0f;
        const float gd = valid ? A.grad_distance[i] : 0.0f;
        const bool want = ok && gd != 0.0f && dist < __int_as_float(0x7f800000);
        const float bound = want ? dist * (1.0f + 1e-5f) + 1e-6f : 0.0f;
        float best;
        LaneHit h;
        NoStats stats;
        const int kind = scene_hit<WALK>(A.s, L, world, o, d, ballot(want), best, h, stats, bound);
        int body = -1;
        float g[7] = {0, 0, 0, 0, 0, 0, 0};
        bool has = false;
        if (want && kind == OBJECT_INSTANCE)
            has = hit_pose_terms(A.s, world, B, A.body_of_instance, A.local_rot_scale, A.body_poses, o, d, h, gd, body, g);
        wave_accumulate(has, body, g, acc);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < B * 7; i += blockDim.x) {
        float s = 0.0f;
        for (int wv = 0; wv < BWD_WAVES; ++wv) s += acc_all[(size_t)wv * B * 7 + i];
        A.partial[((size_t)world * A.chunks + chunk) * B * 7 + i] = s;
    }
}

// Sums the parts of each world in order, then projects the unit-quaternion gradient onto the raw
// quaternion: d unit / d q = (I - u u^T) / |q|.
__global__ void pose_grad_finish_kernel(const PoseGradFinishArgs A) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.worlds * A.bodies_per_world) return;
    const int w = i / A.bodies_per_world, b = i % A.bodies_per_world;
    float g[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int p = 0; p < A.parts; ++p) {
        const float *src = A.partial + (((size_t)w * A.parts + p) * A.bodies_per_world + b) * 7;
        for (int k = 0; k < 7; ++k) g[k] += src[k];
    }
    const float *bp = A.body_poses + 7 * (size_t)i;
    const float q[4] = {bp[3], bp[4], bp[5], bp[6]};
    const float qn = sqrtf(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    const float u[4] = {q[0] / qn, q[1] / qn, q[2] / qn, q[3] / qn};
    const float ug = u[0] * g[3] + u[1] * g[4] + u[2] * g[5] + u[3] * g[6];
    float *out = A.grad + 7 * (size_t)i;
    out[0] = g[0];
    out[1] = g[1];
    out[2] = g[2];
    for (int k = 0; k < 4; ++k) out[3 + k] = (g[3 + k] - u[k] * ug) / qn;
}

}  // namespace

// A kernel variant by walk and workgroups per compute unit: the per-lane walk has one variant (LANE_BLOCKS), the
// packet and hybrid walks one per value of the kernel's list; any other value takes the default.
#define NGX_BY_VARIANT3(walk, blocks, def, lane_blocks, v0, v1, v2, launch) \
    switch (walk) { \
    case WALK_LANES: launch(WALK_LANES, lane_blocks); break; \
    case WALK_HYBRID: \
        switch (blocks ? blocks : def) { case v0: launch(WALK_HYBRID, v0); break; case v1: launch(WALK_HYBRID, v1); break; \
                                         default: launch(WALK_HYBRID, v2); break; } \
        break; \
    default: \
        switch (blocks ? blocks : def) { case v0: launch(WALK_PACKET, v0); break; case v1: launch(WALK_PACKET, v1); break; \
                                         default: launch(WALK_PACKET, v2); break; } \
        break; \
    }

// The persistent grid: as many workgroups as the device holds at once (the kernel's occupancy times the
// compute units), a multiple of 8 so every XCD gets the same number.
// No more workgroups than the tiles need (4 wavefronts each), and at least 8, one per XCD.
static int capped_grid(const int resident, const long long tiles) {
    const long long need = ((tiles + 3) / 4 + 7) / 8 * 8;
    return (int)(need < resident ? need : resident);
}

template <int WALK, int BLOCKS>
static int persistent_grid(const int compute_units, const long long tiles) {
    static const int per_cu = [] {
        int n = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, render_persistent_kernel<WALK, BLOCKS>, 256, 0) != hipSuccess || n < 1) n = 1;
        return n;
    }();
    return capped_grid((per_cu * compute_units + 7) / 8 * 8, tiles);
}

hipError_t launch_render(const RenderArgs &a, hipStream_t stream, const int compute_units) {
    if (a.total_tiles <= 0) return hipSuccess;
    if (a.tile_counters && compute_units > 0) {
        if (a.total_tiles >= INT32_MAX / 2) return hipErrorInvalidValue;
        zero_counters_kernel<<<1, 8 * COUNTER_SLICES, 0, stream>>>(a.tile_counters);
#define NGX_LAUNCH(W, B) render_persistent_kernel<W, B><<<persistent_grid<W, B>(compute_units, a.total_tiles), 256, 0, stream>>>(a)
        NGX_BY_VARIANT3(a.walk, a.blocks, RENDER_BLOCKS_DEFAULT, 1, 5, 6, 7, NGX_LAUNCH)
#undef NGX_LAUNCH
        return hipGetLastError();
    }
    const long long blocks = (a.total_tiles + 3) / 4;
#define NGX_LAUNCH(W, B) render_kernel<W, B><<<(unsigned)blocks, 256, 0, stream>>>(a)
    NGX_BY_VARIANT3(a.walk, a.blocks, RENDER_BLOCKS_DEFAULT, 1, 5, 6, 7, NGX_LAUNCH)
#undef NGX_LAUNCH
    return hipGetLastError();
}

#ifdef NGX_TRACE_STATS
hipError_t launch_render_stats(const RenderArgs &a, int *out, hipStream_t stream) {
    if (a.total_tiles <= 0) return hipSuccess;
    const long long blocks = (a.total_tiles + 3) / 4;
#define NGX_LAUNCH(W) render_stats_kernel<W><<<(unsigned)blocks, 256, 0, stream>>>(a, out)
    switch (a.walk) {
    case WALK_LANES: NGX_LAUNCH(WALK_LANES); break;
    case WALK_HYBRID: NGX_LAUNCH(WALK_HYBRID); break;
    default: NGX_LAUNCH(WALK_PACKET); break;
    }
#undef NGX_LAUNCH
    return hipGetLastError();
}
#endif

template <int WALK, int BLOCKS>
static int rgb_persistent_grid(const int compute_units, const long long tiles) {
    static const int per_cu = [] {
        int n = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, rgb_persistent_kernel<WALK, BLOCKS>, 256, 0) != hipSuccess || n < 1) n = 1;
        return n;
    }();
    return capped_grid((per_cu * compute_units + 7) / 8 * 8, tiles);
}

// The RGB kernels' variants: 4, 5, 6 or 7 workgroups per compute unit for the packet and hybrid walks, 4 for the
// per-lane walk; defaults 4 for the packet walk and 7 for the hybrid walk.
#define NGX_RGB_VARIANT(walk, blocks, launch) \
    switch (walk) { \
    case WALK_LANES: launch(WALK_LANES, 4); break; \
    case WALK_HYBRID: \
        switch (blocks ? blocks : RGB_BLOCKS_DEFAULT_HYBRID) { case 4: launch(WALK_HYBRID, 4); break; \
            case 5: launch(WALK_HYBRID, 5); break; case 6: launch(WALK_HYBRID, 6); break; default: launch(WALK_HYBRID, 7); break; } \
        break; \
    default: \
        switch (blocks ? blocks : RGB_BLOCKS_DEFAULT_PACKET) { case 5: launch(WALK_PACKET, 5); break; \
            case 6: launch(WALK_PACKET, 6); break; case 7: launch(WALK_PACKET, 7); break; default: launch(WALK_PACKET, 4); break; } \
        break; \
    }

hipError_t launch_rgb(const RgbArgs &a, hipStream_t stream, const int compute_units) {
    if (a.r.total_tiles <= 0) return hipSuccess;
    if (a.spp < 1 || a.shade.light_count < 0 || a.shade.light_count > MAX_LIGHTS) return hipErrorInvalidValue;
    if (a.r.tile_counters && compute_units > 0) {
        if (a.r.total_tiles >= INT32_MAX / 2) return hipErrorInvalidValue;
        zero_counters_kernel<<<1, 8 * COUNTER_SLICES, 0, stream>>>(a.r.tile_counters);
#define NGX_LAUNCH(W, B) rgb_persistent_kernel<W, B><<<rgb_persistent_grid<W, B>(compute_units, a.r.total_tiles), 256, 0, stream>>>(a)
        NGX_RGB_VARIANT(a.r.walk, a.blocks, NGX_LAUNCH)
#undef NGX_LAUNCH
        return hipGetLastError();
    }
    const long long blocks = (a.r.total_tiles + 3) / 4;
#define NGX_LAUNCH(W, B) rgb_kernel<W, B><<<(unsigned)blocks, 256, 0, stream>>>(a)
    NGX_RGB_VARIANT(a.r.walk, a.blocks, NGX_LAUNCH)
#undef NGX_LAUNCH
    return hipGetLastError();
}

template <auto KERNEL>
static int resident_grid(const int compute_units, const long long tiles) {
    static const int per_cu = [] {
        int n = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, KERNEL, 256, 0) != hipSuccess || n < 1) n = 1;
        return n;
    }();
    return capped_grid((per_cu * compute_units + 7) / 8 * 8, tiles);
}

// The two-pass picture, persistent launches of the packet and hybrid walks only: the hit pass built like the depth
// kernel (5, 6 or 7 workgroups per compute unit: hit_blocks, 0 for the default), the shading pass like the RGB
// kernel (4 to 7: a.blocks, 0 for 4).
// The shading pass, built for one light where the scene has one.
template <int WALK, int BLOCKS>
static void launch_shade(const RgbArgs &a, hipStream_t stream, const int compute_units) {
    if (a.shade.light_count == 1)
        rgb_shade_kernel<WALK, BLOCKS, true><<<resident_grid<rgb_shade_kernel<WALK, BLOCKS, true>>(compute_units, a.r.total_tiles), 256, 0,
                                               stream>>>(a);
    else
        rgb_shade_kernel<WALK, BLOCKS, false><<<resident_grid<rgb_shade_kernel<WALK, BLOCKS, false>>(compute_units, a.r.total_tiles), 256,
                                                0, stream>>>(a);
}

hipError_t launch_rgb_two_pass(const RgbArgs &a, const int hit_blocks, hipStream_t stream, const int compute_units) {
    if (a.r.total_tiles <= 0) return hipSuccess;
    if (a.spp != 1 || a.shade.light_count < 0 || a.shade.light_count > MAX_LIGHTS || a.hits == nullptr ||
        a.r.tile_counters == nullptr || compute_units <= 0 || a.r.walk == WALK_LANES || a.r.total_tiles >= INT32_MAX / 2)
        return hipErrorInvalidValue;
    const bool hybrid = a.r.walk == WALK_HYBRID;
    zero_counters_kernel<<<1, 8 * COUNTER_SLICES, 0, stream>>>(a.r.tile_counters);
#define NGX_LAUNCH(W, B) rgb_hit_kernel<W, B><<<resident_grid<rgb_hit_kernel<W, B>>(compute_units, a.r.total_tiles), 256, 0, stream>>>(a)
    switch (hit_blocks ? hit_blocks : RENDER_BLOCKS_DEFAULT) {
    case 5: if (hybrid) NGX_LAUNCH(WALK_HYBRID, 5); else NGX_LAUNCH(WALK_PACKET, 5); break;
    case 6: if (hybrid) NGX_LAUNCH(WALK_HYBRID, 6); else NGX_LAUNCH(WALK_PACKET, 6); break;
    default: if (hybrid) NGX_LAUNCH(WALK_HYBRID, 7); else NGX_LAUNCH(WALK_PACKET, 7); break;
    }
#undef NGX_LAUNCH
    zero_counters_kernel<<<1, 8 * COUNTER_SLICES, 0, stream>>>(a.r.tile_counters);
#define NGX_LAUNCH(W, B) launch_shade<W, B>(a, stream, compute_units)
    switch (a.blocks ? a.blocks : a.shade.light_count == 1 ? SHADE_ONE_LIGHT_BLOCKS_DEFAULT : SHADE_BLOCKS_DEFAULT) {
    case 4: if (hybrid) NGX_LAUNCH(WALK_HYBRID, 4); else NGX_LAUNCH(WALK_PACKET, 4); break;
    case 5: if (hybrid) NGX_LAUNCH(WALK_HYBRID, 5); else NGX_LAUNCH(WALK_PACKET, 5); break;
    case 6: if (hybrid) NGX_LAUNCH(WALK_HYBRID, 6); else NGX_LAUNCH(WALK_PACKET, 6); break;
    default: if (hybrid) NGX_LAUNCH(WALK_HYBRID, 7); else NGX_LAUNCH(WALK_PACKET, 7); break;
    }
#undef NGX_LAUNCH
    return hipGetLastError();
}

hipError_t launch_rays(const RayArgs &a, hipStream_t stream) {
    if (a.count <= 0) return hipSuccess;
    const long long waves = (a.count + 63) / 64;
    const long long blocks = (waves + 3) / 4;
#define NGX_LAUNCH(W, B) ray_kernel<W, B><<<(unsigned)blocks, 256, 0, stream>>>(a)
    if (a.walk == WALK_SHARED) {
        switch (a.blocks ? a.blocks : RAY_BLOCKS_DEFAULT) {
        case 5: NGX_LAUNCH(WALK_SHARED, 5); break;
        case 7: NGX_LAUNCH(WALK_SHARED, 7); break;
        default: NGX_LAUNCH(WALK_SHARED, 6); break;
        }
    } else {
        NGX_BY_VARIANT3(a.walk, a.blocks, RAY_BLOCKS_DEFAULT, 6, 5, 6, 7, NGX_LAUNCH)
    }
#undef NGX_LAUNCH
    return hipGetLastError();
}

hipError_t launch_depth_backward(const DepthBackwardArgs &a, hipStream_t stream) {
    if (a.views <= 0) return hipSuccess;
    if (a.bodies_per_world <= 0 || a.bodies_per_world > BACKWARD_MAX_BODIES) return hipErrorInvalidValue;
    const size_t lds = sizeof(float) * BWD_WAVES * 7 * (size_t)a.bodies_per_world;
#define NGX_LAUNCH(W, B) depth_backward_kernel<W, B><<<a.views * a.parts, 64 * BWD_WAVES, lds, stream>>>(a)
    NGX_BY_VARIANT3(a.walk, a.blocks, BACKWARD_BLOCKS_DEFAULT, 5, 4, 5, 6, NGX_LAUNCH)
#undef NGX_LAUNCH
    return hipGetLastError();
}

hipError_t launch_range_backward(const RangeBackwardArgs &a, hipStream_t stream) {
    if (a.worlds <= 0 || a.chunks <= 0) return hipSuccess;
    if (a.bodies_per_world <= 0 || a.bodies_per_world > BACKWARD_MAX_BODIES) return hipErrorInvalidValue;
    const size_t lds = sizeof(float) * BWD_WAVES * 7 * (size_t)a.bodies_per_world;
#define NGX_LAUNCH(W) range_backward_kernel<W><<<a.worlds * a.chunks, 64 * BWD_WAVES, lds, stream>>>(a)
    switch (a.walk) {
    case WALK_LANES: NGX_LAUNCH(WALK_LANES); break;
    case WALK_HYBRID: NGX_LAUNCH(WALK_HYBRID); break;
    case WALK_SHARED: NGX_LAUNCH(WALK_SHARED); break;
    default: NGX_LAUNCH(WALK_PACKET); break;
    }
#undef NGX_LAUNCH
    return hipGetLastError();
}

hipError_t launch_pose_grad_finish(const PoseGradFinishArgs &a, hipStream_t stream) {
    const int n = a.worlds * a.bodies_per_world;
    if (n <= 0) return hipSuccess;
    pose_grad_finish_kernel<<<(n + 255) / 256, 256, 0, stream>>>(a);
    return hipGetLastError();
}

// The allocation-free depth backward's partial sums (ngx_pose_backward_depth_noalloc). The name `scratch` is
// a host-side stand-in: kernels and copies take the device address that noalloc_scratch() looks up, one per
// device, since each GPU holds its own copy of the array.
static constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mutex;
static float *g_scratch_ptr[MAX_DEVICES] = {};

hipError_t noalloc_scratch(float **out) {
    *out = nullptr;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[dev] == nullptr) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch));
        if (e != hipSuccess) return e;
        if (p == nullptr) return hipErrorInvalidSymbol;
        g_scratch_ptr[dev] = static_cast<float *>(p);
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

}  // namespace ngx
