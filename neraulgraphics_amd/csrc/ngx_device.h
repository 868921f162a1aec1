// Device-side layouts and arithmetic shared by every kernel of the MI355X backend.
//
// Every float expression here is written in the operation order of the reference's shared core
// (src/geom_dev.h, src/device_scene.h, src/mesh_dev.h, src/render_core.h) and the library is compiled
// with -ffp-contract=off and IEEE division and square root, so a hit distance computed here is the
// same IEEE single the reference's CPU backend computes. The one deliberate difference from the
// reference's CUDA build is normalize(): 1/sqrtf as the CPU does, never a reciprocal square root
// approximation.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ngx {

constexpr float NO_HIT = 1e30f;
constexpr int OBJECT_NONE = 0;
constexpr int OBJECT_SPHERE = 1;
constexpr int OBJECT_FLOOR = 33;
constexpr int OBJECT_INSTANCE = 34;
constexpr int MAX_SPHERES = 32;

// The reference's 32-byte node layouts (src/bvh.h:30-49), kept bit for bit: the top-level build writes
// TlasNode so its topology can be compared with the reference algorithm's, and the bottom level is
// built on the host in Node layout first.
struct alignas(16) TlasNode {
    float lo[3];
    int32_t left;
    float hi[3];
    int32_t right;
};
struct alignas(16) BlasNode {
    float lo[3];
    int32_t a;
    float hi[3];
    int32_t count;
};
static_assert(sizeof(TlasNode) == 32 && sizeof(BlasNode) == 32, "32-byte nodes");

// The traversal layout: one 64-byte record per interior node holding both children's boxes, so a visit
// is one scalar load of 16 dwords instead of three node reads. Child references: >= 0 an interior node
// of the same array; < 0 a leaf, ~ref being an instance (top level) or the first triangle of a run of
// `count` (bottom level).
struct alignas(64) WideNode {
    float l_lo[3], l_hi[3];
    float r_lo[3], r_hi[3];
    int32_t left, right;
    int32_t left_count, right_count;
};
static_assert(sizeof(WideNode) == 64, "64-byte wide node");

// What the tracer needs of an instance: world-to-mesh rows and which mesh and template entry it is.
struct alignas(16) InstanceRec {
    float4 inv[3];
    int32_t mesh;
    int32_t tmpl;       // index in the world's instance list: the object ID is OBJECT_INSTANCE + tmpl
    int32_t mesh_tris;  // the mesh's triangle count and the root of its skip layout, copied here so the walk
    int32_t skip_root;  // reads them with the instance's rows instead of with loads of their own
};
static_assert(sizeof(InstanceRec) == 64, "64-byte instance record");

// Per mesh: its bottom-level root, as the box the reference tests first and the reference to descend.
struct alignas(16) MeshRoot {
    float lo[3];
    int32_t ref;      // >= 0 interior wide node; < 0: a leaf run starting at triangle ~ref
    float hi[3];
    int32_t count;    // triangles in the run when ref < 0
};
static_assert(sizeof(MeshRoot) == 32, "32-byte mesh root");

// The bottom level again, for the per-lane walk: one 64-byte record per node of the reference layout (same
// indices), with the next node to visit after this node's subtree for each of the eight ray-direction octants
// (octant bit k set when the ray's direction component k is >= 0). Children are visited in octant order:
// the child whose centre is lower along `axis` first when the direction is positive along it.
struct alignas(64) SkipNode {
    float lo[3];
    int32_t a;        // interior: left child (the right is a + 1); leaf: first triangle
    float hi[3];
    int32_t meta;     // leaf: triangle count (> 0); interior: -(1 + 2 * axis + flip), flip set when the left child's centre is higher
    int32_t escape[8];
};
static_assert(sizeof(SkipNode) == 64, "64-byte skip node");

// ------------------------------------------------------------------ float3 arithmetic --

__device__ __forceinline__ float3 mk3(float x, float y, float z) { return make_float3(x, y, z); }
__device__ __forceinline__ float3 add3(const float3 a, const float3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ float3 sub3(const float3 a, const float3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float3 mul3(const float3 a, const float s) { return mk3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float dot3(const float3 a, const float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float3 cross3(const float3 a, const float3 b) {
    return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float3 neg3(const float3 a) { return mk3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ float3 xyz(const float4 v) { return mk3(v.x, v.y, v.z); }

// 1/x rounded to nearest, as IEEE division rounds it, in three instructions: the hardware reciprocal
// estimate and one Newton step with fused multiply-adds. Exhaustively equal to 1.0f / x for every
// 2^-126 <= |x| < 2^126 (checked on the device by ngx_debug_check_reciprocal); outside that range the
// division is done.
__device__ __forceinline__ float rcp_rn(const float x) {
    if (__builtin_expect(!(fabsf(x) >= 0x1p-126f && fabsf(x) < 0x1p126f), 0)) return 1.0f / x;
    const float r = __builtin_amdgcn_rcpf(x);
    const float e = __builtin_fmaf(-x, r, 1.0f);
    return __builtin_fmaf(e, r, r);
}

__device__ __forceinline__ float3 normalize3(const float3 v) { return mul3(v, rcp_rn(sqrtf(dot3(v, v)))); }

// 1/x that never gives infinity (mesh_dev.h:58-61).
__device__ __forceinline__ float safe_rcp(const float x) { return rcp_rn(fabsf(x) < 1e-12f ? copysignf(1e-12f, x) : x); }
__device__ __forceinline__ float3 safe_inverse(const float3 d) { return mk3(safe_rcp(d.x), safe_rcp(d.y), safe_rcp(d.z)); }

// Slab test (mesh_dev.h:66-74): entry distance, or NO_HIT when missed or beyond tmax.
__device__ __forceinline__ float box_entry(const float lx, const float ly, const float lz, const float hx, const float hy,
                                           const float hz, const float3 o, const float3 inv_d, const float tmax) {
    const float tx0 = (lx - o.x) * inv_d.x, tx1 = (hx - o.x) * inv_d.x;
    const float ty0 = (ly - o.y) * inv_d.y, ty1 = (hy - o.y) * inv_d.y;
    const float tz0 = (lz - o.z) * inv_d.z, tz1 = (hz - o.z) * inv_d.z;
    const float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), 0.0f));
    const float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), tmax));
    return tn <= tf ? tn : NO_HIT;
}

// The slab test's decision alone, for walks that only need to know whether a box is entered (and, for
// ordering, its entry distance tn). lim is box_limit(tmax): the lane's closest distance clamped to the
// largest float below NO_HIT, so that tn <= min(far, lim) holds exactly when box_entry(...) < NO_HIT
// (tn and far are never NaN: the minimum and maximum return the other operand of a NaN, and the last
// terms, 0 and tmax, are numbers). Computing lim once per node keeps the per-box work to the slab test.
constexpr float BELOW_NO_HIT = __builtin_bit_cast(float, __builtin_bit_cast(unsigned, NO_HIT) - 1u);
__device__ __forceinline__ float box_limit(const float tmax) { return fminf(tmax, BELOW_NO_HIT); }
__device__ __forceinline__ bool box_hit(const float lx, const float ly, const float lz, const float hx, const float hy,
                                        const float hz, const float3 o, const float3 inv_d, const float lim, float &tn) {
    const float tx0 = (lx - o.x) * inv_d.x, tx1 = (hx - o.x) * inv_d.x;
    const float ty0 = (ly - o.y) * inv_d.y, ty1 = (hy - o.y) * inv_d.y;
    const float tz0 = (lz - o.z) * inv_d.z, tz1 = (hz - o.z) * inv_d.z;
    tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), 0.0f));
    const float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), lim));
    return tn <= tf;
}

// Triangle j of the packed array (v0, e1, e2: 9 floats).
__device__ __forceinline__ void load_tri(const float *tris, const size_t j, float3 &v0, float3 &e1, float3 &e2) {
    const float *p = tris + 9 * j;
    v0 = make_float3(p[0], p[1], p[2]);
    e1 = make_float3(p[3], p[4], p[5]);
    e2 = make_float3(p[6], p[7], p[8]);
}

// Two-sided Moller-Trumbore (mesh_dev.h:80-99), triangle given as v0, e1, e2.
__device__ __forceinline__ bool triangle_hit(const float3 v0, const float3 e1, const float3 e2, const float3 o, const float3 d,
                                             const float tmax, float &t, float &bu, float &bv) {
    const float3 p = cross3(d, e2);
    const float det = dot3(e1, p);
    if (fabsf(det) < 1e-14f) return false;
    const float inv_det = rcp_rn(det);
    const float3 s = sub3(o, v0);
    const float u = dot3(s, p) * inv_det;
    if (u < 0.0f || u > 1.0f) return false;
    const float3 q = cross3(s, e1);
    const float v = dot3(d, q) * inv_det;
    if (v < 0.0f || u + v > 1.0f) return false;
    t = dot3(e2, q) * inv_det;
    bu = u;
    bv = v;
    return t > 0.0f && t < tmax;
}

// ------------------------------------------------------------------ wave primitives --

__device__ __forceinline__ uint64_t ballot(const bool p) { return __builtin_amdgcn_ballot_w64(p); }
__device__ __forceinline__ int lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }
__device__ __forceinline__ int uniform(const int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint32_t uniform_u(const uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }
__device__ __forceinline__ float uniform_f(const float v) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ __forceinline__ int popc64(const uint64_t m) { return __popcll(m); }
// Whether this lane's bit is set in a wave-uniform mask: the mask becomes the execution mask, no vector work.
__device__ __forceinline__ bool in_mask(const uint64_t m) { return __builtin_amdgcn_inverse_ballot_w64(m); }
__device__ __forceinline__ int first_lane(const uint64_t m) { return __ffsll((unsigned long long)m) - 1; }
// A copy of v held in vector registers even when v is wave-uniform. A vector instruction reads at most one
// scalar operand, so a uniform ray origin met with a node's box (scalar too) would otherwise be copied into a
// vector register before every subtraction; the copy is made once here instead.
__device__ __forceinline__ float in_vgpr(const float v) {
    float r;
    asm("v_mov_b32 %0, %1" : "=v"(r) : "v"(v));
    return r;
}
__device__ __forceinline__ float3 in_vgpr3(const float3 v) { return make_float3(in_vgpr(v.x), in_vgpr(v.y), in_vgpr(v.z)); }
// Inclusive prefix sum over the wavefront (every lane active), one DPP add a stage: within each row of 16 lanes by
// shifts of 1, 2, 4 and 8 (a lane shifted in from outside the row reads 0), then lane 15's sum added to rows 1 and 3 and
// lane 31's to rows 2 and 3.
__device__ __forceinline__ int wave_prefix_sum(int v) {
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, true);   // row_shr:1
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, true);   // row_shr:2
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, true);   // row_shr:4
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, true);   // row_shr:8
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);  // row_bcast:15, rows 1 and 3
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);  // row_bcast:31, rows 2 and 3
    return v;
}
// The sum of v over the wavefront (every lane active), by wave_prefix_sum's stages (lane 63's inclusive sum, read from
// every lane): register to register, with no LDS round trip, in one fixed order.
__device__ __forceinline__ float wave_total(float v) {
#define NGX_DPP_ADD(ctrl, row_mask, bound) \
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), ctrl, row_mask, 0xf, bound))
    NGX_DPP_ADD(0x111, 0xf, true);   // row_shr:1
    NGX_DPP_ADD(0x112, 0xf, true);   // row_shr:2
    NGX_DPP_ADD(0x114, 0xf, true);   // row_shr:4
    NGX_DPP_ADD(0x118, 0xf, true);   // row_shr:8
    NGX_DPP_ADD(0x142, 0xa, false);  // row_bcast:15, rows 1 and 3
    NGX_DPP_ADD(0x143, 0xc, false);  // row_bcast:31, rows 2 and 3
#undef NGX_DPP_ADD
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}
// A vector register holding some value the compiler neither computes nor knows, for the operands of lanes whose results
// are masked out (idle lanes of a per-lane walk): no instruction is spent on it. An uninitialized variable would not do:
// its value is undefined to the compiler, which may then fold it through the masks (an undefined value ANDed with a mask
// may be taken as 0), while an empty asm statement's output is a definite value it cannot see.
__device__ __forceinline__ float unknown_f() {
    float r;
    asm volatile("" : "=v"(r));
    return r;
}
__device__ __forceinline__ int unknown_i() {
    int r;
    asm volatile("" : "=v"(r));
    return r;
}

// Loads through the scalar data cache: the address must be the same in every lane of the wave. The
// constant address space is how the compiler is told the data is read-only for the whole kernel.
typedef float v4f __attribute__((ext_vector_type(4)));
typedef int v4i __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(4))) const v4f const_v4f;
typedef __attribute__((address_space(4))) const v4i const_v4i;
__device__ __forceinline__ float4 sload4(const float4 *p) {
    const v4f v = *(const const_v4f *)((uintptr_t)p);
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ int4 sload4i(const int4 *p) {
    const v4i v = *(const const_v4i *)((uintptr_t)p);
    return make_int4(v.x, v.y, v.z, v.w);
}

}  // namespace ngx
