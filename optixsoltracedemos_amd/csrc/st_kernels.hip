// st_kernels.hip - gfx950 kernels of the SolTrace-style tracer.
//
//   k_frames / k_frames_batch  FP64 element frames, reference AABBs, FP32 packing
//                              (lib/element.cpp:77-275, lib/math_util.h:21-59)
//   k_refit_levels / k_refit   bottom-up BVH refit: all levels in one launch (one block per
//                              step), or one launch per level for large trees
//                              (replaces optixAccelBuild UPDATE, lib/geometry_manager.cpp:249-262)
//   k_sun_plane                FP64 sun plane (lib/sun_utils.cu:53-179,
//                              lib/geometry_manager.cpp:97-148)
//   k_trace                    sun sampling + BVH traversal + intersection +
//                              shading + flux binning (cuda/sun.cu, cuda/intersection.cu,
//                              cuda/materials.cu, flux_map_calculator.py:95-179)
//   k_flux_watts               FP64 power normalisation
//
// Trace design: persistent waves pull fixed-size chunks of rays from a global
// counter; each lane carries one ray through all its bounces with the BVH stack
// in LDS ([depth][lane], conflict-free); statistics live in registers and are
// flushed once per chunk with one atomic per wave.
#include <hip/hip_runtime.h>

#include <atomic>

#include "st_internal.h"

namespace st {

// ---------------------------------------------------------------------------
// FP32 vector helpers
struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 v3(const float4& q) { return V3{q.x, q.y, q.z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 operator*(float s, V3 a) { return v3(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ V3 operator-(V3 a) { return v3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ float dot(V3 a, V3 b) { return fmaf(a.x, b.x, fmaf(a.y, b.y, a.z * b.z)); }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ V3 normalize(V3 a) { return rsqrtf(dot(a, a)) * a; }
__device__ __forceinline__ V3 fma3(float s, V3 a, V3 b) { return v3(fmaf(s, a.x, b.x), fmaf(s, a.y, b.y), fmaf(s, a.z, b.z)); }

// ---------------------------------------------------------------------------
// Global-memory accessors. The job's pointers reach the kernel inside a struct
// (by value or from a device array), so the compiler cannot prove they are
// global and would emit flat instructions, whose completion is tracked
// together with LDS traffic (the traversal stack). These casts make every
// scene read, record write and histogram atomic a global-segment instruction.
typedef float f4v __attribute__((ext_vector_type(4)));
typedef float f2v __attribute__((ext_vector_type(2)));
#define ST_GLOBAL __attribute__((address_space(1)))
#define ST_CONST __attribute__((address_space(4)))  // read-only for the whole launch: scalar loads
__device__ __forceinline__ float4 ld4(const float4* p) {
    const f4v v = *(const ST_GLOBAL f4v*)p;
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ float2 ld2(const void* p) {
    const f2v v = *(const ST_GLOBAL f2v*)p;
    return make_float2(v.x, v.y);
}
__device__ __forceinline__ uint32_t ld1(const uint32_t* p) { return *(const ST_GLOBAL uint32_t*)p; }
__device__ __forceinline__ void st4(float4* p, float4 v) { *(ST_GLOBAL f4v*)p = f4v{v.x, v.y, v.z, v.w}; }
__device__ __forceinline__ void st1(uint32_t* p, uint32_t v) { *(ST_GLOBAL uint32_t*)p = v; }
__device__ __forceinline__ void gadd(void* p, unsigned long long v) {
    __hip_atomic_fetch_add((ST_GLOBAL unsigned long long*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t gadd32(uint32_t* p, uint32_t v) {
    return __hip_atomic_fetch_add((ST_GLOBAL uint32_t*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------
// Sampling

// basis of cuda/sun.cu:77-80: u = normalize(cross(|w.x| > 0.99 ? y : x, w)), v = w x u
__device__ __forceinline__ void cone_basis(V3 w, V3& u, V3& v) {
    V3 a = fabsf(w.x) > 0.99f ? v3(0.f, 1.f, 0.f) : v3(1.f, 0.f, 0.f);
    u = normalize(cross(a, w));
    v = cross(w, u);
}

// sin and cos of 2 pi x, x in [0, 1): the hardware sine and cosine take their
// argument in revolutions (error ~1e-6 absolute, which moves a direction
// sampled in a cone of half angle theta by ~1e-6 theta)
__device__ __forceinline__ void sincos_2pi(float x, float& s, float& c) {
    s = __builtin_amdgcn_sinf(x);
    c = __builtin_amdgcn_cosf(x);
}

// sin and cos of an angle that is small in practice (cone and error angles of
// a few mrad): Taylor series through x^9 / x^10 for |x| < 0.35, whose
// truncation error (< 1e-9 relative) is below FP32 rounding; sincosf beyond
__device__ __forceinline__ void sincos_small(float x, float& s, float& c) {
    if (fabsf(x) < 0.35f) {
        const float x2 = x * x;
        s = fmaf(x * x2, fmaf(x2, fmaf(x2, fmaf(x2, 2.7557319e-6f, -1.9841270e-4f), 8.3333333e-3f), -0.16666667f), x);
        c = fmaf(x2, fmaf(x2, fmaf(x2, fmaf(x2, 2.4801587e-5f, -1.3888889e-3f), 4.1666667e-2f), -0.5f), 1.f);
    } else {
        sincosf(x, &s, &c);
    }
}

// sqrt(-2 ln u) of a uniform u in (0, 1] (Box-Muller radius), from the
// hardware log2 and square root
__device__ __forceinline__ float box_muller_radius(float u) {
    return __builtin_amdgcn_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(u));
}

// isotropic 2-D Gaussian deviation (polar Box-Muller): theta = sigma sqrt(-2 ln xr)
__device__ __forceinline__ V3 perturb(V3 a, float sigma, float xr, float xp) {
    V3 u, v;
    cone_basis(a, u, v);
    float st, ct, sp, cp;
    sincos_small(sigma * box_muller_radius(xr), st, ct);
    sincos_2pi(xp, sp, cp);
    return normalize(fma3(ct, a, st * fma3(cp, u, sp * v)));
}

__device__ __forceinline__ float halton2(uint32_t i) {
    return (float)((double)__brev(i) * 2.3283064365386962890625e-10);
}
// Base-3 radical inverse num / 3^digits of i, five digits at a time: a block
// r of five ternary digits (least significant first) contributes its reversal
// kRev5[r]. Trailing zero digits scale num and den alike, so the quotient, a
// correctly rounded double division of the same rational, is the digit-by-digit
// one's exactly; i < 2^32 has at most 21 digits (five blocks, den = 3^25).
struct Rev5Table {
    uint8_t v[243];
    constexpr Rev5Table() : v() {
        for (int r = 0; r < 243; ++r) {
            int x = r, rev = 0;
            for (int d = 0; d < 5; ++d) {
                rev = rev * 3 + x % 3;
                x /= 3;
            }
            v[r] = (uint8_t)rev;
        }
    }
};
__constant__ Rev5Table kRev5 = Rev5Table();
__device__ __forceinline__ float halton3(uint32_t i) {
    uint64_t num = 0;
#pragma unroll
    for (int b = 0; b < 5; ++b) {
        const uint32_t q = i / 243u;
        num = num * 243u + kRev5.v[i - 243u * q];
        i = q;
    }
    return (float)((double)num / 847288609443.0);  // 3^25
}

// ---------------------------------------------------------------------------
// Intersections (FP32). Same root selection and bounds rules as the
// reference programs; the roots are evaluated in cancellation-free form.

// cuda/intersection.cu:47-85
__device__ __forceinline__ bool isect_flat(const float4& q0, const float4& q1, const float4& q2, const float4& q3,
                                          V3 o, V3 d, float tmin, float tmax, float& tout) {
    V3 c = v3(q0), n = v3(q3);
    V3 oc = c - o;
    float t = dot(n, oc) * __builtin_amdgcn_rcpf(dot(d, n));
    if (!(t > tmin && t < tmax)) return false;
    V3 w = fma3(t, d, -oc);
    float x = dot(v3(q1), w), y = dot(v3(q2), w);
    if (!(fabsf(x) <= q0.w && fabsf(y) <= q1.w)) return false;
    tout = t;
    return true;
}

// cuda/intersection.cu:318-448
__device__ __forceinline__ bool isect_parabolic(const float4& q0, const float4& q1, const float4& q2,
                                               const float4& q3, V3 o, V3 d, float tmin, float tmax, float& tout) {
    V3 r = o - v3(q0);
    V3 e1 = v3(q1), e2 = v3(q2), n = v3(q3);
    float cx = q2.w, cy = q3.w;
    float ox = dot(r, e1), oy = dot(r, e2), oz = dot(r, n);
    float dx = dot(d, e1), dy = dot(d, e2), dz = dot(d, n);
    float A = 0.5f * cx * dx * dx + 0.5f * cy * dy * dy;
    float B = cx * ox * dx + cy * oy * dy - dz;
    float C = 0.5f * cx * ox * ox + 0.5f * cy * oy * oy - oz;
    float t = 0.f;
    bool valid = false;
    if (fabsf(A) < 1e-12f) {
        t = -C * __builtin_amdgcn_rcpf(B);
        valid = t > 0.f;
    } else {
        float disc = B * B - 4.f * A * C;
        if (disc >= 0.f) {
            float sq = sqrtf(disc);
            float qq = B >= 0.f ? -0.5f * (B + sq) : -0.5f * (B - sq);
            float ra = qq * __builtin_amdgcn_rcpf(A), rb = qq != 0.f ? C * __builtin_amdgcn_rcpf(qq) : ra;
            float t1 = B >= 0.f ? ra : rb;
            float t2 = B >= 0.f ? rb : ra;
            if (t1 > 0.f && t1 < t2) {
                t = t1;
                valid = true;
            } else if (t2 > 0.f) {
                t = t2;
                valid = true;
            }
        }
    }
    if (!valid || t < tmin || t > tmax) return false;
    float x = fmaf(t, dx, ox), y = fmaf(t, dy, oy);
    if (x < -q0.w || x > q0.w || y < -q1.w || y > q1.w) return false;
    tout = t;
    return true;
}

// cuda/intersection.cu:177-296 (capped cylinder, axis = local y)
__device__ __forceinline__ bool isect_cylinder(const float4& q0, const float4& q1, const float4& q2,
                                              const float4& q3, V3 o, V3 d, float tmin, float tmax, float& tout,
                                              int& cap) {
    V3 r = o - v3(q0);
    V3 ax = v3(q1), ay = v3(q2), az = v3(q3);
    float rad = q0.w, hh = q1.w;
    float ox = dot(r, ax), oy = dot(r, ay), oz = dot(r, az);
    float dx = dot(d, ax), dy = dot(d, ay), dz = dot(d, az);
    float tbig = tmax + 1.f;
    float t_wall = tbig;
    float A = dx * dx + dz * dz;
    if (A > 0.f) {
        const float rA = __builtin_amdgcn_rcpf(A);
        float tc = -(ox * dx + oz * dz) * rA;
        float px = fmaf(tc, dx, ox), pz = fmaf(tc, dz, oz);
        float h2 = rad * rad - (px * px + pz * pz);
        if (h2 >= 0.f) {
            float dt = sqrtf(h2 * rA);
            float t1 = tc - dt, t2 = tc + dt;
            if (t1 > tmin && t1 < tmax && fabsf(fmaf(t1, dy, oy)) <= hh) t_wall = t1;
            else if (t2 > tmin && t2 < tmax && fabsf(fmaf(t2, dy, oy)) <= hh) t_wall = t2;
        }
    }
    float t_cap = tbig;
    if (fabsf(dy) > 1e-6f) {
        float inv = __builtin_amdgcn_rcpf(dy);
        float tb = (-hh - oy) * inv, tt = (hh - oy) * inv;
        float bx = fmaf(tb, dx, ox), bz = fmaf(tb, dz, oz);
        if (tb > tmin && tb < tmax && bx * bx + bz * bz <= rad * rad) t_cap = tb;
        float cx = fmaf(tt, dx, ox), cz = fmaf(tt, dz, oz);
        if (tt > tmin && tt < tmax && cx * cx + cz * cz <= rad * rad) t_cap = fminf(t_cap, tt);
    }
    float t = fminf(t_wall, t_cap);
    if (t >= tmax || t <= tmin) return false;
    cap = !(t == t_wall);
    tout = t;
    return true;
}

__device__ __forceinline__ bool isect_record(uint32_t kind, const float4& q0, const float4& q1, const float4& q2,
                                            const float4& q3, V3 o, V3 d, float tmin, float tmax, float& t, int& cap) {
    cap = 0;
    if (kind == kKindFlat) return isect_flat(q0, q1, q2, q3, o, d, tmin, tmax, t);
    if (kind == kKindParabolic) return isect_parabolic(q0, q1, q2, q3, o, d, tmin, tmax, t);
    return isect_cylinder(q0, q1, q2, q3, o, d, tmin, tmax, t, cap);
}
__device__ __forceinline__ bool isect_prim(const float4* __restrict__ prims, uint32_t kind, int slot, V3 o, V3 d,
                                          float tmin, float tmax, float& t, int& cap) {
    const float4 q0 = ld4(prims + 4 * slot + 0), q1 = ld4(prims + 4 * slot + 1), q2 = ld4(prims + 4 * slot + 2),
                 q3 = ld4(prims + 4 * slot + 3);
    return isect_record(kind, q0, q1, q2, q3, o, d, tmin, tmax, t, cap);
}

// ---------------------------------------------------------------------------
// Closest hit over the BVH: smallest t, equal t -> lower element index.
__device__ __forceinline__ float next_up(float t) { return __int_as_float(__float_as_int(t) + 1); }

// Event counters: ST_STAT_* 0..14, [15] = hits on the receiver kept aside
// (agg_slot). Each wave owns kNumCounters u32 in LDS; every lane that reaches
// a counting site contributes one and the wave adds the number of such lanes
// with a single LDS atomic. The vote and the atomic are scalar and LDS work,
// which issue beside the vector work of other waves; counters kept per lane
// in vector registers measured slower.
constexpr int kNumCounters = 16;
constexpr int kAggCounter = 15;
constexpr uint32_t kMaxChunkRays = 2048;

__device__ __forceinline__ void count_if(uint32_t* wc, int k, bool pred) {
    const unsigned long long m = __ballot(pred);
    const unsigned long long live = __ballot(1);
    if (m && (int)(threadIdx.x & 63) == __ffsll((long long)live) - 1) atomicAdd(wc + k, (uint32_t)__popcll(m));
}
__device__ __forceinline__ void count(uint32_t* wc, int k) { count_if(wc, k, true); }

// One candidate of the closest-hit search. The result over a set of candidates
// (smallest t, equal t -> lower element index; next_up keeps equal-t hits in
// play) does not depend on the order the candidates are tested in.
__device__ __forceinline__ void test_prim(const float4* __restrict__ prims, const uint32_t* __restrict__ meta,
                                          int slot, V3 o, V3 d, float tmin, float& tbest, int& bslot, int& bcap,
                                          uint32_t* wc) {
    const uint32_t m = ld1(meta + slot);
    if (wc) count(wc, ST_STAT_FLAT_TESTS + (int)meta_kind(m));
    float t;
    int cap;
    const float lim = bslot >= 0 ? next_up(tbest) : tbest;
    if (isect_prim(prims, meta_kind(m), slot, o, d, tmin, lim, t, cap)) {
        if (bslot < 0 || t < tbest || (t == tbest && meta_elem(m) < meta_elem(ld1(meta + bslot)))) {
            tbest = t;
            bslot = slot;
            bcap = cap;
        }
    }
}

// A flat candidate of the closest-hit search: test_prim without the kind
// dispatch (the primitive is known to be flat).
__device__ __forceinline__ void test_flat(const float4* __restrict__ prims, const uint32_t* __restrict__ meta, int slot,
                                          V3 o, V3 d, float tmin, float& tbest, int& bslot, int& bcap, uint32_t* wc) {
    if (wc) count(wc, ST_STAT_FLAT_TESTS);
    const float lim = bslot >= 0 ? next_up(tbest) : tbest;
    const float4 q0 = ld4(prims + 4 * slot + 0), q1 = ld4(prims + 4 * slot + 1), q2 = ld4(prims + 4 * slot + 2),
                 q3 = ld4(prims + 4 * slot + 3);
    float t;
    if (isect_flat(q0, q1, q2, q3, o, d, tmin, lim, t)) {
        if (bslot < 0 || t < tbest || (t == tbest && meta_elem(ld1(meta + slot)) < meta_elem(ld1(meta + bslot)))) {
            tbest = t;
            bslot = slot;
            bcap = 0;
        }
    }
}

// Closest hit over a candidate list: a sun ray's sun-plane cell list
// (items[b, e)), or a reflected ray's list of its mirror (first reflections)
// or of its (first, second) mirror pair (second reflections), stored as
// items[b] = number of flat candidates, then those, then the curved ones: the
// lanes of a wave hold lists of different mirrors, and with the kinds apart
// the wave runs one flat loop and one curved loop instead of both tests in
// every step. Every primitive such a ray can reach is listed (conservative
// binning, k_grid_* / k_cone_lists), so the result equals a full BVH
// traversal. bk = list index of the winner.
// The same search over a list that every lane of the wave walks (the sun-plane
// cell of a wave-coherent group of sun rays): b, e and items are wave-uniform,
// so entries, meta words and primitive records come through scalar loads, one
// per wave instead of one per lane, and the loop runs on the scalar unit.
__device__ __forceinline__ void closest_in_shared_list(const JobDev& J, const int32_t* items, uint32_t b, uint32_t e,
                                                       bool headed, V3 o, V3 d, float tmin, float& tbest, int& bslot,
                                                       int& bcap, uint32_t& bk, uint32_t* wk) {
    const ST_CONST uint32_t* it = (const ST_CONST uint32_t*)items;
    const ST_CONST uint32_t* M = (const ST_CONST uint32_t*)J.meta;
    const ST_CONST f4v* P = (const ST_CONST f4v*)J.prims;
    uint32_t k = b, fe = b;  // headed: flat entries [b + 1, fe)
    if (headed) {
        fe = b + 1 + it[b];
        k = b + 1;
    }
    for (; k < e; ++k) {
        const int slot = (int)it[k];
        const uint32_t m = M[slot];
        const uint32_t kind = k < fe ? (uint32_t)kKindFlat : meta_kind(m);
        if (wk) count(wk, ST_STAT_FLAT_TESTS + (int)kind);
        const f4v a0 = P[4 * slot + 0], a1 = P[4 * slot + 1], a2 = P[4 * slot + 2], a3 = P[4 * slot + 3];
        const float4 q0 = make_float4(a0.x, a0.y, a0.z, a0.w), q1 = make_float4(a1.x, a1.y, a1.z, a1.w),
                     q2 = make_float4(a2.x, a2.y, a2.z, a2.w), q3 = make_float4(a3.x, a3.y, a3.z, a3.w);
        const float lim = bslot >= 0 ? next_up(tbest) : tbest;
        float t;
        int cap;
        if (isect_record(kind, q0, q1, q2, q3, o, d, tmin, lim, t, cap)) {
            if (bslot < 0 || t < tbest || (t == tbest && meta_elem(m) < meta_elem(ld1(J.meta + bslot)))) {
                tbest = t;
                bslot = slot;
                bcap = cap;
                bk = k;
            }
        }
    }
}

__device__ __forceinline__ void closest_in_list(const JobDev& J, const int32_t* __restrict__ items, uint32_t b,
                                                uint32_t e, bool headed, V3 o, V3 d, float tmin, float& tbest,
                                                int& bslot, int& bcap, uint32_t& bk, uint32_t* wc) {
    uint32_t* wk = (J.flags & ST_TRACE_COUNT_WORK) ? wc : nullptr;
    {
        const uint64_t ip = (uint64_t)(uintptr_t)items;
        const uint32_t b0 = __builtin_amdgcn_readfirstlane(b), e0 = __builtin_amdgcn_readfirstlane(e);
        const uint32_t i0 = __builtin_amdgcn_readfirstlane((uint32_t)ip),
                       i1 = __builtin_amdgcn_readfirstlane((uint32_t)(ip >> 32));
        const int h0 = __builtin_amdgcn_readfirstlane((int)headed);
        if (!__ballot(b != b0 || e != e0 || (uint32_t)ip != i0 || (uint32_t)(ip >> 32) != i1 || (int)headed != h0)) {
            closest_in_shared_list(J, (const int32_t*)(uintptr_t)(((uint64_t)i1 << 32) | i0), b0, e0, h0 != 0, o, d,
                                   tmin, tbest, bslot, bcap, bk, wk);
            return;
        }
    }
    uint32_t k = b;
    if (headed) {
        const uint32_t fe = b + 1 + ld1((const uint32_t*)items + b);
        for (k = b + 1; k < fe; ++k) {
            const int prev = bslot;
            test_flat(J.prims, J.meta, (int)ld1((const uint32_t*)items + k), o, d, tmin, tbest, bslot, bcap, wk);
            if (bslot != prev) bk = k;
        }
    }
    for (; k < e; ++k) {
        const int prev = bslot;
        test_prim(J.prims, J.meta, (int)ld1((const uint32_t*)items + k), o, d, tmin, tbest, bslot, bcap, wk);
        if (bslot != prev) bk = k;
    }
}

// Candidate list of a segment (lb > le: none, BVH traversal) and the key that
// locates the next segment's list: depth 0 (sun ray): sun-plane cell; depth 1:
// the mirror it reflected off (key = element); depth 2: the pair (first mirror,
// second mirror) (key = index of the second mirror in the first one's list);
// deeper: none. Reflection lists start with their flat-candidate count.
struct SegList {
    const int32_t* items;
    uint32_t lb, le;
    bool headed;
};
__device__ __forceinline__ SegList no_list() { return SegList{nullptr, 1u, 0u, false}; }
__device__ __forceinline__ SegList cell_list(const JobDev& J, uint32_t cell) {
    if (!J.grid_start) return no_list();
    return SegList{J.grid_items, ld1(J.grid_start + cell), ld1(J.grid_start + cell + 1), false};
}
__device__ __forceinline__ SegList next_list(const JobDev& J, int depth, int key) {
    if (key < 0) return no_list();
    if (depth == 1 && J.bounce_range) {
        const float2 r = ld2(J.bounce_range + key);
        return SegList{J.bounce_items, __float_as_uint(r.x), __float_as_uint(r.y), true};
    }
    if (depth == 2 && J.pair_range) {
        const float2 r = ld2(J.pair_range + key);
        return SegList{J.pair_items, __float_as_uint(r.x), __float_as_uint(r.y), true};
    }
    return no_list();
}

template <int STRIDE>
__device__ __forceinline__ void traverse(const float4* __restrict__ nodes, const float4* __restrict__ prims,
                                         const uint32_t* __restrict__ meta, V3 o, V3 d, float tmin, float& tbest,
                                         int& bslot, int& bcap, int32_t* __restrict__ stk, uint32_t* wk) {
    const float ix = 1.f / (fabsf(d.x) > 1e-20f ? d.x : copysignf(1e-20f, d.x));
    const float iy = 1.f / (fabsf(d.y) > 1e-20f ? d.y : copysignf(1e-20f, d.y));
    const float iz = 1.f / (fabsf(d.z) > 1e-20f ? d.z : copysignf(1e-20f, d.z));
    const float ox = o.x * ix, oy = o.y * iy, oz = o.z * iz;
    int sp = 0;
    int32_t node = 0;
    for (;;) {
        while (node >= 0) {
            if (wk) count(wk, ST_STAT_NODE_VISITS);
            const float4 n0 = ld4(nodes + 4 * node + 0), n1 = ld4(nodes + 4 * node + 1),
                         n2 = ld4(nodes + 4 * node + 2);
            const float2 n3 = ld2(nodes + 4 * node + 3);
            float a0 = fmaf(n0.x, ix, -ox), a1 = fmaf(n0.y, ix, -ox);
            float b0 = fmaf(n0.z, iy, -oy), b1 = fmaf(n0.w, iy, -oy);
            float c0 = fmaf(n2.x, iz, -oz), c1 = fmaf(n2.y, iz, -oz);
            float lo0 = fmaxf(fmaxf(fminf(a0, a1), fminf(b0, b1)), fmaxf(fminf(c0, c1), tmin));
            float hi0 = fminf(fminf(fmaxf(a0, a1), fmaxf(b0, b1)), fminf(fmaxf(c0, c1), tbest));
            float e0 = fmaf(n1.x, ix, -ox), e1 = fmaf(n1.y, ix, -ox);
            float f0 = fmaf(n1.z, iy, -oy), f1 = fmaf(n1.w, iy, -oy);
            float g0 = fmaf(n2.z, iz, -oz), g1 = fmaf(n2.w, iz, -oz);
            float lo1 = fmaxf(fmaxf(fminf(e0, e1), fminf(f0, f1)), fmaxf(fminf(g0, g1), tmin));
            float hi1 = fminf(fminf(fmaxf(e0, e1), fmaxf(f0, f1)), fminf(fmaxf(g0, g1), tbest));
            const bool h0 = lo0 <= hi0, h1 = lo1 <= hi1;
            const int32_t ch0 = __float_as_int(n3.x), ch1 = __float_as_int(n3.y);
            if (h0 && h1) {
                const bool swap = lo1 < lo0;
                stk[sp * STRIDE] = swap ? ch0 : ch1;
                ++sp;
                node = swap ? ch1 : ch0;
            } else if (h0) {
                node = ch0;
            } else if (h1) {
                node = ch1;
            } else {
                if (sp == 0) return;
                node = stk[--sp * STRIDE];
            }
        }
        const int32_t start = leaf_start(node), cnt = leaf_count(node);
        for (int k = 0; k < cnt; ++k) test_prim(prims, meta, start + k, o, d, tmin, tbest, bslot, bcap, wk);
        if (sp == 0) return;
        node = stk[--sp * STRIDE];
    }
}

// ---------------------------------------------------------------------------

// emit(idx, n) for every group of the calling lanes that share an index, with
// n = the group's size, for up to ROUNDS groups; lanes left after that emit
// (idx, 1) each (lanes may diverge around the call). Coherent waves share few
// indices (the mirrors a wave's rays hit), so most calls end after a round or
// two.
template <int ROUNDS, class Emit>
__device__ __forceinline__ void add_one_grouped(uint32_t idx, Emit emit) {
    const int lane = (int)(threadIdx.x & 63);
    bool done = false;
    for (int r = 0; r < ROUNDS; ++r) {
        if (!__ballot(!done)) return;
        if (!done) {
            // the first lane still to add leads; readfirstlane reads it
            // without a cross-lane shuffle through LDS
            const uint32_t lidx = __builtin_amdgcn_readfirstlane(idx);
            const bool mine = idx == lidx;
            const unsigned long long same = __ballot(mine);
            if (lane == __ffsll((long long)__ballot(1)) - 1) emit(lidx, (uint32_t)__popcll(same));
            done = mine;
        }
    }
    if (!done) emit(idx, 1u);
}

// Per-wave log of pending output atomics in LDS: word 0 = entries, then one
// u32 per entry: kind << 31 (1 = flux bin, 0 = element hit) | (count - 1) << 25
// | index. On this architecture a wave's vector-memory counter retires loads,
// stores and atomics in issue order, so an atomic sent to the memory-side
// units from inside the trace would hold up the wave's next load; logged, the
// atomics go out in bursts (kLogCap - kLogHeadroom entries or at the chunk's
// end), each a full-width vector atomic. One trace step adds at most
// kLogHeadroom entries (a lane adds at most two).
constexpr uint32_t kLogCap = 320, kLogHeadroom = 128;
// rank of the calling lane among the set lanes of m below it
__device__ __forceinline__ uint32_t lane_rank(unsigned long long m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
// the calling wave's copy of the flux map
__device__ __forceinline__ int64_t* wave_flux(const JobDev& J) {
    const uint32_t w = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    return J.flux + (size_t)(w & J.flux_copy_mask) * J.flux_bins;
}
__device__ __forceinline__ void log_add(uint32_t* log, const JobDev& J, bool flux, uint32_t idx, uint32_t n) {
    if (idx < (1u << 25) && n <= 64u) {
        const uint32_t pos = atomicAdd(log, 1u);
        if (pos < kLogCap) {
            log[1 + pos] = ((uint32_t)flux << 31) | ((n - 1u) << 25) | idx;
            return;
        }
    }
    gadd(flux ? (void*)(wave_flux(J) + idx) : (void*)(J.hits + idx), (unsigned long long)n);
}
// sends the first n entries; every lane of the wave calls it (uniform control flow)
__device__ __forceinline__ void log_send(const uint32_t* log, const JobDev& J, uint32_t n) {
    int64_t* const fl = J.flux ? wave_flux(J) : nullptr;
    for (uint32_t i = threadIdx.x & 63; i < n; i += 64) {
        const uint32_t e = log[1 + i], idx = e & 0x1FFFFFFu, cnt = ((e >> 25) & 63u) + 1u;
        gadd((e >> 31) ? (void*)(fl + idx) : (void*)(J.hits + idx), (unsigned long long)cnt);
    }
}
__device__ __forceinline__ void log_flush(uint32_t* log, const JobDev& J) {
    log_send(log, J, min(log[0], kLogCap));
    if ((threadIdx.x & 63) == 0) log[0] = 0;
}

// The queued form logs where the wave has reconverged instead: a segment
// leaves at most one pending element hit and one pending flux bin per lane
// (kNoEntry: none), and the wave appends them after the segment with every
// lane present. The log's length is then wave-uniform and lives in a scalar
// register (len), so no LDS atomic with its round trip reserves the slots.
// Element hits are grouped by element (a coherent wave hits few mirrors);
// flux bins take one entry per lane, grouped over two rounds on small maps,
// where a wave's receiver hits share bins. Adds at most 128 entries.
constexpr uint32_t kNoEntry = 0xFFFFFFFFu;
__device__ __forceinline__ void log_pending(uint32_t* log, uint32_t& len, const JobDev& J, uint32_t ph, uint32_t pf) {
    const uint32_t lane = threadIdx.x & 63u;
    bool done = ph == kNoEntry;
    for (unsigned long long left = __ballot(!done); left; left = __ballot(!done)) {
        const uint32_t leader = (uint32_t)__ffsll((long long)left) - 1u;
        const uint32_t e = (uint32_t)__builtin_amdgcn_readlane((int)ph, (int)leader);
        const bool mine = !done && ph == e;
        const uint32_t n = (uint32_t)__popcll(__ballot(mine));
        if (e < (1u << 25)) {
            if (lane == leader) log[1 + len] = ((n - 1u) << 25) | e;
            ++len;
        } else if (lane == leader) {
            gadd(J.hits + e, (unsigned long long)n);
        }
        done = done || mine;
    }
    done = pf == kNoEntry;
    if (J.flux_bins < kFluxDirectBins) {
        for (int r = 0; r < 2; ++r) {
            const unsigned long long left = __ballot(!done);
            if (!left) break;
            const uint32_t leader = (uint32_t)__ffsll((long long)left) - 1u;
            const uint32_t b = (uint32_t)__builtin_amdgcn_readlane((int)pf, (int)leader);
            const bool mine = !done && pf == b;
            const uint32_t n = (uint32_t)__popcll(__ballot(mine));
            if (lane == leader) log[1 + len] = (1u << 31) | ((n - 1u) << 25) | b;
            ++len;
            done = done || mine;
        }
    }
    const unsigned long long m = __ballot(!done);
    if (J.flux_bins <= (1u << 25)) {
        if (!done) log[1 + len + lane_rank(m)] = (1u << 31) | pf;
        len += (uint32_t)__popcll(m);
    } else if (!done) {
        gadd(wave_flux(J) + pf, 1ull);
    }
}

// A ray's whole record at once, where the ray starts: slot 0 = its sun-plane
// origin, every later slot empty (all zero, element 0xFFFFFFFF); hits overwrite
// their slots later, instead of each slot taking a partial cache line as the
// ray advances. A full wave of 64 consecutive rays owns one contiguous block
// of 64 * max_depth slots and writes it lane by lane in order (each store
// instruction covers 1 KiB of records, 256 B of element ids), the origins
// fetched from their lanes; other waves write ray by ray.
__device__ __forceinline__ void record_start(const JobDev& J, uint64_t ray, V3 o) {
    if (!J.records) return;
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t r0 = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(ray >> 32)) << 32) |
                        (uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)ray);
    if (__ballot(ray == r0 + lane) == ~0ull) {
        const uint32_t D = (uint32_t)J.max_depth;
        const uint64_t base = (r0 - J.record_base) * (uint64_t)D;
        // slot idx = lane + 64 j of the block belongs to ray r = idx / D, slot s = idx % D
        const uint32_t dr = 64u / D, ds = 64u % D;
        uint32_t r = lane / D, s = lane - r * D;
        for (uint32_t j = 0; j < D; ++j) {
            const float x = __shfl(o.x, (int)r), y = __shfl(o.y, (int)r), z = __shfl(o.z, (int)r);
            st4(J.records + base + lane + 64u * j, s == 0 ? make_float4(0.f, x, y, z) : make_float4(0.f, 0.f, 0.f, 0.f));
            if (J.rec_elem) st1(J.rec_elem + base + lane + 64u * j, s == 0 ? 0xFFFFFFFEu : 0xFFFFFFFFu);
            r += dr;
            s += ds;
            if (s >= D) {
                s -= D;
                ++r;
            }
        }
        return;
    }
    const uint64_t i = (ray - J.record_base) * (uint64_t)J.max_depth;
    st4(J.records + i, make_float4(0.f, o.x, o.y, o.z));
    for (int k = 1; k < J.max_depth; ++k) st4(J.records + i + k, make_float4(0.f, 0.f, 0.f, 0.f));
    if (J.rec_elem) {
        st1(J.rec_elem + i, 0xFFFFFFFEu);
        for (int k = 1; k < J.max_depth; ++k) st1(J.rec_elem + i + k, 0xFFFFFFFFu);
    }
}

__device__ __forceinline__ void record(const JobDev& J, uint64_t ray, int slot, float tag, V3 p, uint32_t elem) {
    if (J.records) {
        const uint64_t i = (ray - J.record_base) * (uint64_t)J.max_depth + (uint64_t)slot;
        st4(J.records + i, make_float4(tag, p.x, p.y, p.z));
        if (J.rec_elem) st1(J.rec_elem + i, elem);
    }
}

// Sun sample of one ray: origin on the sun plane, direction in the sun cone
// (cuda/sun.cu:94-134). Counts the ray and writes its record slot 0.
__device__ __forceinline__ void sun_ray(const JobDev& J, const ST_CONST SunFrame& S, uint64_t ray, uint32_t* wc,
                                        V3& o, V3& d, uint32_t& cell, uint32_t cell_hint = 0xFFFFFFFFu) {
    const uint32_t rlo = (uint32_t)ray, rhi = (uint32_t)(ray >> 32);
    // ---- sun sample: cuda/sun.cu:94-134
    const U4 u = philox4x32_10<true>(U4{rlo, rhi, 0u, J.job_id}, J.seed_lo, J.seed_hi);
    float a, b;
    if (S.position_sampling == ST_POS_STRATIFIED) {
        // the 64 rays of a group share a cell and a wave's rays normally form
        // one group: its cell is computed once, from wave-uniform values (scalar
        // unit); lanes of a second group (unaligned ray ranges) compute theirs
        const uint64_t g = ray >> 6;
        const uint64_t g0 = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(g >> 32)) << 32) |
                            (uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)g);
        // cell_hint: the caller's cell of group g0
        cell = cell_hint != 0xFFFFFFFFu ? cell_hint : stratified_cell_of_group(g0, J.seed_lo, J.seed_hi, J.job_id);
        if (g != g0) cell = stratified_cell_of_group(g, J.seed_lo, J.seed_hi, J.job_id);
        a = ((float)(cell & 255u) + u01(u.x)) * (1.f / 256.f);
        b = ((float)(cell >> 8) + u01(u.y)) * (1.f / 256.f);
    } else {
        if (S.position_sampling == ST_POS_HALTON) {
            a = halton2(rlo);
            b = halton3(rlo);
        } else {
            a = u01(u.x);
            b = u01(u.y);
        }
        // the grid cell of the sampled plane coordinates (a * 256 is exact in
        // FP32; b * rows may round up across a row boundary: the footprints'
        // margins cover that)
        const uint32_t rows = (uint32_t)S.grid_rows;
        cell = (min((uint32_t)(b * (float)rows), rows - 1u) << 8) | min((uint32_t)(a * 256.f), 255u);
    }
    o = v3(fmaf(a, S.e1[0], fmaf(b, S.e2[0], S.o[0])), fmaf(a, S.e1[1], fmaf(b, S.e2[1], S.o[1])),
           fmaf(a, S.e1[2], fmaf(b, S.e2[2], S.o[2])));
    {
        const V3 w = v3(S.w[0], S.w[1], S.w[2]), su = v3(S.u[0], S.u[1], S.u[2]), sv = v3(S.v[0], S.v[1], S.v[2]);
        float sp, cp, st, ct;
        sincos_2pi(u01(u.z), sp, cp);
        if (S.shape == ST_SUN_GAUSSIAN) {
            sincos_small(S.sigma * box_muller_radius(u01(u.w)), st, ct);
        } else {
            const float s = S.vers * (1.f - u01(u.w));
            ct = 1.f - s;
            st = __builtin_amdgcn_sqrtf(s * (2.f - s));
        }
        d = normalize(fma3(ct, w, st * fma3(cp, su, sp * sv)));
    }
    record_start(J, ray, o);
}

// One segment of a ray: closest hit from o along d, then the hit element's
// program: a receiver's here; a mirror's (shade_mirror) is left to the caller,
// to run before the next segment. Returns true when the ray hit a mirror and
// continues (o = hit point, tmin, depth and key updated for the next segment,
// pend = the mirror's primitive for shade_mirror, d unchanged), false when it
// ended.
// DEFER: element hits and flux bins are left in ph / pf for log_pending
// instead of being logged here.
template <int STRIDE, bool DEFER>
__device__ __forceinline__ bool trace_segment(const JobDev& J, uint64_t ray, int32_t* stk, uint32_t* wc,
                                              uint32_t* wh, uint32_t* wl, V3& o, V3& d, float& tmin, int& depth,
                                              int& last, SegList L, int& key, int& pend, uint32_t& ph,
                                              uint32_t& pf) {
    float t = kTmax;
    int slot = -1, cap = 0;
    uint32_t bk = 0xFFFFFFFFu;
    const bool listed = L.lb <= L.le;
    if (listed) closest_in_list(J, L.items, L.lb, L.le, L.headed, o, d, tmin, t, slot, cap, bk, wc);
    else traverse<STRIDE>(J.nodes, J.prims, J.meta, o, d, tmin, t, slot, cap, stk,
                          (J.flags & ST_TRACE_COUNT_WORK) ? wc : nullptr);
    if (slot < 0) {
        // __miss__ms, cuda/materials.cu:236-253: derived at the chunk's end, and
        // counted here too when work is counted (the flush then keeps the count,
        // so a ST_TRACE_COUNT_WORK run checks the derivation)
        if (J.flags & ST_TRACE_COUNT_WORK) count(wc, ST_STAT_ESCAPED);
        return false;
    }
    const int nd = depth + 1;
    const uint32_t m = ld1(J.meta + slot);
    const uint32_t kind = meta_kind(m), elem = meta_elem(m);
    const V3 p = fma3(t, d, o);
    if (meta_receiver(m)) {
        const float4 q0 = ld4(J.prims + 4 * slot + 0), q1 = ld4(J.prims + 4 * slot + 1),
                     q2 = ld4(J.prims + 4 * slot + 2), q3 = ld4(J.prims + 4 * slot + 3);
        const V3 r = o - v3(q0);
        // cuda/materials.cu:86-121 (flat: front face only) and :123-163 (cylinder)
        if (kind == kKindFlat && !(dot(d, v3(q3)) < 0.f)) {
            count(wc, ST_STAT_BACK_FACE);
            return false;
        }
        if (nd < J.max_depth) {
            record(J, ray, nd, 2.f, p, elem);
            last = nd;
            const bool fe = slot == J.flux_slot;
            count_if(wc, kAggCounter, slot == J.agg_slot);
            if (slot != J.agg_slot && J.hits) {
                if (wh && elem < (uint32_t)kLdsHitElements) atomicAdd(wh + elem, 1u);
                else if (DEFER) ph = elem;
                else log_add(wl, J, false, elem, 1u);
            }
            count(wc, ST_STAT_RECEIVER_HITS);
            const bool direct = depth == 0;
            count_if(wc, ST_STAT_DIRECT_SUN, direct);
            count_if(wc, ST_STAT_CAP_HITS, cap != 0);
            if (fe && J.flux && !cap && (!direct || (J.flags & ST_TRACE_FLUX_DIRECT_SUN))) {
                // scripts/post_processing/flux_map_calculator.py:95-179
                const float lx = fmaf(t, dot(d, v3(q1)), dot(r, v3(q1)));
                const float ly = fmaf(t, dot(d, v3(q2)), dot(r, v3(q2)));
                float fx, fy;
                // bins per unit length of the receiver's local axes
                const float by = (float)J.flux_ny * 0.5f * __builtin_amdgcn_rcpf(q1.w);
                if (kind == kKindCylinder) {
                    const float lz = fmaf(t, dot(d, v3(q3)), dot(r, v3(q3)));
                    float th = atan2_poly(-lz, lx);
                    if (th < 0.f) th += 6.283185307179586f;
                    fx = th * (float)J.flux_nx * 0.15915494309189535f;
                } else {
                    // the map's x runs along -x_b: the scripts take the receiver as the
                    // parallelogram v1 = -W x_b, v2 = H y_b of lib/element.cpp:136-139
                    fx = (q0.w - lx) * ((float)J.flux_nx * 0.5f * __builtin_amdgcn_rcpf(q0.w));
                }
                fy = (ly + q1.w) * by;
                int ix = (int)floorf(fx), iy = (int)floorf(fy);
                ix = min(max(ix, 0), J.flux_nx - 1);
                iy = min(max(iy, 0), J.flux_ny - 1);
                // a wave's hits share bins only on small maps
                if (DEFER)
                    pf = (uint32_t)(iy * J.flux_nx + ix);
                else if (J.flux_bins < kFluxDirectBins)
                    add_one_grouped<2>((uint32_t)(iy * J.flux_nx + ix),
                                       [&](uint32_t b, uint32_t n) { log_add(wl, J, true, b, n); });
                else
                    log_add(wl, J, true, (uint32_t)(iy * J.flux_nx + ix), 1u);
                count(wc, ST_STAT_FLUX_BINNED);
            }
        } else {
            count(wc, ST_STAT_TRUNCATED);
        }
        return false;
    }
    // mirror: cuda/materials.cu:26-84, 170-231
    if (nd >= J.max_depth) {
        count(wc, ST_STAT_TRUNCATED);
        return false;
    }
    record(J, ray, nd, 1.f, p, elem);
    last = nd;
    if (J.hits) {
        if (wh && elem < (uint32_t)kLdsHitElements) atomicAdd(wh + elem, 1u);
        else if (DEFER) ph = elem;
        else add_one_grouped<64>(elem, [&](uint32_t e, uint32_t n) { log_add(wl, J, false, e, n); });
    }
    count(wc, ST_STAT_MIRROR_HITS);
    // the mirror's program runs as the first step of the ray's next segment
    // (shade_mirror): d stays the incoming direction until then
    o = p;
    tmin = kTminBounce;
    pend = slot | (cap ? (int)0x80000000u : 0);
    // key of the next segment's list: a first reflection's list depends on its
    // mirror only; a second reflection's on where it sits in that list
    key = nd == 1 ? (int)elem : (nd == 2 && listed) ? (int)bk : -1;
    depth = nd;
    return true;
}

// The program of a mirror hit (cuda/materials.cu:26-84, 170-231) of a ray at
// its hit point o with incoming direction d on primitive pend & 0x7FFFFFFF
// (sign bit: an end-cap hit), bounce depth nd: reflectivity draw, surface
// normal, slope error on the normal, reflection, specular error on the
// reflected ray. Returns false when the ray is absorbed; otherwise d is the
// reflected direction. The queued forms run it at the start of the ray's next
// segment, where the wave holds only continuing rays, instead of in the step
// that found the hit, where the lanes whose ray ended would sit idle through it.
__device__ __forceinline__ bool shade_mirror(const JobDev& J, uint64_t ray, int pend, int nd, V3 o, V3& d,
                                             uint32_t* wc) {
    const uint32_t rlo = (uint32_t)ray, rhi = (uint32_t)(ray >> 32);
    const int slot = pend & 0x7FFFFFFF;
    const bool cap = pend < 0;
    const uint32_t m = ld1(J.meta + slot);
    const uint32_t kind = meta_kind(m), elem = meta_elem(m);
    // draws of bounce nd: block 2nd-1 = (reflectivity, slope xr, slope xp,
    // specular xr), block 2nd = (specular xp, -, -, -), each drawn only when
    // one of its optics terms is active
    const float4 op = ld4(J.optics + elem);
    U4 ue = U4{0u, 0u, 0u, 0u};
    if (op.x < 1.f || op.y > 0.f || op.z > 0.f)
        ue = philox4x32_10<true>(U4{rlo, rhi, 2u * (uint32_t)nd - 1u, J.job_id}, J.seed_lo, J.seed_hi);
    if (op.x < 1.f && u01(ue.x) >= op.x) {
        count(wc, ST_STAT_ABSORBED);
        return false;
    }
    V3 n;
    if (kind == kKindFlat) {
        n = v3(ld4(J.prims + 4 * slot + 3));
    } else {
        const float4 q0 = ld4(J.prims + 4 * slot + 0), q1 = ld4(J.prims + 4 * slot + 1),
                     q2 = ld4(J.prims + 4 * slot + 2), q3 = ld4(J.prims + 4 * slot + 3);
        const V3 ax = v3(q1), ay = v3(q2), az = v3(q3);
        const V3 r = o - v3(q0);
        const float lx = dot(r, ax), ly = dot(r, ay);
        V3 nl;
        if (kind == kKindParabolic) {
            nl = normalize(v3(-q2.w * lx, -q3.w * ly, 1.f));
        } else {
            const float lz = dot(r, az);
            nl = cap ? v3(0.f, ly < 0.f ? -1.f : 1.f, 0.f) : normalize(v3(lx, 0.f, lz));
        }
        n = normalize(fma3(nl.x, ax, fma3(nl.y, ay, nl.z * az)));
    }
    if (dot(d, n) > 0.f) n = -n;  // faceforward(n, -d, n)
    if (op.y > 0.f) n = perturb(n, op.y, u01(ue.y), u01(ue.z));
    V3 rd = normalize(fma3(-2.f * dot(n, d), n, d));
    if (op.z > 0.f) {
        const U4 us = philox4x32_10<true>(U4{rlo, rhi, 2u * (uint32_t)nd, J.job_id}, J.seed_lo, J.seed_hi);
        rd = perturb(rd, op.z, u01(ue.w), u01(us.x));
    }
    d = rd;
    return true;
}

// Residue modulo kHaltonCells = 2^8 * 3^5 of the indices whose Halton
// positions lie in cell k of a row-by-row walk (serpentine) over the 256 x 243
// cells: the cell's column reversed in 8 bits is the index modulo 2^8 (the
// base-2 radical inverse reverses the low bits), its row reversed in 5 ternary
// digits the index modulo 3^5; the Chinese remainder theorem joins the two
// (59 = 243^-1 mod 256, 187 = 256^-1 mod 243).
__device__ __forceinline__ uint32_t halton_residue(uint32_t k) {
    const uint32_t cy = k >> 8, cx = (cy & 1u) ? 255u - (k & 255u) : (k & 255u);
    const uint32_t x = __brev(cx) >> 24;
    uint32_t y = 0, t = cy;
    for (int d = 0; d < 5; ++d) {
        const uint32_t q = t / 3u;
        y = 3u * y + (t - 3u * q);
        t = q;
    }
    return (x * (243u * 59u) + y * (256u * 187u)) % kHaltonCells;
}

// Ray index of slot s of a job's range; slots are traced in order. Inside the
// range's whole blocks of kHaltonBlock aligned indices (Halton positions,
// kOrderHalton) a block is traced cell by cell, neighbouring cells after one
// another: its 64 slots 64k..64k+63 hold the indices base + r + j * kHaltonCells,
// j < 64, r = halton_residue(k), which share a Halton cell.
// Inside whole stratified sweeps (kOrderStratified) slots 64k..64k+63 hold the
// group whose cell is the k-th of the sweep's Hilbert curve. Elsewhere slot s
// is index ray_begin + s. Either way the 64 slots from any multiple of 64 hold
// consecutive indices when the order is stratified.
__device__ __forceinline__ uint64_t ray_of_slot(const JobDev& J, uint64_t s) {
    if (J.order_kind == kOrderStratified && s >= J.order_head) {
        const uint64_t t = s - J.order_head;
        if ((t >> 22) < (uint64_t)J.order_blocks) {
            const uint64_t sweep = ((J.ray_begin + J.order_head) >> 22) + (t >> 22);
            const uint64_t g = stratified_group_of_cell(sweep, hilbert_cell((uint32_t)(t >> 6) & 0xFFFFu), J.seed_lo,
                                                        J.seed_hi, J.job_id);
            return (g << 6) | (t & 63u);
        }
    }
    if (J.order_kind == kOrderHalton && s >= J.order_head) {
        const uint64_t t = s - J.order_head;
        if (t < (uint64_t)J.order_blocks * kHaltonBlock) {
            const uint32_t u = (uint32_t)t;  // the range lies below 2^32
            const uint32_t q = u / kHaltonBlock, r = u - q * kHaltonBlock;
            // the 64 slots of a group share k: its residue on the scalar unit
            const uint32_t k = r >> 6, k0 = __builtin_amdgcn_readfirstlane(k);
            uint32_t res = halton_residue(k0);
            if (k != k0) res = halton_residue(k);
            return J.ray_begin + J.order_head + (uint64_t)q * kHaltonBlock + (uint64_t)(r & 63u) * kHaltonCells + res;
        }
    }
    return J.ray_begin + s;
}

// First ray index of the 64 slots from s (a multiple of 64 slots in a job whose
// order is not Halton's: they hold consecutive indices) and, for stratified
// positions, the sun-plane cell of their group (else 0xFFFFFFFF).
__device__ __forceinline__ uint64_t group_of_slot(const JobDev& J, const ST_CONST SunFrame& S, uint64_t s,
                                                  uint32_t* cell) {
    *cell = 0xFFFFFFFFu;
    if (J.order_kind == kOrderStratified && s >= J.order_head) {
        const uint64_t t = s - J.order_head;
        if ((t >> 22) < (uint64_t)J.order_blocks) {
            const uint64_t sweep = ((J.ray_begin + J.order_head) >> 22) + (t >> 22);
            *cell = hilbert_cell((uint32_t)(t >> 6) & 0xFFFFu);
            return stratified_group_of_cell(sweep, *cell, J.seed_lo, J.seed_hi, J.job_id) << 6;
        }
    }
    const uint64_t r = J.ray_begin + s;
    if (S.position_sampling == ST_POS_STRATIFIED) *cell = stratified_cell_of_group(r >> 6, J.seed_lo, J.seed_hi, J.job_id);
    return r;
}

// empty slots are all-zero (the reference memsets the buffer, lib/soltrace_system.cpp:125)
// slots after the ray's last recorded one stay as record_start left them (empty)
__device__ __forceinline__ void finish_ray(const JobDev&, uint64_t, int) {}

// work_counter[0] hands out chunks, work_counter[1] counts waves that ran out
// of work; the last wave to leave resets both to zero, so the next launch on
// the stream starts from a clean counter without a memset (graph-safe).
// jobs == nullptr: a single job passed by value in `single`.
// STACK: traversal stack entries per lane (>= BVH depth); MINW: waves per SIMD
// the register allocation must allow (LDS: STACK * BLOCK * 4 bytes per block).
// QUEUED = false: each lane carries one ray through all its segments and the
// wave moves to the next 64 rays when its slowest ray ends, so after the first
// segment the lanes whose ray ended sit idle. QUEUED = true: sun rays are traced
// 64 at a time; a ray that continues after a mirror is parked in the wave's LDS
// queue (hit point, incoming direction, ray offset, depth, list key, mirror)
// and the queue is traced in full waves of 64 continuing rays (the rest at the
// chunk's end), each step running the mirror's program (shade_mirror) and then
// the segment it starts.
// Both give identical results: every per-ray quantity depends on the ray index
// and its own path only, and all accumulations are integer sums.
constexpr int kQueueCap = 64;  // < 64 parked between steps (lanes refill from the queue first)

// tag of a parked ray: offset in the chunk (< 2^11) | depth << 11 | (key + 1)
// << 15, key = the list key of its next segment (0: none, or a key >= 2^17 - 1,
// whose segment is then traced through the BVH). A parked ray's last record
// slot is its depth.
__device__ __forceinline__ uint32_t queue_tag(uint32_t offset, int depth, int key) {
    const uint32_t f = (uint32_t)(key + 1) < (1u << 17) ? (uint32_t)(key + 1) : 0u;
    return offset | ((uint32_t)depth << 11) | (f << 15);
}

// LHITS > 0: element hit counts of scenes with at most LHITS elements are
// kept per wave in LDS and added to the output once per chunk.
// STACK = 0: the traversal stack (kStackDepth entries) is the lane's private
// memory instead of LDS. A BVH walk then goes through the scratch path, but the
// LDS left to each wave (parked rays, log, counters) is small enough for 8
// waves per SIMD: the form for scenes whose rays rarely walk the BVH.
template <int BLOCK, int STACK, int MINW, int QUEUED, int LHITS>
__global__ __launch_bounds__(BLOCK, MINW) void k_trace(JobDev single, const JobDev* __restrict__ jobs,
                                                       const uint64_t* __restrict__ chunk_prefix, int32_t n_jobs,
                                                       uint64_t total_chunks, uint32_t* __restrict__ work_counter,
                                                       int32_t chunk_rays, int32_t static_chunks) {
    constexpr int SS = STACK > 0 ? BLOCK : 1;  // stride of the stack's entries
    __shared__ int32_t lds_stack[STACK > 0 ? STACK * BLOCK : 1];
    __shared__ uint32_t lds_counters[BLOCK / 64][kNumCounters];
    __shared__ float lds_queue[QUEUED ? BLOCK / 64 : 1][QUEUED ? 8 * kQueueCap : 1];
    __shared__ uint32_t lds_log[BLOCK / 64][1 + kLogCap];
    __shared__ uint32_t lds_hits[LHITS > 0 ? BLOCK / 64 : 1][LHITS > 0 ? LHITS : 1];
    // queued form (not the LDS-stack variants with LDS hit counts, whose LDS
    // is full): first ray index and
    // stratified cell of each group of 64 slots of the chunk, computed for the
    // whole chunk at its start, 32 groups in 32 lanes (every order but
    // Halton's, whose groups' indices are not consecutive)
    constexpr bool TABLE = QUEUED && (LHITS == 0 || STACK == 0);
    __shared__ uint64_t lds_bases[TABLE ? BLOCK / 64 : 1][TABLE ? kMaxChunkRays / 64 : 1];
    __shared__ uint32_t lds_cells[TABLE ? BLOCK / 64 : 1][TABLE ? kMaxChunkRays / 64 : 1];
    int32_t private_stack[STACK > 0 ? 1 : kStackDepth];
    int32_t* stk = STACK > 0 ? lds_stack + threadIdx.x : private_stack;
    const int lane = threadIdx.x & 63;
    uint32_t* wc = lds_counters[threadIdx.x >> 6];
    if (lane < kNumCounters) wc[lane] = 0;
    uint32_t* wh = LHITS > 0 ? lds_hits[threadIdx.x >> 6] : nullptr;
    uint32_t* wl = lds_log[threadIdx.x >> 6];
    if (lane == 0) wl[0] = 0;
    if (LHITS > 0)
        for (int i = lane; i < LHITS; i += 64) wh[i] = 0;
    const uint32_t waves = gridDim.x * (BLOCK / 64);
    const uint32_t wave_id = __builtin_amdgcn_readfirstlane(blockIdx.x * (BLOCK / 64) + (threadIdx.x >> 6));
    // one chunk of job J (the chunk's index within the job: cj). It is
    // expanded at two call sites, one per source of the job record, so that
    // each copy reads J from one address space (constant: scalar loads, never
    // reloaded after the kernel's own stores); one copy over a pointer that
    // may be either would read every field through generic loads.
    // A job's counts (event counters, LDS hit counts) to its outputs. Rays,
    // segments and escapes are not counted as they happen: every ray traced
    // since the last flush has ended, each in exactly one of escaped, absorbed,
    // back face, truncated or receiver hit, and each segment after the first
    // follows a mirror hit that was not absorbed.
    auto flush_counts = [&](const JobDev& J, uint32_t rays) __attribute__((always_inline)) {
        if (LHITS > 0)
            for (int i = lane; i < LHITS; i += 64) {
                const uint32_t h = wh[i];
                if (h) {
                    wh[i] = 0;
                    if (J.hits) gadd(J.hits + i, (unsigned long long)h);
                }
            }
        uint32_t v = lane < kNumCounters ? wc[lane] : 0u;
        if (lane == ST_STAT_RAYS) v = rays;
        if (lane == ST_STAT_SEGMENTS) v = rays + wc[ST_STAT_MIRROR_HITS] - wc[ST_STAT_ABSORBED];
        if (lane == ST_STAT_ESCAPED && !(J.flags & ST_TRACE_COUNT_WORK))
            v = rays - wc[ST_STAT_ABSORBED] - wc[ST_STAT_BACK_FACE] - wc[ST_STAT_TRUNCATED] - wc[ST_STAT_RECEIVER_HITS];
        if (lane < kNumCounters) wc[lane] = 0;
        if (v) {
            if (lane < kAggCounter) {
                if (J.stats) gadd(J.stats + lane, (unsigned long long)v);
            } else if (lane == kAggCounter && J.hits && J.agg_slot >= 0) {
                gadd(J.hits + meta_elem(ld1(J.meta + J.agg_slot)), (unsigned long long)v);
            }
        }
    };
    // A launch of one job keeps its counts in the wave's LDS across the wave's
    // chunks and flushes them every kFlushRays rays (a ray adds at most a few
    // events: the u32 counters stay far from overflow) and when the wave ends,
    // instead of at every chunk: each flush is ~10 device atomics on the same
    // 16 addresses from every wave. Launches of several jobs, and jobs that
    // count work (tens of node visits and tests per segment), flush per chunk.
    constexpr uint32_t kFlushRays = 1u << 22;
    uint32_t wave_rays = 0;  // rays of the single job since the wave's last flush (wave-uniform)
    auto run_chunk = [&](const JobDev& J, uint64_t cj) __attribute__((always_inline)) {
        const ST_CONST SunFrame& S = *(const ST_CONST SunFrame*)J.sun;
        const uint64_t first = cj * (uint64_t)chunk_rays;
        const uint64_t end = min(first + (uint64_t)chunk_rays, J.ray_count);
        uint32_t len = 0;  // queued form: entries in the wave's log (wave-uniform)
        if (QUEUED) {
            float* q = lds_queue[threadIdx.x >> 6];
            uint64_t* bases = lds_bases[TABLE ? threadIdx.x >> 6 : 0];
            uint32_t* cells = lds_cells[TABLE ? threadIdx.x >> 6 : 0];
            const bool table = TABLE && J.order_kind != kOrderHalton;
            if (table && (uint32_t)lane < (kMaxChunkRays / 64)) {
                const uint64_t s0 = first + 64u * (uint32_t)lane;
                if (s0 < end) {
                    uint32_t c0;
                    bases[lane] = group_of_slot(J, S, s0, &c0);
                    cells[lane] = c0;
                }
            }
            // ray index of the slot at offset u of this chunk
            auto chunk_ray = [&](uint32_t u) -> uint64_t {
                return table ? bases[u >> 6] + (u & 63u) : ray_of_slot(J, first + u);
            };
            uint32_t qn = 0;  // wave-uniform queue length, < kQueueCap between steps
            bool alive = false;  // this lane holds a ray with a next segment
            V3 o = v3(0.f, 0.f, 0.f), d = o;
            uint32_t off = 0;    // its offset in the chunk
            int depth = 0, key = -1;
            int pend = 0;        // the mirror it hit last, for shade_mirror
            // lanes without a ray take parked ones (up to `want` in total)
            auto refill = [&](uint32_t want) {
                const unsigned long long dead = __ballot(!alive);
                const uint32_t take = min(qn, want);
                if (!alive) {
                    const uint32_t r = lane_rank(dead);
                    if (r < take) {
                        const uint32_t i = qn - take + r;
                        o = v3(q[0 * kQueueCap + i], q[1 * kQueueCap + i], q[2 * kQueueCap + i]);
                        d = v3(q[3 * kQueueCap + i], q[4 * kQueueCap + i], q[5 * kQueueCap + i]);
                        const uint32_t tag = __float_as_uint(q[6 * kQueueCap + i]);
                        off = tag & 0x7FFu;
                        depth = (int)((tag >> 11) & 15u);
                        key = (int)(tag >> 15) - 1;
                        pend = (int)__float_as_uint(q[7 * kQueueCap + i]);
                        alive = true;
                    }
                }
                qn -= take;
            };
            // sends the log when the next step could overfill it
            auto flush_if_full = [&]() {
                if (len > kLogCap - kLogHeadroom) {
                    log_send(wl, J, len);
                    len = 0;
                }
            };
            // the next segment of every held ray. QUEUED == 2: a segment
            // without a candidate list walks the BVH (tens of nodes) where a
            // list walk takes a few tests, so a step traces either list
            // segments or BVH segments; the BVH ones wait (keeping their rays)
            // until half the wave, or every held ray, needs the BVH.
            auto step = [&]() {
                flush_if_full();
                uint32_t ph = kNoEntry, pf = kNoEntry;
                if constexpr (QUEUED == 2) {
                    const SegList L = alive ? next_list(J, depth, key) : no_list();
                    const bool bvh = alive && L.lb > L.le;
                    const unsigned long long bm = __ballot(bvh);
                    if (alive && bvh == (__popcll(bm) >= 32 || bm == __ballot(alive))) {
                        const uint64_t ray = chunk_ray(off);
                        int last = depth;
                        float tmin = kTminBounce;
                        alive = shade_mirror(J, ray, pend, depth, o, d, wc) &&
                                trace_segment<SS, true>(J, ray, stk, wc, wh, wl, o, d, tmin, depth, last, L, key,
                                                           pend, ph, pf);
                        if (!alive) finish_ray(J, ray, last);
                    }
                } else {
                    if (alive) {
                        const uint64_t ray = chunk_ray(off);
                        int last = depth;
                        float tmin = kTminBounce;
                        alive = shade_mirror(J, ray, pend, depth, o, d, wc) &&
                                trace_segment<SS, true>(J, ray, stk, wc, wh, wl, o, d, tmin, depth, last,
                                                           next_list(J, depth, key), key, pend, ph, pf);
                        if (!alive) finish_ray(J, ray, last);
                    }
                }
                log_pending(wl, len, J, ph, pf);
            };
            for (uint64_t b = first; b < end; b += 64) {
                flush_if_full();
                const uint64_t i = b + (uint64_t)lane;
                alive = false;
                uint32_t ph = kNoEntry, pf = kNoEntry;
                // the group's first index and cell from the chunk's table (b is a
                // multiple of 64 slots)
                const uint32_t gi = (uint32_t)(b - first) >> 6;
                const uint64_t base = table ? bases[gi] : 0;
                const uint32_t hint = table ? cells[gi] : 0xFFFFFFFFu;
                if (i < end) {
                    const uint64_t ray = table ? base + (uint64_t)lane : ray_of_slot(J, i);
                    uint32_t cell;
                    sun_ray(J, S, ray, wc, o, d, cell, hint);
                    float tmin = kTminPrimary;
                    int last = 0;
                    depth = 0;
                    key = -1;
                    off = (uint32_t)(i - first);
                    alive = trace_segment<SS, true>(J, ray, stk, wc, wh, wl, o, d, tmin, depth, last,
                                                       cell_list(J, cell), key, pend, ph, pf);
                    if (!alive) finish_ray(J, ray, last);
                }
                log_pending(wl, len, J, ph, pf);
                // full waves while held and parked rays fill one
                for (;;) {
                    const uint32_t na = (uint32_t)__popcll(__ballot(alive));
                    if (na + qn < 64u) break;
                    refill(64u - na);
                    step();
                }
                // park the rest (fewer than 64 in all)
                const unsigned long long m = __ballot(alive);
                if (alive) {
                    const uint32_t i2 = qn + lane_rank(m);
                    q[0 * kQueueCap + i2] = o.x;
                    q[1 * kQueueCap + i2] = o.y;
                    q[2 * kQueueCap + i2] = o.z;
                    q[3 * kQueueCap + i2] = d.x;
                    q[4 * kQueueCap + i2] = d.y;
                    q[5 * kQueueCap + i2] = d.z;
                    q[6 * kQueueCap + i2] = __uint_as_float(queue_tag(off, depth, key));
                    q[7 * kQueueCap + i2] = __uint_as_float((uint32_t)pend);
                }
                qn += (uint32_t)__popcll(m);
                alive = false;
            }
            // chunk end: partial waves until every ray is done
            while (qn > 0 || __ballot(alive)) {
                refill(64u);
                step();
            }
        } else {
            for (uint64_t i0 = first; i0 < end; i0 += 64) {
                // a batch of 64 whole rays can add more entries than the log
                // holds (two per segment): flush when half full; entries that
                // find it full go out as direct atomics
                if (wl[0] > kLogCap / 2) log_flush(wl, J);
                const uint64_t i = i0 + (uint64_t)lane;
                const uint64_t base = J.order_kind == kOrderHalton ? 0 : ray_of_slot(J, i0);
                if (i >= end) continue;
                const uint64_t ray = J.order_kind == kOrderHalton ? ray_of_slot(J, i) : base + (uint64_t)lane;
                V3 o, d;
                uint32_t cell;
                sun_ray(J, S, ray, wc, o, d, cell);
                float tmin = kTminPrimary;
                int depth = 0, last = 0, key = -1, pend = 0;
                uint32_t ph = kNoEntry, pf = kNoEntry;  // unused: this form logs as it goes
                if (trace_segment<SS, false>(J, ray, stk, wc, wh, wl, o, d, tmin, depth, last, cell_list(J, cell),
                                                key, pend, ph, pf))
                    while (shade_mirror(J, ray, pend, depth, o, d, wc) &&
                           trace_segment<SS, false>(J, ray, stk, wc, wh, wl, o, d, tmin, depth, last,
                                                       next_list(J, depth, key), key, pend, ph, pf)) {
                    }
                finish_ray(J, ray, last);
            }
        }
        // the chunk's log goes out (same wave wrote it: LDS order suffices)
        if (QUEUED) log_send(wl, J, len);
        else log_flush(wl, J);
        const uint32_t rays = (uint32_t)(end - first);
        if (!jobs && !(J.flags & ST_TRACE_COUNT_WORK)) {
            wave_rays += rays;
            if (wave_rays >= kFlushRays) {
                flush_counts(J, wave_rays);
                wave_rays = 0;
            }
        } else {
            flush_counts(J, rays);
        }
    };
    for (bool first_pass = true;; first_pass = false) {
        // static_chunks: every wave of the (fully resident) grid traces chunk
        // wave_id only, and the shared counter is not touched
        uint64_t c = wave_id;
        if (static_chunks) {
            if (!first_pass) break;
        } else {
            uint32_t f = 0;
            if (lane == 0) f = gadd32(&work_counter[0], 1u);
            c = __builtin_amdgcn_readfirstlane(f);
        }
        if (c >= total_chunks) break;
        if (jobs) {
            int lo = 0, hi = n_jobs;  // chunk_prefix[lo] <= c < chunk_prefix[hi]
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (chunk_prefix[mid] <= c) lo = mid;
                else hi = mid;
            }
            lo = __builtin_amdgcn_readfirstlane(lo);
            run_chunk(*(const JobDev*)(const ST_CONST JobDev*)(jobs + lo), c - chunk_prefix[lo]);
        } else {
            // `single` is the first argument: offset 0 of the kernel argument
            // segment. Read there through an opaque pointer, its fields are
            // loaded at their uses as the job records' are; named directly, the
            // compiler hoists every field into scalar registers, more than fit,
            // and reloads the spilled ones with vector instructions (its
            // address would be that of a private copy).
            const ST_CONST JobDev* sp = (const ST_CONST JobDev*)__builtin_amdgcn_kernarg_segment_ptr();
            asm volatile("" : "+s"(sp));
            run_chunk(*(const JobDev*)sp, c);
        }
    }
    if (!jobs && wave_rays > 0) {
        const ST_CONST JobDev* sp = (const ST_CONST JobDev*)__builtin_amdgcn_kernarg_segment_ptr();
        asm volatile("" : "+s"(sp));
        flush_counts(*(const JobDev*)sp, wave_rays);
    }
    if (!static_chunks && lane == 0) {
        __threadfence();
        if (gadd32(&work_counter[1], 1u) == waves - 1u) {
            atomicExch(&work_counter[0], 0u);
            atomicExch(&work_counter[1], 0u);
        }
    }
}

// ---------------------------------------------------------------------------
__global__ void k_frames(const st_element_f64* __restrict__ elems, int32_t n, const int32_t* __restrict__ elem_slot,
                         const double* __restrict__ aim_override, const double* __restrict__ zrot_override,
                         double* __restrict__ frames, double* __restrict__ ref_aabb, float* __restrict__ bounds_slot,
                         float4* __restrict__ prims, uint32_t* __restrict__ meta, float4* __restrict__ optics) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const st_element_f64 E = elems[e];
    double aim[3] = {E.aim_point[0], E.aim_point[1], E.aim_point[2]};
    if (aim_override) for (int k = 0; k < 3; ++k) aim[k] = aim_override[3 * e + k];
    const double zrot = zrot_override ? zrot_override[e] : E.zrot_deg;
    double l2g[9];
    element_frame(E, aim, zrot, l2g);
    if (frames) for (int k = 0; k < 9; ++k) frames[9 * (size_t)e + k] = l2g[k];
    double lo[3], hi[3];
    element_ref_aabb(E, l2g, lo, hi);
    for (int k = 0; k < 3; ++k) {
        ref_aabb[6 * (size_t)e + k] = lo[k];
        ref_aabb[6 * (size_t)e + 3 + k] = hi[k];
    }
    const int slot = elem_slot[e];
    float blo[3], bhi[3];
    element_bounds(E, l2g, blo, bhi);
    for (int k = 0; k < 3; ++k) {
        bounds_slot[6 * (size_t)slot + k] = blo[k];
        bounds_slot[6 * (size_t)slot + 3 + k] = bhi[k];
    }
    float4 q[4];
    element_pack(E, l2g, q);
    for (int k = 0; k < 4; ++k) prims[4 * (size_t)slot + k] = q[k];
    if (meta) meta[slot] = meta_pack((uint32_t)e, element_kind(E), E.role == ST_ROLE_RECEIVER ? 1u : 0u);
    if (optics) optics[e] = make_float4((float)E.reflectivity, (float)E.slope_error, (float)E.spec_error, 0.f);
}

__global__ void k_frames_batch(const st_element_f64* __restrict__ elems, int32_t n,
                               const int32_t* __restrict__ elem_slot, int32_t steps,
                               const double* __restrict__ sun_units, const double* __restrict__ target,
                               double* __restrict__ ref_aabb, float* __restrict__ bounds_slot,
                               float4* __restrict__ prims, double* __restrict__ frames) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (int64_t)n * steps) return;
    const int k = (int)(g / n), e = (int)(g % n);
    const st_element_f64 E = elems[e];
    double aim[3] = {E.aim_point[0], E.aim_point[1], E.aim_point[2]};
    if (E.role == ST_ROLE_MIRROR) {
        const double s[3] = {sun_units[3 * k], sun_units[3 * k + 1], sun_units[3 * k + 2]};
        bisector_aim_point(E.origin, s, target, aim);
    }
    double l2g[9];
    element_frame(E, aim, E.zrot_deg, l2g);
    if (frames)
        for (int j = 0; j < 9; ++j) frames[((size_t)k * n + e) * 9 + j] = l2g[j];
    double lo[3], hi[3];
    element_ref_aabb(E, l2g, lo, hi);
    double* ra = ref_aabb + ((size_t)k * n + e) * 6;
    for (int j = 0; j < 3; ++j) {
        ra[j] = lo[j];
        ra[3 + j] = hi[j];
    }
    const int slot = elem_slot[e];
    float blo[3], bhi[3];
    element_bounds(E, l2g, blo, bhi);
    float* bs = bounds_slot + ((size_t)k * n + slot) * 6;
    for (int j = 0; j < 3; ++j) {
        bs[j] = blo[j];
        bs[3 + j] = bhi[j];
    }
    float4 q[4];
    element_pack(E, l2g, q);
    float4* pq = prims + ((size_t)k * n + slot) * 4;
    for (int j = 0; j < 4; ++j) pq[j] = q[j];
}

// FP64 bisector aim of every mirror at target (receivers keep their aim point)
__global__ void k_aim(st_element_f64* __restrict__ elems, int32_t n, const double* __restrict__ sun_unit,
                      const double* __restrict__ target) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n || elems[e].role != ST_ROLE_MIRROR) return;
    const double s[3] = {sun_unit[0], sun_unit[1], sun_unit[2]};
    double aim[3];
    bisector_aim_point(elems[e].origin, s, target, aim);
    for (int k = 0; k < 3; ++k) elems[e].aim_point[k] = aim[k];
}

__global__ void k_aim_copy(const st_element_f64* __restrict__ src, st_element_f64* __restrict__ dst, int32_t n,
                           const double* __restrict__ sun_unit, const double* __restrict__ target) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    st_element_f64 E = src[e];
    if (E.role == ST_ROLE_MIRROR) {
        const double s[3] = {sun_unit[0], sun_unit[1], sun_unit[2]};
        bisector_aim_point(E.origin, s, target, E.aim_point);
    }
    dst[e] = E;
}

__device__ __forceinline__ void child_box(int32_t code, const float4* __restrict__ nodes,
                                          const float* __restrict__ bounds, float lo[3], float hi[3]) {
    if (code >= 0) {
        const float4 n0 = nodes[4 * code + 0], n1 = nodes[4 * code + 1], n2 = nodes[4 * code + 2];
        lo[0] = fminf(n0.x, n1.x);
        hi[0] = fmaxf(n0.y, n1.y);
        lo[1] = fminf(n0.z, n1.z);
        hi[1] = fmaxf(n0.w, n1.w);
        lo[2] = fminf(n2.x, n2.z);
        hi[2] = fmaxf(n2.y, n2.w);
    } else {
        const int32_t s = leaf_start(code), c = leaf_count(code);
        for (int k = 0; k < 3; ++k) {
            lo[k] = 3.0e38f;
            hi[k] = -3.0e38f;
        }
        for (int i = 0; i < c; ++i)
            for (int k = 0; k < 3; ++k) {
                lo[k] = fminf(lo[k], bounds[6 * (s + i) + k]);
                hi[k] = fmaxf(hi[k], bounds[6 * (s + i) + 3 + k]);
            }
    }
}

__global__ void k_refit(const int32_t* __restrict__ child, const int32_t* __restrict__ level_nodes, int32_t begin,
                        int32_t count, int32_t n_nodes, int32_t n_prims, int32_t steps,
                        const float* __restrict__ bounds_slot, float4* __restrict__ nodes) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (int64_t)count * steps) return;
    const int k = (int)(g / count), i = (int)(g % count);
    const int32_t node = level_nodes[begin + i];
    float4* N = nodes + (size_t)k * n_nodes * 4;
    const float* B = bounds_slot + (size_t)k * n_prims * 6;
    const int32_t c0 = child[2 * node], c1 = child[2 * node + 1];
    float l0[3], h0[3], l1[3], h1[3];
    child_box(c0, N, B, l0, h0);
    child_box(c1, N, B, l1, h1);
    N[4 * node + 0] = make_float4(l0[0], h0[0], l0[1], h0[1]);
    N[4 * node + 1] = make_float4(l1[0], h1[0], l1[1], h1[1]);
    N[4 * node + 2] = make_float4(l0[2], h0[2], l1[2], h1[2]);
    N[4 * node + 3] = make_float4(__int_as_float(c0), __int_as_float(c1), 0.f, 0.f);
}

// The whole refit in one launch: one block per step walks the levels bottom-up,
// a barrier between levels (a level reads the boxes the level below wrote, in
// this block). For BVHs of at most kRefitOneLaunchNodes nodes and
// kRefitMaxLevels levels; larger ones take k_refit, one launch per level.
constexpr int32_t kRefitMaxLevels = 32, kRefitOneLaunchNodes = 1 << 16;
struct RefitLevels {
    int32_t offset[kRefitMaxLevels + 1];  // level l's nodes: level_nodes[offset[l], offset[l + 1])
};
__global__ __launch_bounds__(1024) void k_refit_levels(const int32_t* __restrict__ child,
                                                       const int32_t* __restrict__ level_nodes, RefitLevels lv,
                                                       int32_t levels, int32_t n_nodes, int32_t n_prims,
                                                       const float* __restrict__ bounds_slot, float4* __restrict__ nodes) {
    const int k = blockIdx.x;
    float4* N = nodes + (size_t)k * n_nodes * 4;
    const float* B = bounds_slot + (size_t)k * n_prims * 6;
    for (int l = levels - 1; l >= 0; --l) {
        for (int32_t i = lv.offset[l] + (int32_t)threadIdx.x; i < lv.offset[l + 1]; i += (int32_t)blockDim.x) {
            const int32_t node = level_nodes[i];
            const int32_t c0 = child[2 * node], c1 = child[2 * node + 1];
            float l0[3], h0[3], l1[3], h1[3];
            child_box(c0, N, B, l0, h0);
            child_box(c1, N, B, l1, h1);
            N[4 * node + 0] = make_float4(l0[0], h0[0], l0[1], h0[1]);
            N[4 * node + 1] = make_float4(l1[0], h1[0], l1[1], h1[1]);
            N[4 * node + 2] = make_float4(l0[2], h0[2], l1[2], h1[2]);
            N[4 * node + 3] = make_float4(__int_as_float(c0), __int_as_float(c1), 0.f, 0.f);
        }
        __syncthreads();
    }
}

// FP64 sun plane, one block per step (lib/sun_utils.cu:53-179 + geometry_manager.cpp:97-148)
__global__ __launch_bounds__(256) void k_sun_plane(const double* __restrict__ ref_aabb, int32_t n,
                                                   const double* __restrict__ sun_units, st_sun_desc sun,
                                                   SunFrame* __restrict__ out) {
    __shared__ double red[5][256];
    const int k = blockIdx.x, tid = threadIdx.x;
    const double s[3] = {sun_units[3 * k], sun_units[3 * k + 1], sun_units[3 * k + 2]};
    const double angle = sun.shape == ST_SUN_GAUSSIAN ? 4.0 * sun.param : sun.param;
    const double tana = tan(angle);
    double ax[3] = {fabs(s[0]) < 0.9 ? 1.0 : 0.0, fabs(s[0]) < 0.9 ? 0.0 : 1.0, 0.0};
    double u[3] = {ax[1] * s[2] - ax[2] * s[1], ax[2] * s[0] - ax[0] * s[2], ax[0] * s[1] - ax[1] * s[0]};
    double lu = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    for (int j = 0; j < 3; ++j) u[j] /= lu;
    double v[3] = {s[1] * u[2] - s[2] * u[1], s[2] * u[0] - s[0] * u[2], s[0] * u[1] - s[1] * u[0]};
    double lv = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    for (int j = 0; j < 3; ++j) v[j] /= lv;
    double dmax = 0.0, umin = 1e300, umax = -1e300, vmin = 1e300, vmax = -1e300;
    const double* A = ref_aabb + (size_t)k * n * 6;
    for (int e = tid; e < n; e += blockDim.x) {
        for (int c = 0; c < 8; ++c) {
            const double p[3] = {A[6 * e + ((c & 1) ? 3 : 0)], A[6 * e + 1 + ((c & 2) ? 3 : 0)],
                                 A[6 * e + 2 + ((c & 4) ? 3 : 0)]};
            const double ps = fabs(p[0] * s[0] + p[1] * s[1] + p[2] * s[2]);
            dmax = fmax(dmax, ps);
            const double buf = ps * tana;
            const double pu = p[0] * u[0] + p[1] * u[1] + p[2] * u[2];
            const double pv = p[0] * v[0] + p[1] * v[1] + p[2] * v[2];
            umin = fmin(umin, pu - buf);
            umax = fmax(umax, pu + buf);
            vmin = fmin(vmin, pv - buf);
            vmax = fmax(vmax, pv + buf);
        }
    }
    red[0][tid] = dmax;
    red[1][tid] = umin;
    red[2][tid] = -umax;
    red[3][tid] = vmin;
    red[4][tid] = -vmax;
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if (tid < st) {
            red[0][tid] = fmax(red[0][tid], red[0][tid + st]);
            for (int j = 1; j < 5; ++j) red[j][tid] = fmin(red[j][tid], red[j][tid + st]);
        }
        __syncthreads();
    }
    if (tid != 0) return;
    const double d = red[0][0];
    umin = red[1][0];
    umax = -red[2][0];
    vmin = red[3][0];
    vmax = -red[4][0];
    SunFrame F;
    for (int j = 0; j < 3; ++j) {
        F.v0[j] = umin * u[j] + vmin * v[j] + d * s[j];
        F.v1[j] = umax * u[j] + vmin * v[j] + d * s[j];
        F.v2[j] = umax * u[j] + vmax * v[j] + d * s[j];
        F.v3[j] = umin * u[j] + vmax * v[j] + d * s[j];
    }
    double e1[3], e2[3];
    for (int j = 0; j < 3; ++j) {
        e1[j] = F.v1[j] - F.v0[j];
        e2[j] = F.v3[j] - F.v0[j];
    }
    for (int j = 0; j < 3; ++j) F.sun[j] = s[j];
    F.distance = d;
    F.area = sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]) * sqrt(e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2]);
    F.box_half_angle = angle;
    F.dni = sun.dni > 0 ? sun.dni : 1000.0;
    // sampling basis around w = -s (cuda/sun.cu:76-80), FP64 then rounded
    const double w[3] = {-s[0], -s[1], -s[2]};
    const double a2[3] = {fabs(w[0]) > 0.99 ? 0.0 : 1.0, fabs(w[0]) > 0.99 ? 1.0 : 0.0, 0.0};
    double su[3] = {a2[1] * w[2] - a2[2] * w[1], a2[2] * w[0] - a2[0] * w[2], a2[0] * w[1] - a2[1] * w[0]};
    const double lsu = sqrt(su[0] * su[0] + su[1] * su[1] + su[2] * su[2]);
    for (int j = 0; j < 3; ++j) su[j] /= lsu;
    const double sv[3] = {w[1] * su[2] - w[2] * su[1], w[2] * su[0] - w[0] * su[2], w[0] * su[1] - w[1] * su[0]};
    for (int j = 0; j < 3; ++j) {
        F.o[j] = (float)F.v0[j];
        F.e1[j] = (float)e1[j];
        F.e2[j] = (float)e2[j];
        F.w[j] = (float)w[j];
        F.u[j] = (float)su[j];
        F.v[j] = (float)sv[j];
    }
    const double hs = sin(0.5 * sun.param);
    F.vers = (float)(2.0 * hs * hs);
    F.sigma = (float)sun.param;
    F.shape = sun.shape;
    F.position_sampling = sun.position_sampling;
    F.grid_rows = sun.position_sampling == ST_POS_HALTON ? kHaltonRows : kGridRes;
    F.reserved = 0;
    out[k] = F;
}

// ---------------------------------------------------------------------------
// kGridParts threads per (step, primitive): part p visits every kGridParts-th
// cell of the primitive's footprint, so large footprints do not serialize
constexpr int kGridParts = 16;
// bounds and prims of step k at k * n * 6 floats and k * n * 4 float4; meta shared
__global__ void k_grid_count(const float* __restrict__ bounds, const float4* __restrict__ prims,
                             const uint32_t* __restrict__ meta, int32_t n, int32_t steps,
                             const SunFrame* __restrict__ sun, double tan_max, uint32_t* __restrict__ counts) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (int64_t)n * steps * kGridParts) return;
    const int part = (int)(g % kGridParts);
    const int64_t e = g / kGridParts;
    const int k = (int)(e / n), slot = (int)(e % n);
    Footprint fp;
    if (!grid_range(bounds + (size_t)k * n * 6, prims + (size_t)k * n * 4, meta, slot, sun[k], tan_max, fp)) return;
    const int* r = fp.r;
    uint32_t* C = counts + (size_t)k * kGridCells;
    const int w = r[1] - r[0] + 1, cells = w * (r[3] - r[2] + 1);
    for (int i = part; i < cells; i += kGridParts) {
        const int ix = r[0] + i % w, iy = r[2] + i / w;
        if (cell_meets(fp, ix, iy)) atomicAdd(&C[iy * kGridRes + ix], 1u);
    }
}

// exclusive prefix sum of the kGridRes^2 counts of step blockIdx.x (one block
// of 1024 per step); start[cells] = total, stats[0] = total, stats[1] = longest
// list. With jobs: the step's job gets the grid when its lists fit item_cap and
// are short enough to beat the BVH, else none (its sun rays traverse the BVH).
__global__ __launch_bounds__(1024) void k_grid_scan(const uint32_t* __restrict__ counts_all,
                                                    uint32_t* __restrict__ start_all, uint32_t* __restrict__ stats_all,
                                                    JobDev* __restrict__ jobs, const int32_t* __restrict__ items_all,
                                                    uint32_t item_cap) {
    // each of the 16 waves owns 4,096 consecutive cells: their sum and longest
    // list (coalesced 16-byte loads), then, offset by the waves before it, their
    // prefix sums (a wave-wide scan of each lane's four counts per 256 cells)
    constexpr int kWaves = 16, kSeg = kGridCells / kWaves, kIters = kSeg / 256;
    static_assert(kGridCells % (kWaves * 256) == 0, "grid cells per wave");
    __shared__ uint32_t wsum[kWaves], wmax[kWaves];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6, k = blockIdx.x;
    const uint4* counts = (const uint4*)(counts_all + (size_t)k * kGridCells + (size_t)wv * kSeg);
    uint32_t* start = start_all + (size_t)k * (kGridCells + 1);
    uint32_t sum = 0, mx = 0;
    for (int i = 0; i < kIters; ++i) {
        const uint4 c = counts[i * 64 + lane];
        sum += c.x + c.y + c.z + c.w;
        mx = max(mx, max(max(c.x, c.y), max(c.z, c.w)));
    }
    for (int d = 32; d > 0; d >>= 1) {
        sum += __shfl_xor(sum, d);
        mx = max(mx, (uint32_t)__shfl_xor(mx, d));
    }
    if (lane == 0) {
        wsum[wv] = sum;
        wmax[wv] = mx;
    }
    __syncthreads();
    uint32_t carry = 0, total = 0, longest = 0;
    for (int j = 0; j < kWaves; ++j) {
        if (j < wv) carry += wsum[j];
        total += wsum[j];
        longest = max(longest, wmax[j]);
    }
    uint32_t* out = start + (size_t)wv * kSeg;
    for (int i = 0; i < kIters; ++i) {
        const uint4 c = counts[i * 64 + lane];
        const uint32_t own = c.x + c.y + c.z + c.w;
        uint32_t inc = own;  // inclusive scan over the wave's lanes
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t v = __shfl_up(inc, d);
            if (lane >= d) inc += v;
        }
        const uint32_t ex = carry + inc - own;
        uint32_t* o = out + (size_t)(i * 64 + lane) * 4;
        o[0] = ex;
        o[1] = ex + c.x;
        o[2] = ex + c.x + c.y;
        o[3] = ex + c.x + c.y + c.z;
        carry += __shfl(inc, 63);
    }
    if (t == 0) {
        start[kGridCells] = total;
        stats_all[2 * k] = total;
        stats_all[2 * k + 1] = longest;
        if (jobs) {
            const bool ok = grid_lists_pay(total, longest) && total <= item_cap;
            jobs[k].grid_start = ok ? start : nullptr;
            jobs[k].grid_items = ok ? items_all + (size_t)k * item_cap : nullptr;
        }
    }
}

// items of step k at items_all + k * item_cap; a step whose lists overflow
// item_cap is skipped (its job gets no grid)
__global__ void k_grid_fill(const float* __restrict__ bounds, const float4* __restrict__ prims,
                            const uint32_t* __restrict__ meta, int32_t n, int32_t steps,
                            const SunFrame* __restrict__ sun, double tan_max, const uint32_t* __restrict__ start_all,
                            uint32_t* __restrict__ cursor_all, int32_t* __restrict__ items_all, uint32_t item_cap) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (int64_t)n * steps * kGridParts) return;
    const int part = (int)(g % kGridParts);
    const int64_t e = g / kGridParts;
    const int k = (int)(e / n), slot = (int)(e % n);
    const uint32_t* start = start_all + (size_t)k * (kGridCells + 1);
    if (start[kGridCells] > item_cap) return;
    Footprint fp;
    if (!grid_range(bounds + (size_t)k * n * 6, prims + (size_t)k * n * 4, meta, slot, sun[k], tan_max, fp)) return;
    const int* r = fp.r;
    uint32_t* cursor = cursor_all + (size_t)k * kGridCells;
    int32_t* items = items_all + (size_t)k * item_cap;
    const int w = r[1] - r[0] + 1, cells = w * (r[3] - r[2] + 1);
    for (int i = part; i < cells; i += kGridParts) {
        const int ix = r[0] + i % w, iy = r[2] + i / w;
        if (!cell_meets(fp, ix, iy)) continue;
        const int c = iy * kGridRes + ix;
        items[start[c] + atomicAdd(&cursor[c], 1u)] = slot;
    }
}

// ---------------------------------------------------------------------------
// Candidate lists of first reflections. A sun ray reflected off mirror e
// leaves from e's box B_e in a direction within phi of a = reflect(-sun, n_e):
// phi = sun-cone angle + 2 * (slope-error bound + surface-normal spread) +
// specular-error bound, each error bounded by sigma * sqrt(-2 ln 2^-24) =
// 5.7686 sigma (the smallest uniform). Such a ray can reach a box B only if the
// cone {t u : t >= 0, angle(u, a) <= phi} meets B - B_e (Minkowski
// difference); the cone is covered by the boxes of 32 pieces [t_i, t_i+1]
// (geometric in t) and tested against the BVH, so the list holds every
// primitive the ray can hit. Elements whose cone is wide (cylinders, large
// errors) or whose list exceeds kMaxBounceList get none.
__device__ __forceinline__ bool bounce_cone(const st_element_f64& E, double theta_sun, double* phi) {
    if (E.role != ST_ROLE_MIRROR || E.surface == ST_SURFACE_CYLINDER) return false;
    double spread = 0.0;
    if (E.surface == ST_SURFACE_PARABOLIC) {
        const double sx = E.curv_x * 0.5 * E.dim_x, sy = E.curv_y * 0.5 * E.dim_y;
        spread = atan(sqrt(sx * sx + sy * sy));
    }
    const double p = (theta_sun + 2.0 * (5.7686 * E.slope_error + spread) + 5.7686 * E.spec_error) * 1.001 + 1e-5;
    *phi = p;
    return p < 0.35;
}

// whether a ray leaving box hb in a direction within the cone C can reach the
// box [lo, hi]
__device__ __forceinline__ bool cone_reaches(const Cone& C, const float* hb, const float lo[3], const float hi[3]) {
    const double dlo[3] = {(double)lo[0] - hb[3], (double)lo[1] - hb[4], (double)lo[2] - hb[5]};
    const double dhi[3] = {(double)hi[0] - hb[0], (double)hi[1] - hb[1], (double)hi[2] - hb[2]};
    return cone_meets_box(C, dlo, dhi);
}

// Primitives whose boxes a ray leaving box hb in a direction within the cone C
// can reach, found by a BVH walk; the first `cap` are written to items[0 ...].
// Returns the count, or cap + 1 as soon as it exceeds cap. stack: this lane's
// column of an LDS array, entry i at stack[i * STRIDE].
template <int STRIDE>
__device__ __forceinline__ uint32_t cone_walk(const float4* __restrict__ nodes, const float* __restrict__ bounds_slot,
                                              const float* hb, const Cone& C, int32_t* __restrict__ items,
                                              uint32_t cap, int32_t* __restrict__ stack) {
    auto meets = [&](const float lo[3], const float hi[3]) __attribute__((always_inline)) {
        return cone_reaches(C, hb, lo, hi);
    };
    uint32_t count = 0;  // the stack holds <= one pending sibling per level (depth <= 31)
    int sp = 0;
    int32_t node = 0;
    for (;;) {
        const float4 n0 = nodes[4 * node + 0], n1 = nodes[4 * node + 1], n2 = nodes[4 * node + 2], n3 = nodes[4 * node + 3];
        const float lo0[3] = {n0.x, n0.z, n2.x}, hi0[3] = {n0.y, n0.w, n2.y};
        const float lo1[3] = {n1.x, n1.z, n2.z}, hi1[3] = {n1.y, n1.w, n2.w};
        const int32_t ch0 = __float_as_int(n3.x), ch1 = __float_as_int(n3.y);
        const bool hit0 = meets(lo0, hi0), hit1 = ch1 != ch0 && meets(lo1, hi1);  // a lone primitive sits under both
        // an inner child is pushed; a leaf child's primitives are tested now
        auto visit = [&](int32_t ch) __attribute__((always_inline)) -> bool {
            if (ch >= 0) {
                stack[(sp++) * STRIDE] = ch;
                return true;
            }
            const int32_t st = leaf_start(ch), cnt = leaf_count(ch);
            for (int i = 0; i < cnt; ++i) {
                const float* b = bounds_slot + 6 * (size_t)(st + i);
                if (!meets(b, b + 3)) continue;
                if (count == cap) return false;
                items[count++] = st + i;
            }
            return true;
        };
        if (hit0 && !visit(ch0)) return cap + 1;
        if (hit1 && !visit(ch1)) return cap + 1;
        if (sp == 0) break;
        node = stack[(--sp) * STRIDE];
    }
    return count;
}

// reflection axis reflect(-sun, n_e) of a sun ray off element e (unit)
__device__ __forceinline__ void sun_reflection_axis(const double* __restrict__ frames, int e, const double sun[3], double a[3]) {
    const double nz[3] = {frames[9 * (size_t)e + 2], frames[9 * (size_t)e + 5], frames[9 * (size_t)e + 8]};
    const double din[3] = {-sun[0], -sun[1], -sun[2]};
    const double dn = din[0] * nz[0] + din[1] * nz[1] + din[2] * nz[2];
    for (int k = 0; k < 3; ++k) a[k] = din[k] - 2.0 * dn * nz[k];
    const double la = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    for (int k = 0; k < 3; ++k) a[k] /= la;
}

// Reflection-list candidates of a flat mirror culled against its rectangle
// (rect_sweep_misses on the candidate's box; for a flat candidate also
// rect_sweep_misses_rect on its own rectangle, element frames of this scene
// state); returns the number kept, compacted in place.
__device__ __forceinline__ uint32_t cull_by_rectangle(const st_element_f64* __restrict__ elems,
                                                      const st_element_f64& E, const double* __restrict__ frames,
                                                      const double* __restrict__ f, const uint32_t* __restrict__ meta,
                                                      const Cone& C, double phi, const float* __restrict__ bounds_slot,
                                                      int32_t* items, uint32_t len) {
    RectSweep S;
    rect_sweep_init(S, E.origin, 0.5 * E.dim_x, 0.5 * E.dim_y, f, C.a, phi);
    uint32_t kept = 0;
    for (uint32_t i = 0; i < len; ++i) {
        const int32_t v = items[i];
        if (rect_sweep_misses(S, bounds_slot + 6 * (size_t)v)) continue;
        const uint32_t m = meta[v];
        if (meta_kind(m) == kKindFlat) {
            const uint32_t e = meta_elem(m);
            RectCand R;
            rect_cand_init(R, elems[e].origin, 0.5 * elems[e].dim_x, 0.5 * elems[e].dim_y, frames + 9 * (size_t)e);
            if (rect_sweep_misses_rect(S, R)) continue;
        }
        items[kept++] = v;
    }
    return kept;
}

// Radius of the sphere around element E's origin that holds its surface in
// every orientation: the distance to the farthest corner of its local box
// (paraboloid sag and cylinder depth included).
__device__ __forceinline__ double element_radius(const st_element_f64& E) {
    const double w = E.dim_x, h = E.dim_y;
    double zlo = 0, zhi = 0;
    if (E.surface == ST_SURFACE_PARABOLIC) {
        const double sx = 0.5 * E.curv_x * (w / 2) * (w / 2), sy = 0.5 * E.curv_y * (h / 2) * (h / 2);
        zlo = fmin(0.0, sx) + fmin(0.0, sy);
        zhi = fmax(0.0, sx) + fmax(0.0, sy);
    } else if (E.surface == ST_SURFACE_CYLINDER) {
        zlo = -w / 2;
        zhi = w / 2;
    }
    return sqrt(0.25 * w * w + 0.25 * h * h + fmax(zlo * zlo, zhi * zhi));
}

// Bounding spheres of the elements (origin, element_radius + 1 mm), for
// k_reaim_lists
__global__ void k_element_spheres(const st_element_f64* __restrict__ elems, int32_t n, double4* __restrict__ out) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const st_element_f64& E = elems[e];
    out[e] = make_double4(E.origin[0], E.origin[1], E.origin[2], element_radius(E) + 1e-3);
}

// First-reflection candidates of mirrors that every state re-aims at `target`
// by the bisector rule. Such a mirror's reflection axis is the direction t
// from its origin to the target whatever the sun, so its first reflections
// leave its sphere (k_element_spheres) in directions within phi (bounce_cone)
// of t, and reach element B only if B's sphere subtends that cone: the angle
// between t and d = c_B - c_A is at most phi + asin(R / |d|), R = r_A + r_B (or
// |d| <= R). The spheres' 1 mm and phi's 1e-6 rad margins exceed the FP32
// rounding of the traced rays. One wave per mirror tests every element, the
// block's spheres staged through LDS; the list (flat candidates first, header
// = their count, as k_cone_lists writes them) goes to scratch[A *
// kMaxBounceList ...], range[A] = its (begin, end), (1, 0) for none or longer
// than kMaxBounceList.
constexpr int kReaimBlock = 1024;
__global__ __launch_bounds__(kReaimBlock) void k_reaim_lists(const st_element_f64* __restrict__ elems, int32_t n,
                                                             const double4* __restrict__ spheres,
                                                             const int32_t* __restrict__ elem_slot,
                                                             const uint32_t* __restrict__ meta,
                                                             const double* __restrict__ target, double theta_sun,
                                                             uint2* __restrict__ range, int32_t* __restrict__ scratch) {
    __shared__ double4 lds_spheres[kReaimBlock];
    __shared__ int32_t lds_items[kReaimBlock / 64][kMaxBounceList];
    const int lane = threadIdx.x & 63;
    const int A = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
    int32_t* items = lds_items[threadIdx.x >> 6];
    double t[3] = {0, 0, 0}, cp = 1.0, sp = 0.0;
    double4 a = make_double4(0, 0, 0, 0);
    bool active = false;  // wave-uniform: A is a mirror with a cone
    if (A < n) {
        const st_element_f64& E = elems[A];
        double phi = 0.0;
        a = spheres[A];
        const double v[3] = {target[0] - a.x, target[1] - a.y, target[2] - a.z};
        const double lv = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        active = bounce_cone(E, theta_sun, &phi) && lv > 0.0;
        if (active) {
            for (int j = 0; j < 3; ++j) t[j] = v[j] / lv;
            cp = cos(phi + 1e-6);
            sp = sin(phi + 1e-6);
        } else if (lane == 0) {
            range[A] = make_uint2(1u, 0u);
        }
    }
    uint32_t count = 0;
    bool over = false;
    for (int b0 = 0; b0 < n; b0 += kReaimBlock) {  // every wave of the block runs every tile
        __syncthreads();
        if (b0 + (int)threadIdx.x < n) lds_spheres[threadIdx.x] = spheres[b0 + threadIdx.x];
        __syncthreads();
        if (!active || over) continue;
        const int tile = min(kReaimBlock, n - b0);
        for (int j0 = 0; j0 < tile; j0 += 64) {
            const int j = j0 + lane;
            bool pass = false;
            if (j < tile) {
                const double4 c = lds_spheres[j];
                const double ca[3] = {a.x, a.y, a.z}, cb[3] = {c.x, c.y, c.z};
                pass = sphere_cone_reaches(ca, a.w, t, cp, sp, cb, c.w);
            }
            const unsigned long long m = __ballot(pass);
            const uint32_t cnt = (uint32_t)__popcll(m);
            if (count + cnt > kMaxBounceList - 1) {
                over = true;
                break;
            }
            if (pass) items[count + lane_rank(m)] = elem_slot[b0 + j];
            count += cnt;
        }
    }
    if (!active) return;
    if (over) {
        if (lane == 0) range[A] = make_uint2(1u, 0u);
        return;
    }
    // flat candidates first
    int32_t* list = scratch + (size_t)A * kMaxBounceList;
    const bool have = (uint32_t)lane < count;
    const int32_t slot = have ? items[lane] : 0;
    const bool flat = have && meta_kind(meta[slot]) == kKindFlat;
    const unsigned long long mf = __ballot(flat), mc = __ballot(have && !flat);
    const uint32_t nf = (uint32_t)__popcll(mf);
    if (have) list[1 + (flat ? lane_rank(mf) : nf + lane_rank(mc))] = slot;
    if (lane == 0) {
        list[0] = (int32_t)nf;
        range[A] = make_uint2((uint32_t)A * kMaxBounceList, (uint32_t)A * kMaxBounceList + count + 1);
    }
}

// Task t of scene state k = t / tasks_per_step, i = t % tasks_per_step: first
// reflections off element i (pair_items == null), or second reflections of
// rays that left element pair_owner[t] (< 0: none) and hit the primitive
// pair_items[t]. The list goes to scratch[t * kMaxBounceList ...] as its
// number of flat candidates followed by the flat, then the curved candidates;
// counts[t] (if given) gets its length with that header, kNoBounceList for
// none or longer than kMaxBounceList, and range[t] (if given) its (begin, end)
// in the scratch, (1, 0) for none.
// sup_range / sup_items (first reflections): candidate lists that hold every
// primitive the element's first reflections can reach (k_reaim_lists); an
// element that has one tests its candidates' boxes instead of walking the BVH
// (the same leaf test, so the list holds every primitive the walk would).
constexpr int kConeBlock = 64;
__global__ __launch_bounds__(kConeBlock) void k_cone_lists(const st_element_f64* __restrict__ elems, int32_t tasks,
                                                           int32_t tasks_per_step,
                             int32_t n, int32_t n_nodes, const int32_t* __restrict__ elem_slot,
                             const uint32_t* __restrict__ meta, const double* __restrict__ frames_all,
                             const float* __restrict__ bounds_all, const float4* __restrict__ nodes_all,
                             const SunFrame* __restrict__ sun_all, double theta_sun,
                             const int32_t* __restrict__ pair_items, const int32_t* __restrict__ pair_owner,
                             uint32_t* __restrict__ counts, uint2* __restrict__ range, int32_t* __restrict__ scratch,
                             const uint2* __restrict__ sup_range, const int32_t* __restrict__ sup_items) {
    __shared__ int32_t lds_stack[(kStackDepth + 1) * kConeBlock];
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= tasks) return;
    const int k = g / tasks_per_step;
    // first reflections: thread g computes the list of the element whose
    // primitive sits at leaf slot g % n, so a wave's 64 mirrors are BVH
    // neighbours, whose cones walk nearly the same nodes; the list still goes
    // to task t = k * n + element
    const int t = pair_items ? g : k * tasks_per_step + (int)meta_elem(meta[g - k * tasks_per_step]);
    const double* frames = frames_all + (size_t)k * n * 9;
    const float* bounds_slot = bounds_all + (size_t)k * n * 6;
    const float4* nodes = nodes_all + (size_t)k * n_nodes * 4;
    const SunFrame* sun = sun_all + k;
    auto none = [&]() {
        if (counts) counts[t] = kNoBounceList;
        if (range) range[t] = make_uint2(1u, 0u);
    };
    double a[3], phi = 0.0;
    int from = t - k * tasks_per_step;  // element the rays leave from
    if (!pair_items) {
        if (!bounce_cone(elems[from], theta_sun, &phi)) {
            none();
            return;
        }
        sun_reflection_axis(frames, from, sun->sun, a);
    } else {
        const int ea = pair_owner[t];
        if (ea < 0) {  // a list header, not a candidate
            none();
            return;
        }
        from = (int)meta_elem(meta[pair_items[t]]);
        double phi_a = 0.0, phi_b = 0.0;
        if (!bounce_cone(elems[ea], theta_sun, &phi_a) || !bounce_cone(elems[from], 0.0, &phi_b)) {
            none();
            return;
        }
        // incidence within phi_a of a_A; reflection off the second mirror adds
        // its own error and normal-spread terms (phi_b with no sun term)
        double aa[3];
        sun_reflection_axis(frames, ea, sun->sun, aa);
        const double nb[3] = {frames[9 * (size_t)from + 2], frames[9 * (size_t)from + 5], frames[9 * (size_t)from + 8]};
        const double dn = aa[0] * nb[0] + aa[1] * nb[1] + aa[2] * nb[2];
        for (int j = 0; j < 3; ++j) a[j] = aa[j] - 2.0 * dn * nb[j];
        const double la = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
        for (int j = 0; j < 3; ++j) a[j] /= la;
        phi = phi_a + phi_b;
        if (phi >= 0.35) {
            none();
            return;
        }
    }
    Cone C;
    for (int j = 0; j < 3; ++j) C.a[j] = a[j];
    C.tp = tan(phi);
    C.cp = cos(phi);
    C.sp = sin(phi);
    // pieces reach past the scene's extent: the root's two child boxes bound it
    const float4 r0 = nodes[0], r1 = nodes[1], r2 = nodes[2];
    const double ex = fmax(r0.y, r1.y) - fmin(r0.x, r1.x), ey = fmax(r0.w, r1.w) - fmin(r0.z, r1.z);
    const double ez = fmax(r2.y, r2.w) - fmin(r2.x, r2.z);
    const double reach = 2.0 * sqrt(ex * ex + ey * ey + ez * ez) + 2.0 * kConeFirst;
    C.ratio = pow(reach / kConeFirst, 1.0 / (kConePieces - 2));
    int32_t* list = scratch + (size_t)t * kMaxBounceList;
    const float* hb = bounds_slot + 6 * (size_t)elem_slot[from];
    uint32_t len = 0;
    const uint2 sr = sup_range && !pair_items ? sup_range[from] : make_uint2(1u, 0u);
    if (sr.x <= sr.y) {  // the element's candidates over every re-aimed state (entry sr.x is its header)
        for (uint32_t i = sr.x + 1; i < sr.y && len < kMaxBounceList; ++i) {
            const int32_t v = sup_items[i];
            const float* b = bounds_slot + 6 * (size_t)v;
            if (cone_reaches(C, hb, b, b + 3)) {
                if (len < kMaxBounceList - 1) list[1 + len] = v;
                ++len;
            }
        }
    } else {
        len = cone_walk<kConeBlock>(nodes, bounds_slot, hb, C, list + 1, kMaxBounceList - 1, lds_stack + threadIdx.x);
    }
    if (len > kMaxBounceList - 1) {
        none();
        return;
    }
    if (elems[from].surface == ST_SURFACE_FLAT)
        len = cull_by_rectangle(elems, elems[from], frames, frames + 9 * (size_t)from, meta, C, phi, bounds_slot,
                                list + 1, len);
    uint32_t nf = 0;  // flat candidates first
    for (uint32_t i = 0; i < len; ++i) {
        const int32_t v = list[1 + i];
        if (meta_kind(meta[v]) == kKindFlat) {
            list[1 + i] = list[1 + nf];
            list[1 + nf] = v;
            ++nf;
        }
    }
    list[0] = (int32_t)nf;
    if (counts) counts[t] = len + 1;
    if (range) range[t] = make_uint2((uint32_t)t * kMaxBounceList, (uint32_t)t * kMaxBounceList + len + 1);
}

// items[range[t].x + k] = scratch[t * kMaxBounceList + k] for every listed task
__global__ void k_list_compact(int32_t tasks, const uint2* __restrict__ range, const int32_t* __restrict__ scratch,
                               int32_t* __restrict__ items) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (int64_t)tasks * kMaxBounceList) return;
    const int t = (int)(g / kMaxBounceList), k = (int)(g % kMaxBounceList);
    const uint2 r = range[t];
    if (r.x <= r.y && (uint32_t)k < r.y - r.x) items[r.x + k] = scratch[g];
}

__global__ void k_fill_u64(uint64_t* p, uint64_t n, uint64_t value) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = value;
}

// out[j][i] += sum of copies[j][c][i] over the copies c, which are cleared
__global__ void k_flux_gather(int64_t* __restrict__ copies, uint32_t n_copies, uint64_t bins, uint64_t total,
                              int64_t* __restrict__ out) {
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    int64_t* c = copies + (g / bins) * n_copies * bins + g % bins;
    int64_t sum = 0;
    for (uint32_t k = 0; k < n_copies; ++k) {
        sum += c[(size_t)k * bins];
        c[(size_t)k * bins] = 0;
    }
    out[g] += sum;
}

__global__ void k_flux_watts(const int64_t* __restrict__ flux, int32_t nbins, const SunFrame* __restrict__ sun,
                             double num_rays, double bin_area, double* __restrict__ out, double* ppr) {
    const double power_per_ray = sun->dni * sun->area / num_rays;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0 && ppr) *ppr = power_per_ray;
    if (i < nbins && out) out[i] = (double)flux[i] * power_per_ray / bin_area;
}

// ---------------------------------------------------------------------------
// VALU issue calibration: kCalChains independent FP32 FMA chains per lane
// (plain v_fma_f32; packed FP32 is disabled in this build), so the issue rate
// of plain vector instructions bounds the time.
constexpr int kCalChains = 8, kCalIters = 4096;
__global__ __launch_bounds__(256) void k_valu_calibrate(float* __restrict__ out, float a, float b) {
    float x[kCalChains];
#pragma unroll
    for (int c = 0; c < kCalChains; ++c) x[c] = (float)threadIdx.x * 1e-3f + (float)c;
    for (int i = 0; i < kCalIters; ++i) {
#pragma unroll
        for (int c = 0; c < kCalChains; ++c) x[c] = fmaf(x[c], a, b);
    }
    float sum = 0.f;
#pragma unroll
    for (int c = 0; c < kCalChains; ++c) sum += x[c];
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = sum;
}

hipError_t launch_valu_calibrate(hipStream_t s, float* out, int blocks, double* lane_ops) {
    hipLaunchKernelGGL(k_valu_calibrate, dim3(blocks), dim3(256), 0, s, out, 1.0001f, 0.5f);
    *lane_ops = (double)blocks * 256.0 * kCalIters * kCalChains;
    return hipGetLastError();
}

int device_cu_count() {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) return 256;
    return cus;
}

hipError_t launch_frames(hipStream_t s, const st_element_f64* elems, int32_t n, const int32_t* elem_slot,
                         const double* aim_override, const double* zrot_override, double* frames, double* ref_aabb,
                         float* bounds_slot, float4* prims, uint32_t* meta, float4* optics) {
    hipLaunchKernelGGL(k_frames, dim3((n + 127) / 128), dim3(128), 0, s, elems, n, elem_slot, aim_override,
                       zrot_override, frames, ref_aabb, bounds_slot, prims, meta, optics);
    return hipGetLastError();
}

hipError_t launch_aim(hipStream_t s, st_element_f64* elems, int32_t n, const double* sun_unit, const double* target) {
    hipLaunchKernelGGL(k_aim, dim3((n + 127) / 128), dim3(128), 0, s, elems, n, sun_unit, target);
    return hipGetLastError();
}

constexpr int kRoundGather = 32;
struct RoundResults {
    const uint64_t* stats[kRoundGather];
    const SunFrame* sun[kRoundGather];
};
// one block per job: its statistics and sun frame, word by word
__global__ void k_round_results(RoundResults r, uint64_t* __restrict__ stats_out, SunFrame* __restrict__ frames_out) {
    const int j = blockIdx.x, t = threadIdx.x;
    if (t < ST_STAT_COUNT) stats_out[(size_t)j * ST_STAT_COUNT + t] = r.stats[j][t];
    constexpr int kWords = (int)(sizeof(SunFrame) / sizeof(uint32_t));
    static_assert(sizeof(SunFrame) % sizeof(uint32_t) == 0, "sun frame words");
    const uint32_t* src = (const uint32_t*)r.sun[j];
    uint32_t* dst = (uint32_t*)(frames_out + j);
    for (int i = t; i < kWords; i += (int)blockDim.x) dst[i] = src[i];
}

hipError_t launch_round_results(hipStream_t s, const uint64_t* const* stats, const SunFrame* const* suns, int32_t count,
                                uint64_t* stats_out, SunFrame* frames_out) {
    for (int32_t b = 0; b < count; b += kRoundGather) {
        const int32_t c = std::min<int32_t>(kRoundGather, count - b);
        RoundResults r{};
        for (int32_t i = 0; i < c; ++i) {
            r.stats[i] = stats[b + i];
            r.sun[i] = suns[b + i];
        }
        hipLaunchKernelGGL(k_round_results, dim3((unsigned)c), dim3(64), 0, s, r, stats_out + (size_t)b * ST_STAT_COUNT,
                           frames_out + b);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_aim_copy(hipStream_t s, const st_element_f64* src, st_element_f64* dst, int32_t n,
                           const double* sun_unit, const double* target) {
    hipLaunchKernelGGL(k_aim_copy, dim3((n + 127) / 128), dim3(128), 0, s, src, dst, n, sun_unit, target);
    return hipGetLastError();
}

hipError_t launch_frames_batch(hipStream_t s, const st_element_f64* elems, int32_t n, const int32_t* elem_slot,
                               int32_t steps, const double* sun_units, const double* target, double* ref_aabb,
                               float* bounds_slot, float4* prims, double* frames) {
    const int64_t total = (int64_t)n * steps;
    hipLaunchKernelGGL(k_frames_batch, dim3((unsigned)((total + 127) / 128)), dim3(128), 0, s, elems, n, elem_slot,
                       steps, sun_units, target, ref_aabb, bounds_slot, prims, frames);
    return hipGetLastError();
}

hipError_t launch_refit(hipStream_t s, const int32_t* child, const int32_t* level_nodes,
                        const int32_t* level_offsets_host, int32_t levels, int32_t n_nodes, int32_t n_prims,
                        int32_t steps, const float* bounds_slot, float4* nodes) {
    if (levels <= kRefitMaxLevels && n_nodes <= kRefitOneLaunchNodes) {
        RefitLevels lv{};
        for (int l = 0; l <= levels; ++l) lv.offset[l] = level_offsets_host[l];
        hipLaunchKernelGGL(k_refit_levels, dim3((unsigned)steps), dim3(1024), 0, s, child, level_nodes, lv, levels,
                           n_nodes, n_prims, bounds_slot, nodes);
        return hipGetLastError();
    }
    for (int lv = levels - 1; lv >= 0; --lv) {
        const int32_t b = level_offsets_host[lv], c = level_offsets_host[lv + 1] - b;
        if (c <= 0) continue;
        const int64_t total = (int64_t)c * steps;
        hipLaunchKernelGGL(k_refit, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, child, level_nodes, b, c,
                           n_nodes, n_prims, steps, bounds_slot, nodes);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_sun_plane(hipStream_t s, const double* ref_aabb, int32_t n, int32_t steps, const double* sun_units,
                            const st_sun_desc* sun, SunFrame* out) {
    hipLaunchKernelGGL(k_sun_plane, dim3(steps), dim3(256), 0, s, ref_aabb, n, sun_units, *sun, out);
    return hipGetLastError();
}

template <int BLOCK, int STACK, int MINW, int QUEUED, int LHITS>
static void launch_variant(hipStream_t s, int grid, const JobDev* jobs, const uint64_t* chunk_prefix, int32_t n_jobs,
                           uint64_t total_chunks, uint32_t* work_counter, int chunk, int static_chunks,
                           const JobDev& single) {
    hipLaunchKernelGGL((k_trace<BLOCK, STACK, MINW, QUEUED, LHITS>), dim3(grid), dim3(BLOCK), 0, s, single, jobs,
                       chunk_prefix, n_jobs, total_chunks, work_counter, chunk, static_chunks);
}

template <int BLOCK, int STACK, int MINW, int QUEUED, int LHITS>
static int resident_blocks_per_cu() {
    static std::atomic<int> cached{-1};  // per variant; the library runs on one kind of device
    int nb = cached.load();
    if (nb < 0) {
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, (const void*)k_trace<BLOCK, STACK, MINW, QUEUED, LHITS>,
                                                         BLOCK, 0) != hipSuccess) {
            (void)hipGetLastError();
            nb = 0;
        }
        cached.store(nb);
    }
    return nb;
}

// Q: 0 whole rays, 1 queued, 2 queued with BVH segments grouped
#define ST_VARIANTS                    \
    ST_VARIANT(64, 32, 4, 0, 0)        \
    ST_VARIANT(64, 32, 5, 0, 0)        \
    ST_VARIANT(128, 32, 4, 0, 0)       \
    ST_VARIANT(128, 32, 5, 0, 0)       \
    ST_VARIANT(256, 32, 4, 0, 0)       \
    ST_VARIANT(128, 16, 4, 0, 0)       \
    ST_VARIANT(128, 16, 6, 0, 0)       \
    ST_VARIANT(128, 16, 8, 0, 0)       \
    ST_VARIANT(256, 16, 4, 0, 0)       \
    ST_VARIANT(256, 16, 6, 0, 0)       \
    ST_VARIANT(256, 16, 8, 0, 0)       \
    ST_VARIANT(128, 16, 4, 1, 0)       \
    ST_VARIANT(128, 16, 5, 1, 0)       \
    ST_VARIANT(256, 16, 4, 1, 0)       \
    ST_VARIANT(256, 16, 5, 1, 0)       \
    ST_VARIANT(64, 32, 3, 1, 0)        \
    ST_VARIANT(128, 32, 3, 1, 0)       \
    ST_VARIANT(128, 0, 6, 1, 0)        \
    ST_VARIANT(128, 0, 7, 1, 0)        \
    ST_VARIANT(128, 0, 8, 1, 0)        \
    ST_VARIANT(256, 0, 6, 1, 0)        \
    ST_VARIANT(256, 0, 7, 1, 0)        \
    ST_VARIANT(256, 0, 8, 1, 0)        \
    ST_VARIANT(128, 16, 5, 2, 0)       \
    ST_VARIANT(256, 16, 5, 2, 0)       \
    ST_VARIANT(128, 0, 6, 2, 0)        \
    ST_VARIANT(128, 0, 7, 2, 0)        \
    ST_VARIANT(128, 0, 8, 2, 0)        \
    ST_VARIANT(256, 0, 6, 2, 0)        \
    ST_VARIANT(256, 0, 7, 2, 0)        \
    ST_VARIANT(256, 0, 8, 2, 0)        \
    ST_VARIANT(256, 8, 5, 1, kLdsHitElements) \
    ST_VARIANT(256, 16, 5, 1, kLdsHitElements) \
    ST_VARIANT(256, 0, 6, 1, kLdsHitElements) \
    ST_VARIANT(256, 0, 6, 2, kLdsHitElements) \
    ST_VARIANT(256, 0, 7, 1, kLdsHitElements) \
    ST_VARIANT(256, 0, 7, 2, kLdsHitElements)

// Compiled shapes (the list above). The queued form keeps 2 KB of parked rays
// per wave in LDS on top of the stack, which caps it at 5 waves per SIMD (stack
// 16) or 3 (stack 32); with the stack in private memory (stack 0) LDS allows 8,
// and the register budgets of 6, 7 and 8 waves are compiled (72 registers at 7
// waves hold the queued form's state with two more spilled values than 80 at
// 6; 64 at 8 spill inside the loops).
bool trace_variant_compiled(int block, int stack, int waves, int form, bool lds_hits) {
    const int h = lds_hits ? kLdsHitElements : 0;
#define ST_VARIANT(B, S, W, Q, H) \
    if (block == B && stack == S && waves == W && form == Q && h == H) return true;
    ST_VARIANTS
#undef ST_VARIANT
    return false;
}
bool grouped_variant_exists(int block, int stack, int waves) {
    return trace_variant_compiled(block, stack, waves, 2, false);
}
bool trace_variant_exists(int block, int stack, int waves, bool queued) {
    return trace_variant_compiled(block, stack, waves, queued ? 1 : 0, false);
}

int trace_resident_waves(const LaunchShape& shape) {
    const int st = shape.stack, w = shape.waves, b = shape.block, h = shape.lds_hits ? kLdsHitElements : 0;
    const int r = shape.queued ? (shape.group_bvh ? 2 : 1) : 0;
#define ST_VARIANT(B, S, W, Q, H)                                    \
    if (b == B && st == S && w == W && r == Q && h == H)             \
        return resident_blocks_per_cu<B, S, W, Q, H>() * device_cu_count() * (B / 64);
    ST_VARIANTS
#undef ST_VARIANT
    return 0;
}

hipError_t launch_trace(hipStream_t s, const JobDev* jobs, const uint64_t* chunk_prefix, int32_t n_jobs,
                        uint64_t total_chunks, uint32_t* work_counter, const LaunchShape& shape,
                        const JobDev& single) {
    int grid = shape.grid > 0 ? shape.grid : device_cu_count() * 8;
    const uint64_t waves_needed = total_chunks;
    const int waves_per_block = shape.block / 64;
    const uint64_t max_blocks = (waves_needed + waves_per_block - 1) / waves_per_block;
    if ((uint64_t)grid > max_blocks) grid = (int)(max_blocks > 0 ? max_blocks : 1);
    const int st = shape.stack, w = shape.waves, b = shape.block, c = shape.chunk;
    const int h = shape.lds_hits ? kLdsHitElements : 0;
    const int r = shape.queued ? (shape.group_bvh ? 2 : 1) : 0;
    if (r && (c <= 0 || c > (int)kMaxChunkRays)) return hipErrorInvalidValue;  // parked rays keep an 11-bit offset
    // one chunk per wave without the shared counter, when every wave fits on
    // the device at once (a wave not yet resident would hold its chunk back
    // until others finish)
    const int stat = total_chunks <= (uint64_t)trace_resident_waves(shape) ? 1 : 0;
    if (stat) grid = (int)max_blocks;
#define ST_VARIANT(B, S, W, Q, H)                                                                                 \
    if (b == B && st == S && w == W && r == Q && h == H) {                                                       \
        launch_variant<B, S, W, Q, H>(s, grid, jobs, chunk_prefix, n_jobs, total_chunks, work_counter, c, stat,  \
                                      single);                                                                   \
        return hipGetLastError();                                                                                \
    }
    ST_VARIANTS
#undef ST_VARIANT
    return hipErrorInvalidValue;
}

hipError_t launch_grid_count(hipStream_t s, const float* bounds, const float4* prims, const uint32_t* meta, int32_t n,
                             int32_t steps, const SunFrame* sun, double tan_max, uint32_t* counts, uint32_t* start,
                             uint32_t* stats, JobDev* jobs, const int32_t* items, uint32_t item_cap) {
    const int64_t total = (int64_t)n * steps * kGridParts;
    hipLaunchKernelGGL(k_grid_count, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, bounds, prims, meta, n,
                       steps, sun, tan_max, counts);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_grid_scan, dim3(steps), dim3(1024), 0, s, counts, start, stats, jobs, items, item_cap);
    return hipGetLastError();
}

hipError_t launch_grid_fill(hipStream_t s, const float* bounds, const float4* prims, const uint32_t* meta, int32_t n,
                            int32_t steps, const SunFrame* sun, double tan_max, const uint32_t* start, uint32_t* cursor,
                            int32_t* items, uint32_t item_cap) {
    const int64_t total = (int64_t)n * steps * kGridParts;
    hipLaunchKernelGGL(k_grid_fill, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, bounds, prims, meta, n,
                       steps, sun, tan_max, start, cursor, items, item_cap);
    return hipGetLastError();
}

hipError_t launch_cone_lists(hipStream_t s, const st_element_f64* elems, int32_t tasks, int32_t tasks_per_step,
                             int32_t n, int32_t n_nodes, const int32_t* elem_slot, const uint32_t* meta,
                             const double* frames, const float* bounds_slot, const float4* nodes, const SunFrame* sun,
                             double theta_sun, const int32_t* pair_items, const int32_t* pair_owner, uint32_t* counts,
                             uint2* range, int32_t* scratch, const uint2* sup_range, const int32_t* sup_items) {
    if (tasks <= 0) return hipSuccess;
    if (tasks_per_step <= 0 || tasks % tasks_per_step != 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_cone_lists, dim3((tasks + kConeBlock - 1) / kConeBlock), dim3(kConeBlock), 0, s, elems, tasks,
                       tasks_per_step, n,
                       n_nodes, elem_slot, meta, frames, bounds_slot, nodes, sun, theta_sun, pair_items, pair_owner,
                       counts, range, scratch, sup_range, sup_items);
    return hipGetLastError();
}

hipError_t launch_reaim_lists(hipStream_t s, const st_element_f64* elems, int32_t n, const int32_t* elem_slot,
                              const uint32_t* meta, const double* target, double theta_sun, double4* spheres,
                              uint2* range, int32_t* scratch) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_element_spheres, dim3((n + 255) / 256), dim3(256), 0, s, elems, n, spheres);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_reaim_lists, dim3((unsigned)(((int64_t)n * 64 + kReaimBlock - 1) / kReaimBlock)),
                       dim3(kReaimBlock), 0, s, elems, n, spheres, elem_slot, meta, target, theta_sun, range, scratch);
    return hipGetLastError();
}

hipError_t launch_list_compact(hipStream_t s, int32_t tasks, const uint2* range, const int32_t* scratch,
                               int32_t* items) {
    if (tasks <= 0) return hipSuccess;
    const int64_t total = (int64_t)tasks * kMaxBounceList;
    hipLaunchKernelGGL(k_list_compact, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, tasks, range, scratch,
                       items);
    return hipGetLastError();
}

hipError_t launch_fill_u64(hipStream_t s, void* p, uint64_t n_u64, uint64_t value) {
    if (n_u64 == 0) return hipSuccess;
    uint64_t blocks = (n_u64 + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)blocks), dim3(256), 0, s, (uint64_t*)p, n_u64, value);
    return hipGetLastError();
}

hipError_t launch_flux_gather(hipStream_t s, int64_t* copies, uint32_t n_copies, uint64_t bins, int32_t jobs,
                              int64_t* out) {
    const uint64_t total = bins * (uint64_t)jobs;
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(k_flux_gather, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, copies, n_copies, bins,
                       total, out);
    return hipGetLastError();
}

hipError_t launch_flux_watts(hipStream_t s, const int64_t* flux, int32_t nbins, const SunFrame* sun, double num_rays,
                             double bin_area, double* out, double* power_per_ray_dev) {
    const int n = nbins > 0 ? nbins : 1;
    hipLaunchKernelGGL(k_flux_watts, dim3((n + 255) / 256), dim3(256), 0, s, flux, nbins, sun, num_rays, bin_area, out,
                       power_per_ray_dev);
    return hipGetLastError();
}

}  // namespace st
