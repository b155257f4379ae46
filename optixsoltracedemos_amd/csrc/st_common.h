// st_common.h - definitions shared by host C++ and gfx950 device code.
//
// FP64 element frames and bounds (used on the GPU for aiming / refit and on
// the host for the BVH topology), the FP32 primitive record the trace kernels
// read, the BVH node layout and Philox4x32-10.
#pragma once

#include <math.h>
#include <stdint.h>

#include "../../include/soltrace_hip.h"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define ST_HD __host__ __device__
#define ST_INLINE __host__ __device__ __forceinline__
#else
#define ST_HD
#define ST_INLINE inline
struct float4 { float x, y, z, w; };
#endif

namespace st {

constexpr double kPi = 3.14159265358979323846;
constexpr int kStackDepth = 32;       // traversal stack entries per lane (tree depth <= 31)
constexpr int kMaxLeaf = 4;           // primitives per BVH leaf
constexpr int kMaxDepthLimit = 16;    // largest max_depth accepted by st_trace
constexpr float kTminPrimary = 1e-3f; // cuda/sun.cu:123
constexpr float kTminBounce = 1e-2f;  // cuda/materials.cu:70
constexpr float kTmax = 1e16f;        // cuda/sun.cu:124

// prim_meta word: bits 0-1 surface code (0 flat, 1 parabolic, 2 cylinder),
// bit 2 receiver, bits 8-31 element index
enum : uint32_t { kKindFlat = 0, kKindParabolic = 1, kKindCylinder = 2 };
ST_INLINE uint32_t meta_pack(uint32_t elem, uint32_t kind, uint32_t receiver) {
    return (elem << 8) | (receiver << 2) | kind;
}
ST_INLINE uint32_t meta_elem(uint32_t m) { return m >> 8; }
ST_INLINE uint32_t meta_kind(uint32_t m) { return m & 3u; }
ST_INLINE uint32_t meta_receiver(uint32_t m) { return (m >> 2) & 1u; }
constexpr uint32_t kMaxElements = 1u << 24;

// BVH child pointer: >= 0 inner node index; < 0 leaf = ~((start << 3) | (count - 1))
ST_INLINE int32_t leaf_code(int32_t start, int32_t count) { return ~((start << 3) | (count - 1)); }
ST_INLINE int32_t leaf_start(int32_t c) { return (~c) >> 3; }
ST_INLINE int32_t leaf_count(int32_t c) { return ((~c) & 7) + 1; }

// Node layout (64 B, four float4):
//   n0 = (c0.lo.x, c0.hi.x, c0.lo.y, c0.hi.y)
//   n1 = (c1.lo.x, c1.hi.x, c1.lo.y, c1.hi.y)
//   n2 = (c0.lo.z, c0.hi.z, c1.lo.z, c1.hi.z)
//   n3 = (child0 bits, child1 bits, 0, 0)
//
// Primitive record (64 B, four float4), element frame in FP32:
//   q0 = (center.xyz, hw)   flat: w/2 | parabolic: L1/2 | cylinder: radius
//   q1 = (ax.xyz, hh)       flat: x   | parabolic: e1 = -x_b | cylinder: base_x ; h/2
//   q2 = (ay.xyz, cx)       flat: y   | parabolic: e2 = y_b  | cylinder: axis = base_z x base_x
//   q3 = (az.xyz, cy)       flat: x*y | parabolic: normalize(e2 x e1) | cylinder: base_z

// ---------------------------------------------------------------------------
// FP64 frame: lib/math_util.h:21-59 + lib/element.cpp:77-100. Row-major L2G;
// its columns are the element x, y, z axes.
ST_INLINE void element_frame(const st_element_f64& e, const double aim[3], double zrot_deg, double l2g[9]) {
    double n0 = aim[0] - e.origin[0], n1 = aim[1] - e.origin[1], n2 = aim[2] - e.origin[2];
    double len = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
    if (len > 0) { n0 /= len; n1 /= len; n2 /= len; }
    double yaw = atan2(n0, n2), pitch = asin(n1), roll = zrot_deg * kPi / 180.0;
    double ca = cos(yaw), sa = sin(yaw), cb = cos(pitch), sb = sin(pitch), cg = cos(roll), sg = sin(roll);
    // G2L (Spencer & Murty eq. 2); L2G = transpose
    double g[9] = {ca * cg + sa * sb * sg, -cb * sg, -sa * cg + ca * sb * sg,
                   ca * sg - sa * sb * cg, cb * cg,  -sa * sg - ca * sb * cg,
                   sa * cb,                sb,       ca * cb};
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) l2g[r * 3 + c] = g[c * 3 + r];
}

ST_INLINE void l2g_apply(const double l2g[9], const double o[3], double x, double y, double z, double out[3]) {
    out[0] = l2g[0] * x + l2g[1] * y + l2g[2] * z + o[0];
    out[1] = l2g[3] * x + l2g[4] * y + l2g[5] * z + o[1];
    out[2] = l2g[6] * x + l2g[7] * y + l2g[8] * z + o[2];
}

// Reference AABB (lib/element.cpp:172-275): rectangle = 4 flat corners,
// cylinder = 8 corners of (+-w/2, +-h/2, +-w/2). Feeds the sun plane.
ST_INLINE void element_ref_aabb(const st_element_f64& e, const double l2g[9], double lo[3], double hi[3]) {
    double w = e.dim_x, h = e.dim_y;
    int nc = e.surface == ST_SURFACE_CYLINDER ? 8 : 4;
    for (int k = 0; k < 3; ++k) { lo[k] = 1e300; hi[k] = -1e300; }
    for (int c = 0; c < nc; ++c) {
        double p[3];
        double z = nc == 8 ? ((c & 4) ? w / 2 : -w / 2) : 0.0;
        l2g_apply(l2g, e.origin, (c & 1) ? w / 2 : -w / 2, (c & 2) ? h / 2 : -h / 2, z, p);
        for (int k = 0; k < 3; ++k) { lo[k] = fmin(lo[k], p[k]); hi[k] = fmax(hi[k], p[k]); }
    }
}

// Conservative traversal bounds: the reference box extended by the paraboloid
// sag (the reference leaves it out, lib/element.cpp:181-206) and padded for
// FP32 slab tests (1 mm + 2e-6 |coord|).
ST_INLINE void element_bounds(const st_element_f64& e, const double l2g[9], float lo[3], float hi[3]) {
    double w = e.dim_x, h = e.dim_y, zlo = 0, zhi = 0;
    if (e.surface == ST_SURFACE_PARABOLIC) {
        double sx = 0.5 * e.curv_x * (w / 2) * (w / 2), sy = 0.5 * e.curv_y * (h / 2) * (h / 2);
        zlo = fmin(0.0, sx) + fmin(0.0, sy);
        zhi = fmax(0.0, sx) + fmax(0.0, sy);
    } else if (e.surface == ST_SURFACE_CYLINDER) {
        zlo = -w / 2;
        zhi = w / 2;
    }
    double dlo[3] = {1e300, 1e300, 1e300}, dhi[3] = {-1e300, -1e300, -1e300};
    for (int c = 0; c < 8; ++c) {
        double p[3];
        l2g_apply(l2g, e.origin, (c & 1) ? w / 2 : -w / 2, (c & 2) ? h / 2 : -h / 2, (c & 4) ? zhi : zlo, p);
        for (int k = 0; k < 3; ++k) { dlo[k] = fmin(dlo[k], p[k]); dhi[k] = fmax(dhi[k], p[k]); }
    }
    for (int k = 0; k < 3; ++k) {
        double pad = 1e-3 + 2e-6 * fmax(fabs(dlo[k]), fabs(dhi[k]));
        lo[k] = (float)(dlo[k] - pad);
        hi[k] = (float)(dhi[k] + pad);
    }
}

// FP64 -> FP32 primitive record: lib/element.cpp:114-164 (toDeviceGeometryData)
// with the per-call frame work of cuda/intersection.cu:339-355 hoisted here.
ST_INLINE void element_pack(const st_element_f64& e, const double l2g[9], float4 q[4]) {
    double x[3] = {l2g[0], l2g[3], l2g[6]}, y[3] = {l2g[1], l2g[4], l2g[7]}, z[3] = {l2g[2], l2g[5], l2g[8]};
    double ax[3], ay[3], az[3];
    float hw = (float)(e.dim_x / 2), hh = (float)(e.dim_y / 2), cx = 0.f, cy = 0.f;
    if (e.surface == ST_SURFACE_PARABOLIC) {
        for (int k = 0; k < 3; ++k) { ax[k] = -x[k]; ay[k] = y[k]; }
        // n = normalize(e2 x e1)
        az[0] = ay[1] * ax[2] - ay[2] * ax[1];
        az[1] = ay[2] * ax[0] - ay[0] * ax[2];
        az[2] = ay[0] * ax[1] - ay[1] * ax[0];
        double l = sqrt(az[0] * az[0] + az[1] * az[1] + az[2] * az[2]);
        for (int k = 0; k < 3; ++k) az[k] /= l;
        cx = (float)e.curv_x;
        cy = (float)e.curv_y;
    } else if (e.surface == ST_SURFACE_CYLINDER) {
        for (int k = 0; k < 3; ++k) { ax[k] = x[k]; az[k] = z[k]; }
        // axis = base_z x base_x (cuda/intersection.cu:192-195)
        ay[0] = az[1] * ax[2] - az[2] * ax[1];
        ay[1] = az[2] * ax[0] - az[0] * ax[2];
        ay[2] = az[0] * ax[1] - az[1] * ax[0];
    } else {
        for (int k = 0; k < 3; ++k) { ax[k] = x[k]; ay[k] = y[k]; }
        az[0] = ax[1] * ay[2] - ax[2] * ay[1];
        az[1] = ax[2] * ay[0] - ax[0] * ay[2];
        az[2] = ax[0] * ay[1] - ax[1] * ay[0];
        double l = sqrt(az[0] * az[0] + az[1] * az[1] + az[2] * az[2]);
        for (int k = 0; k < 3; ++k) az[k] /= l;
    }
    q[0] = {(float)e.origin[0], (float)e.origin[1], (float)e.origin[2], hw};
    q[1] = {(float)ax[0], (float)ax[1], (float)ax[2], hh};
    q[2] = {(float)ay[0], (float)ay[1], (float)ay[2], cx};
    q[3] = {(float)az[0], (float)az[1], (float)az[2], cy};
}

ST_INLINE uint32_t element_kind(const st_element_f64& e) {
    return e.surface == ST_SURFACE_PARABOLIC ? kKindParabolic
                                             : (e.surface == ST_SURFACE_CYLINDER ? kKindCylinder : kKindFlat);
}

// Bisector aim (FP64): the mirror normal that reflects the sun centre onto
// target; the aim point is 100 m along it, like the toy STINPUT aims
// (data/stinput/toy_problem_parabolic.stinput:15-17).
ST_INLINE void bisector_aim_point(const double origin[3], const double sun_unit[3], const double target[3],
                                  double aim[3]) {
    double t[3] = {target[0] - origin[0], target[1] - origin[1], target[2] - origin[2]};
    double lt = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
    double n[3];
    for (int k = 0; k < 3; ++k) n[k] = sun_unit[k] + (lt > 0 ? t[k] / lt : 0.0);
    double ln = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (int k = 0; k < 3; ++k) aim[k] = origin[k] + 100.0 * (ln > 0 ? n[k] / ln : sun_unit[k]);
}

// ---------------------------------------------------------------------------
// Sun frame consumed by the trace kernels (written on the device in FP64,
// FP32 copies for the per-ray work).
struct SunFrame {
    double v0[3], v1[3], v2[3], v3[3];
    double sun[3];                  // unit vector towards the sun
    double distance, area, box_half_angle, dni;
    float o[3], e1[3], e2[3];       // v0, v1 - v0, v3 - v0
    float w[3], u[3], v[3];         // sampling basis, cuda/sun.cu:77-80
    float vers;                     // pillbox: 1 - cos(alpha) = 2 sin^2(alpha/2)
    float sigma;                    // gaussian sigma
    int32_t shape, position_sampling;
};

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11): counter (ray_lo, ray_hi, block, job),
// key (seed_lo, seed_hi).
struct U4 { uint32_t x, y, z, w; };

// Each round multiplies through one 32x32 -> 64-bit product (a single
// v_mad_u64_u32 on gfx950, where the separate high and low multiplies would
// each cost as much).
ST_INLINE U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x, p1 = (uint64_t)0xCD9E8D57u * c.z;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

// 23-bit uniform ((x >> 9) + 0.5) / 2^23 in (0,1); u and 1 - u are exact in FP32
ST_INLINE float u01(uint32_t x) { return ((float)(x >> 9) + 0.5f) * (1.0f / 8388608.0f); }

}  // namespace st
