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
    int32_t grid_rows;              // rows of the sun-plane grid along e2 (frame_rows)
    int32_t reserved;
};

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11): counter (ray_lo, ray_hi, block, job),
// key (seed_lo, seed_hi).
struct U4 { uint32_t x, y, z, w; };

// a ^ b ^ c. On the device one v_bitop3_b32 (truth table 0x96), where the
// compiler issues two XORs; a vector instruction, so only for values that
// differ across lanes (uniform ones stay on the scalar unit as plain XORs).
ST_INLINE uint32_t xor3_lanes(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
#else
    return a ^ b ^ c;
#endif
}

// Each round multiplies through one 32x32 -> 64-bit product (a single
// v_mad_u64_u32 on gfx950, where the separate high and low multiplies would
// each cost as much). LANES: a per-lane counter under a wave-uniform key (the
// three-way XORs of a round as two bitop3 instructions instead of four XORs;
// the key is made opaque at each call, so its ten-round schedule is recomputed
// on the scalar unit instead of being hoisted out of the caller's loops into
// registers that do not fit and are reloaded by vector instructions); the
// results are the same either way.
template <bool LANES = false>
ST_INLINE U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
#if defined(__HIP_DEVICE_COMPILE__)
    if (LANES) asm volatile("" : "+s"(k0), "+s"(k1));
#endif
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x, p1 = (uint64_t)0xCD9E8D57u * c.z;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        if (LANES) c = U4{xor3_lanes(hi1, c.y, k0), lo1, xor3_lanes(hi0, c.w, k1), lo0};
        else c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

// cell of the stratified sun-plane sampling: 64 consecutive rays per cell of a
// 256 x 256 grid, cells in a keyed 16-bit Feistel order within each sweep
ST_INLINE uint32_t stratified_cell_of_group(uint64_t group, uint32_t seed_lo, uint32_t seed_hi, uint32_t job) {
    const uint64_t sweep = group >> 16;
    const U4 k = philox4x32_10(U4{(uint32_t)sweep, (uint32_t)(sweep >> 32), 0x53545254u, job}, seed_lo, seed_hi);
    const uint32_t keys[4] = {k.x, k.y, k.z, k.w};
    uint32_t L = ((uint32_t)group >> 8) & 0xFFu, R = (uint32_t)group & 0xFFu;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const uint32_t F = ((R ^ keys[r]) * 0x9E3779B1u) >> 24;
        const uint32_t nl = R;
        R = L ^ F;
        L = nl;
    }
    return (L << 8) | R;
}
// The inverse: the group of sweep `sweep` (bits 16 and up of the result) whose
// cell is `cell` (the Feistel rounds undone in reverse order).
ST_INLINE uint64_t stratified_group_of_cell(uint64_t sweep, uint32_t cell, uint32_t seed_lo, uint32_t seed_hi,
                                            uint32_t job) {
    const U4 k = philox4x32_10(U4{(uint32_t)sweep, (uint32_t)(sweep >> 32), 0x53545254u, job}, seed_lo, seed_hi);
    const uint32_t keys[4] = {k.x, k.y, k.z, k.w};
    uint32_t L = (cell >> 8) & 0xFFu, R = cell & 0xFFu;
#pragma unroll
    for (int r = 3; r >= 0; --r) {
        const uint32_t pr = L;
        L = R ^ (((L ^ keys[r]) * 0x9E3779B1u) >> 24);
        R = pr;
    }
    return (sweep << 16) | (L << 8) | R;
}
// Cell (y << 8 | x) at position h of a Hilbert curve over the 256 x 256 grid
// (the quadrant-by-quadrant construction: each step reads two bits of h and
// rotates or reflects the sub-square it enters).
ST_INLINE uint32_t hilbert_cell(uint32_t h) {
    uint32_t x = 0, y = 0;
    for (uint32_t s = 1; s < 256u; s <<= 1) {
        const uint32_t rx = 1u & (h >> 1), ry = 1u & (h ^ rx);
        if (ry == 0) {
            if (rx == 1) {
                x = s - 1 - x;
                y = s - 1 - y;
            }
            const uint32_t tmp = x;
            x = y;
            y = tmp;
        }
        x += s * rx;
        y += s * ry;
        h >>= 2;
    }
    return (y << 8) | x;
}

// 23-bit uniform ((x >> 9) + 0.5) / 2^23 in (0,1); u and 1 - u are exact in FP32
ST_INLINE float u01(uint32_t x) { return ((float)(x >> 9) + 0.5f) * (1.0f / 8388608.0f); }

// FP32 reciprocal: the hardware v_rcp_f32 on the device (about 1 ulp), a
// division on the host
ST_INLINE float rcp_f32(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}

// atan2(y, x) in (-pi, pi] for the cylindrical flux bin. Octant reduction to
// a = min(|x|,|y|) / max(|x|,|y|) in [0, 1], then the polynomial of
// Abramowitz & Stegun, Handbook of Mathematical Functions, eq. 4.4.49 (from
// C. Hastings, Approximations for Digital Computers, 1955):
//   atan(a) = a (1 + a2 a^2 + a4 a^4 + ... + a16 a^16),  |error| <= 2e-8 on [0, 1],
// so the FP32 result errs by its own rounding, a few 1e-8 rad (checked on
// the host against the libm atan2 by tests/cpp/atan_check.cpp). atan2(0, 0) = 0.
ST_INLINE float atan2_poly(float y, float x) {
    const float ax = fabsf(x), ay = fabsf(y);
    const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
    const float a = mx > 0.f ? mn * rcp_f32(mx) : 0.f;
    const float s = a * a;
    float r = fmaf(s, 0.0028662257f, -0.0161657367f);
    r = fmaf(s, r, 0.0429096138f);
    r = fmaf(s, r, -0.0752896400f);
    r = fmaf(s, r, 0.1065626393f);
    r = fmaf(s, r, -0.1420889944f);
    r = fmaf(s, r, 0.1999355085f);
    r = fmaf(s, r, -0.3333314528f);
    r = fmaf(s, r, 1.0f);
    r *= a;
    if (ay > ax) r = 1.5707963267948966f - r;
    if (x < 0.f) r = 3.141592653589793f - r;
    return copysignf(r, y);
}

// ---------------------------------------------------------------------------
// The conservative tests that build candidate lists: the sun-plane grid's
// footprints, the reflection cone against boxes and the flat-mirror sweep. They
// run on the device (k_grid_count / k_grid_fill, k_cone_lists) and are checked
// against brute-force ray sampling on the host.

constexpr int kGridRes = 256;  // sun-plane candidate grid, the stratified-sampling cells
constexpr int kGridCells = kGridRes * kGridRes;
// Rows of the grid along e2 under Halton positions: 3^5, so that the rays whose
// indices agree modulo 2^8 * 3^5 (one Halton cell: a in [i/256, (i+1)/256),
// b in [j/243, (j+1)/243)) start in one grid cell. Rows 243..255 stay empty.
constexpr int kHaltonRows = 243;
ST_INLINE int frame_rows(const SunFrame& F) { return F.grid_rows > 0 ? F.grid_rows : kGridRes; }

// Sun-plane candidate grid: kGridRes x frame_rows cells over the sun plane's
// (a, b) parameters. A primitive is listed in every cell from which a sun ray
// can reach its traversal box: the box corners projected along the sun onto
// the plane, widened by the largest lateral drift of a ray in the sun cone over
// the axial distance from the plane (tan_max * (distance - corner . sun)) plus
// 1 mm and 1e-6 of the plane's extent for FP32 ray origins.
// A flat primitive is listed more tightly: a sun ray hits it only at its
// rectangle, whose projection along the sun is a parallelogram (centre c,
// half-edge vectors ex, ey, in metres along the plane's unit axes); widened by
// the same margin plus the rectangle's FP32 allowance (1 mm and 2e-6 of its
// coordinates, as the traversal boxes), it bounds the cell range, and a cell
// is listed only if no edge normal of the parallelogram separates the two
// (cell_meets).
struct Footprint {
    int r[4];
    int normals;             // 0: the cell range only; 2: parallelogram test
    double c[2], n[2][2], h[2];  // centre, unit edge normals, reach along each normal
    double cell[2];          // cell size in metres along the plane's axes
};
ST_INLINE bool cell_meets(const Footprint& f, int ix, int iy) {
    const double x = (ix + 0.5) * f.cell[0] - f.c[0], y = (iy + 0.5) * f.cell[1] - f.c[1];
    for (int j = 0; j < f.normals; ++j) {
        const double ext = fabs(f.n[j][0]) * 0.5 * f.cell[0] + fabs(f.n[j][1]) * 0.5 * f.cell[1];
        if (fabs(f.n[j][0] * x + f.n[j][1] * y) > f.h[j] + ext) return false;
    }
    return true;
}
ST_INLINE bool grid_range(const float* __restrict__ bounds, const float4* __restrict__ prims,
                          const uint32_t* __restrict__ meta, int slot, const SunFrame& F, double tan_max,
                          Footprint& fp) {
    int* r = fp.r;
    fp.normals = 0;
    const double e1[3] = {F.v1[0] - F.v0[0], F.v1[1] - F.v0[1], F.v1[2] - F.v0[2]};
    const double e2[3] = {F.v3[0] - F.v0[0], F.v3[1] - F.v0[1], F.v3[2] - F.v0[2]};
    const double l1 = e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2];
    const double l2 = e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2];
    if (!(l1 > 0.0) || !(l2 > 0.0)) return false;
    double amin = 1e300, amax = -1e300, bmin = 1e300, bmax = -1e300, smin = 1e300;
    for (int c = 0; c < 8; ++c) {
        const double p[3] = {(double)bounds[6 * slot + ((c & 1) ? 3 : 0)], (double)bounds[6 * slot + 1 + ((c & 2) ? 3 : 0)],
                             (double)bounds[6 * slot + 2 + ((c & 4) ? 3 : 0)]};
        const double q[3] = {p[0] - F.v0[0], p[1] - F.v0[1], p[2] - F.v0[2]};
        const double a = (q[0] * e1[0] + q[1] * e1[1] + q[2] * e1[2]) / l1;
        const double b = (q[0] * e2[0] + q[1] * e2[1] + q[2] * e2[2]) / l2;
        amin = fmin(amin, a);
        amax = fmax(amax, a);
        bmin = fmin(bmin, b);
        bmax = fmax(bmax, b);
        smin = fmin(smin, p[0] * F.sun[0] + p[1] * F.sun[1] + p[2] * F.sun[2]);
    }
    const double extent = sqrt(l1) + sqrt(l2) + fabs(F.v0[0]) + fabs(F.v0[1]) + fabs(F.v0[2]);
    const double m = fmax(F.distance - smin, 0.0) * tan_max + 1e-3 + 1e-6 * extent;
    const double L1 = sqrt(l1), L2 = sqrt(l2), ma = m / L1, mb = m / L2;
    amin -= ma;
    amax += ma;
    bmin -= mb;
    bmax += mb;
    if (meta_kind(meta[slot]) == kKindFlat) {
        const float4 q0 = prims[4 * slot], q1 = prims[4 * slot + 1], q2 = prims[4 * slot + 2];
        const double hw = q0.w, hh = q1.w;
        const double cen[3] = {q0.x, q0.y, q0.z}, ax[3] = {q1.x, q1.y, q1.z}, ay[3] = {q2.x, q2.y, q2.z};
        const double u1[3] = {e1[0] / L1, e1[1] / L1, e1[2] / L1}, u2[3] = {e2[0] / L2, e2[1] / L2, e2[2] / L2};
        const double q[3] = {cen[0] - F.v0[0], cen[1] - F.v0[1], cen[2] - F.v0[2]};
        auto d3 = [](const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
        const double M = m + 1e-3 + 2e-6 * (fabs(cen[0]) + fabs(cen[1]) + fabs(cen[2]) + hw + hh);
        const double ex[2] = {hw * d3(ax, u1), hw * d3(ax, u2)}, ey[2] = {hh * d3(ay, u1), hh * d3(ay, u2)};
        fp.c[0] = d3(q, u1);
        fp.c[1] = d3(q, u2);
        amin = fmax(amin, (fp.c[0] - fabs(ex[0]) - fabs(ey[0]) - M) / L1);
        amax = fmin(amax, (fp.c[0] + fabs(ex[0]) + fabs(ey[0]) + M) / L1);
        bmin = fmax(bmin, (fp.c[1] - fabs(ex[1]) - fabs(ey[1]) - M) / L2);
        bmax = fmin(bmax, (fp.c[1] + fabs(ex[1]) + fabs(ey[1]) + M) / L2);
        const double lx = sqrt(ex[0] * ex[0] + ex[1] * ex[1]), ly = sqrt(ey[0] * ey[0] + ey[1] * ey[1]);
        // an edge seen end-on gives no normal; a normal that exists bounds the
        // parallelogram by the other edge's reach
        if (lx > 1e-9 && ly > 1e-9) {
            fp.n[0][0] = -ex[1] / lx;
            fp.n[0][1] = ex[0] / lx;
            fp.h[0] = fabs(fp.n[0][0] * ey[0] + fp.n[0][1] * ey[1]) + M;
            fp.n[1][0] = -ey[1] / ly;
            fp.n[1][1] = ey[0] / ly;
            fp.h[1] = fabs(fp.n[1][0] * ex[0] + fp.n[1][1] * ex[1]) + M;
            fp.cell[0] = L1 / kGridRes;
            fp.cell[1] = L2 / frame_rows(F);
            fp.normals = 2;
        }
    }
    if (amax < 0.0 || amin > 1.0 || bmax < 0.0 || bmin > 1.0) return false;
    r[0] = (int)fmax(0.0, floor(amin * kGridRes));
    r[1] = (int)fmin(kGridRes - 1.0, floor(amax * kGridRes));
    const int rows = frame_rows(F);
    r[2] = (int)fmax(0.0, floor(bmin * rows));
    r[3] = (int)fmin(rows - 1.0, floor(bmax * rows));
    return r[0] <= r[1] && r[2] <= r[3];
}

// A flat mirror's rays leave from its rectangle R, not from its whole box, and
// the set they sweep, S = R + {t u : t >= 0, angle(u, a) <= phi}, is convex. A
// candidate box that a plane separates from S cannot be reached: the walk's
// candidates are filtered with planes of normal m and m . u <= 0 for every
// cone direction u, so that the largest m . x over S is the largest over R.
// For any unit w normal to a, m = cos(phi) w - sin(phi) a qualifies; w runs
// over the separating axes of the two outlines seen along a (a x v for R's
// axes and the three world axes, both senses) and the direction from R's
// centre to the box centre normal to a. Also qualifying: -a (boxes wholly
// behind R along a) and the mirror plane, oriented towards a, when every cone
// direction leaves on its side (boxes wholly behind the mirror). R is widened
// by eps (FP32 hit points and the FP32 rectangle lie within it) and phi by
// 1e-6 rad (FP32 directions). R: centre pc, half sizes hw, hh along the
// frame's x and y axes (f: 3x3 row-major, columns x, y, z = normal); a: unit
// cone axis.
struct RectSweep {
    double pc[3], X[3], Y[3];
    double a[3];
    double hw, hh, eps, cph, sph;
    double M[12][3];  // the planes that do not depend on the box
    int nm;
};
ST_INLINE double rs_dot(const double* u, const double* v) { return u[0] * v[0] + u[1] * v[1] + u[2] * v[2]; }
// largest m . x over R widened by eps
ST_INLINE double rs_max_r(const RectSweep& S, const double m[3]) {
    return rs_dot(m, S.pc) + fabs(rs_dot(m, S.X)) * S.hw + fabs(rs_dot(m, S.Y)) * S.hh + S.eps;
}
// smallest m . x over the box b = (lo, hi)
ST_INLINE double rs_min_b(const double m[3], const float* b) {
    double v = 0.0;
    for (int k = 0; k < 3; ++k) v += fmin(m[k] * (double)b[k], m[k] * (double)b[3 + k]);
    return v;
}
// m from a direction v: w = unit part of sign * v normal to a (none if tiny)
ST_INLINE bool rs_plane_of(const RectSweep& S, const double v[3], double sign, double m[3]) {
    const double va = rs_dot(v, S.a);
    const double w[3] = {sign * (v[0] - va * S.a[0]), sign * (v[1] - va * S.a[1]), sign * (v[2] - va * S.a[2])};
    const double l = sqrt(rs_dot(w, w));
    if (l < 1e-9) return false;
    for (int k = 0; k < 3; ++k) m[k] = S.cph * w[k] / l - S.sph * S.a[k];
    return true;
}
ST_INLINE void rect_sweep_init(RectSweep& S, const double pc[3], double hw, double hh, const double* f,
                               const double a[3], double phi) {
    for (int k = 0; k < 3; ++k) {
        S.pc[k] = pc[k];
        S.X[k] = f[3 * k];
        S.Y[k] = f[3 * k + 1];
        S.a[k] = a[k];
    }
    const double N[3] = {f[2], f[5], f[8]};
    S.hw = hw;
    S.hh = hh;
    S.eps = 1e-3 + 2e-6 * (fabs(pc[0]) + fabs(pc[1]) + fabs(pc[2]) + hw + hh);
    const double ph = phi + 1e-6;
    S.cph = cos(ph);
    S.sph = sin(ph);
    const double* A = S.a;
    const double axw[3][3] = {{0.0, -A[2], A[1]}, {A[2], 0.0, -A[0]}, {-A[1], A[0], 0.0}};  // a x world axes
    const double axX[3] = {A[1] * S.X[2] - A[2] * S.X[1], A[2] * S.X[0] - A[0] * S.X[2], A[0] * S.X[1] - A[1] * S.X[0]};
    const double axY[3] = {A[1] * S.Y[2] - A[2] * S.Y[1], A[2] * S.Y[0] - A[0] * S.Y[2], A[0] * S.Y[1] - A[1] * S.Y[0]};
    S.nm = 0;
    for (int sgn = -1; sgn <= 1; sgn += 2) {
        if (rs_plane_of(S, axX, (double)sgn, S.M[S.nm])) ++S.nm;
        if (rs_plane_of(S, axY, (double)sgn, S.M[S.nm])) ++S.nm;
        for (int k = 0; k < 3; ++k)
            if (rs_plane_of(S, axw[k], (double)sgn, S.M[S.nm])) ++S.nm;
    }
    for (int k = 0; k < 3; ++k) S.M[S.nm][k] = -A[k];
    ++S.nm;
    const double na = rs_dot(N, A);
    if (fabs(na) >= S.sph) {
        const double s = na > 0.0 ? -1.0 : 1.0;
        for (int k = 0; k < 3; ++k) S.M[S.nm][k] = s * N[k];
        ++S.nm;
    }
}
// true when a plane proves that no ray of the sweep reaches the box b = (lo, hi)
ST_INLINE bool rect_sweep_misses(const RectSweep& S, const float* b) {
    for (int j = 0; j < S.nm; ++j)
        if (rs_max_r(S, S.M[j]) < rs_min_b(S.M[j], b)) return true;
    const double d[3] = {0.5 * ((double)b[0] + b[3]) - S.pc[0], 0.5 * ((double)b[1] + b[4]) - S.pc[1],
                         0.5 * ((double)b[2] + b[5]) - S.pc[2]};
    double m[3];
    return rs_plane_of(S, d, 1.0, m) && rs_max_r(S, m) < rs_min_b(m, b);
}

// A flat candidate of a reflection list: its rectangle (centre c, unit axes X,
// Y and normal N of its frame, half sizes hw, hh) thickened by eps along each
// of them, eps the FP32 allowance of the traversal boxes (1 mm and 2e-6 of its
// coordinates): a ray's FP32 hit on the candidate lies within it.
struct RectCand {
    double c[3], X[3], Y[3], N[3];
    double hw, hh, eps;
};
ST_INLINE void rect_cand_init(RectCand& R, const double c[3], double hw, double hh, const double* f) {
    for (int k = 0; k < 3; ++k) {
        R.c[k] = c[k];
        R.X[k] = f[3 * k];
        R.Y[k] = f[3 * k + 1];
        R.N[k] = f[3 * k + 2];
    }
    R.hw = hw;
    R.hh = hh;
    R.eps = 1e-3 + 2e-6 * (fabs(c[0]) + fabs(c[1]) + fabs(c[2]) + hw + hh);
}
// smallest m . x over the thickened rectangle
ST_INLINE double rs_min_rect(const double m[3], const RectCand& R) {
    return rs_dot(m, R.c) - fabs(rs_dot(m, R.X)) * (R.hw + R.eps) - fabs(rs_dot(m, R.Y)) * (R.hh + R.eps) -
           fabs(rs_dot(m, R.N)) * R.eps;
}
// rect_sweep_misses for a flat candidate, tighter than its box: the same planes
// against the rectangle itself, and the candidate's own plane, oriented against
// the cone, when every cone direction leaves on one side of it (rays that start
// on that side, as all of R does, never reach it)
ST_INLINE bool rect_sweep_misses_rect(const RectSweep& S, const RectCand& R) {
    for (int j = 0; j < S.nm; ++j)
        if (rs_max_r(S, S.M[j]) < rs_min_rect(S.M[j], R)) return true;
    const double d[3] = {R.c[0] - S.pc[0], R.c[1] - S.pc[1], R.c[2] - S.pc[2]};
    double m[3];
    if (rs_plane_of(S, d, 1.0, m) && rs_max_r(S, m) < rs_min_rect(m, R)) return true;
    const double na = rs_dot(R.N, S.a);
    if (fabs(na) >= S.sph) {
        const double s = na > 0.0 ? -1.0 : 1.0;
        const double mn[3] = {s * R.N[0], s * R.N[1], s * R.N[2]};
        if (rs_max_r(S, mn) < rs_min_rect(mn, R)) return true;
    }
    return false;
}

// Cone from the apex (origin) with unit axis a and half angle phi; pieces of
// the cone are bounded by boxes over [0, first], then [t, t * ratio] up to the
// far corner of the tested box (ratio fixed per launch from the scene size).
// Whether a ray that leaves the sphere (ca, ra) in a direction within angle
// phi of the unit axis t (cp = cos phi, sp = sin phi) can enter the sphere
// (cb, rb): the directions from the first sphere's points to the second's
// fill the cone of half-angle asin(R / |d|) around d = cb - ca (R = ra + rb,
// and every direction when |d| <= R), which meets the ray cone when the angle
// between t and d is at most phi + asin(R / |d|). R grows by 1e-9 of the
// coordinates' size for the FP64 rounding of the test.
ST_INLINE bool sphere_cone_reaches(const double ca[3], double ra, const double t[3], double cp, double sp,
                                   const double cb[3], double rb) {
    const double d[3] = {cb[0] - ca[0], cb[1] - ca[1], cb[2] - ca[2]};
    const double L = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    const double R = ra + rb + 1e-9 * (L + fabs(ca[0]) + fabs(ca[1]) + fabs(ca[2]));
    if (L <= R) return true;
    const double sn = R / L, cs = sqrt(fmax(0.0, 1.0 - sn * sn));
    return d[0] * t[0] + d[1] * t[1] + d[2] * t[2] >= L * (cp * cs - sp * sn);
}

struct Cone {
    double a[3];
    double tp, cp, sp;  // tan, cos, sin of phi
    double ratio;
};
constexpr int kConePieces = 32;
constexpr double kConeFirst = 0.25;

ST_INLINE bool cone_meets_box(const Cone& C, const double lo[3], const double hi[3]) {
    // the cone misses the box if it misses the box's bounding sphere: the
    // sphere spans asin(r / |c|) around its centre direction
    double r2 = 0.0, d2 = 0.0, ac = 0.0, far2 = 0.0, smin = 0.0, smax = 0.0;
    for (int k = 0; k < 3; ++k) {
        const double c = 0.5 * (lo[k] + hi[k]), h = 0.5 * (hi[k] - lo[k]);
        r2 += h * h;
        d2 += c * c;
        ac += C.a[k] * c;
        const double m = fmax(fabs(lo[k]), fabs(hi[k]));
        far2 += m * m;
        smin += fmin(C.a[k] * lo[k], C.a[k] * hi[k]);
        smax += fmax(C.a[k] * lo[k], C.a[k] * hi[k]);
    }
    if (d2 > r2) {
        const double d = sqrt(d2), sb = sqrt(r2) / d, cb = sqrt(fmax(0.0, 1.0 - sb * sb));
        if (ac / d < C.cp * cb - C.sp * sb - 1e-9) return false;
    }
    // piece [t0, t1] holds the cone points whose coordinate along the axis is
    // in [t0, t1]; the box's points have theirs in [smin, smax] (widened for
    // rounding), so only the pieces overlapping that range are tested
    const double slack = 1e-9 * (fabs(smin) + fabs(smax) + 1.0);
    smin -= slack;
    smax += slack;
    if (smax < 0.0) return false;
    const double T = sqrt(far2);
    double t0 = 0.0, t1 = fmin(T, kConeFirst);
    for (int i = 0; i < kConePieces; ++i) {
        if (t0 > smax) break;
        if (t1 >= smin) {
            const double r0 = t0 * C.tp, r1 = t1 * C.tp;
            // the piece's box overlaps the box along one axis (constant
            // indices keep the cone in registers)
            auto overlap = [&](double a, double blo, double bhi) __attribute__((always_inline)) {
                const double c0 = a * t0, c1 = a * t1;
                return !(fmax(c0 + r0, c1 + r1) < blo || fmin(c0 - r0, c1 - r1) > bhi);
            };
            if (overlap(C.a[0], lo[0], hi[0]) && overlap(C.a[1], lo[1], hi[1]) && overlap(C.a[2], lo[2], hi[2]))
                return true;
        }
        if (t1 >= T) break;
        t0 = t1;
        t1 = i == kConePieces - 2 ? T : fmin(T, t1 * C.ratio);
    }
    return false;
}

}  // namespace st
