// math_util.h - FP64 orientation helpers of the drop-in API (lib/math_util.h).
#pragma once

#include <cmath>

#include "vector3d.h"
#include "vector_types.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

namespace mathUtil {

// normal -> (yaw, pitch, roll) in radians, SolTrace convention
// (lib/math_util.h:21-32): yaw = atan2(n_x, n_z), pitch = asin(n_y), roll = zrot.
inline Vector3d normal_to_euler(const Vector3d& normal, double zrot_deg) {
    const Vector3d n = normal.normalized();
    return Vector3d(std::atan2(n[0], n[2]), std::asin(n[1]), zrot_deg * M_PI / 180.0);
}

// global-to-local rotation of Spencer & Murty (eq. 2), lib/math_util.h:41-59
inline Matrix33d get_rotation_matrix_G2L(const Vector3d& euler) {
    const double ca = std::cos(euler[0]), sa = std::sin(euler[0]);
    const double cb = std::cos(euler[1]), sb = std::sin(euler[1]);
    const double cg = std::cos(euler[2]), sg = std::sin(euler[2]);
    return Matrix33d{ca * cg + sa * sb * sg, -cb * sg, -sa * cg + ca * sb * sg,
                     ca * sg - sa * sb * cg, cb * cg,  -sa * sg - ca * sb * cg,
                     sa * cb,                sb,       ca * cb};
}

// p_global = G2L^T p_local + origin
inline Vector3d local_to_global(const Vector3d& p, const Matrix33d& g2l, const Vector3d& origin) {
    return g2l.transpose() * p + origin;
}

// p_local = G2L (p_global - origin)
inline Vector3d global_to_local(const Vector3d& p, const Matrix33d& g2l, const Vector3d& origin) {
    return g2l * (p - origin);
}

inline float3 toFloat3(const Vector3d& v) { return make_float3((float)v[0], (float)v[1], (float)v[2]); }

}  // namespace mathUtil
