// GeometryDataST.h - host view of one element's FP32 primitive, kept for
// source compatibility with Element::toDeviceGeometryData()
// (cuda/GeometryDataST.h:9-170). The device never reads this AoS record: the
// tracer packs its own 64-byte primitive records on the GPU.
#pragma once

#include <cmath>

#include "vector_types.h"

struct GeometryDataST {
    enum Type { PARALLELOGRAM = 0, CYLINDER_Y = 1, RECTANGLE_PARABOLIC = 2, UNKNOWN_TYPE = 3, RECTANGLE_FLAT = 4 };

    // plane (normal, d) from two in-plane vectors and a point on the plane
    static float4 plane_of(float3 a, float3 b, float3 p) {
        float nx = a.y * b.z - a.z * b.y, ny = a.z * b.x - a.x * b.z, nz = a.x * b.y - a.y * b.x;
        const float l = std::sqrt(nx * nx + ny * ny + nz * nz);
        nx /= l;
        ny /= l;
        nz /= l;
        return make_float4(nx, ny, nz, nx * p.x + ny * p.y + nz * p.z);
    }
    static float3 scaled_inverse(float3 v) {  // v / |v|^2
        const float s = 1.0f / (v.x * v.x + v.y * v.y + v.z * v.z);
        return make_float3(v.x * s, v.y * s, v.z * s);
    }

    struct Parallelogram {
        Parallelogram() = default;
        Parallelogram(float3 v1_, float3 v2_, float3 anchor_)
            : plane(plane_of(v1_, v2_, anchor_)), v1(scaled_inverse(v1_)), v2(scaled_inverse(v2_)), anchor(anchor_) {}
        float4 plane{};
        float3 v1{}, v2{}, anchor{};
    };
    struct Rectangle_Flat {
        Rectangle_Flat() = default;
        Rectangle_Flat(float3 c, float3 x_, float3 y_, float w, float h)
            : plane(plane_of(x_, y_, c)), center(c), x(x_), y(y_), width(w), height(h) {}
        float4 plane{};
        float3 center{}, x{}, y{};
        float width = 0, height = 0;
    };
    struct Cylinder_Y {
        Cylinder_Y() = default;
        Cylinder_Y(float3 c, float r, float hh, float3 bx, float3 bz)
            : center(c), radius(r), half_height(hh), base_x(bx), base_z(bz) {}
        float3 center{};
        float radius = 0, half_height = 0;
        float3 base_x{}, base_z{};
    };
    struct Rectangle_Parabolic {
        Rectangle_Parabolic() = default;
        Rectangle_Parabolic(float3 v1_, float3 v2_, float3 anchor_, float cx, float cy)
            : plane(plane_of(v1_, v2_, anchor_)), v1(scaled_inverse(v1_)), v2(scaled_inverse(v2_)), anchor(anchor_),
              curv_x(cx), curv_y(cy) {}
        float4 plane{};
        float3 v1{}, v2{}, anchor{};
        float curv_x = 0, curv_y = 0;
    };

    void setParallelogram(const Parallelogram& p) { type = PARALLELOGRAM; parallelogram = p; }
    const Parallelogram& getParallelogram() const { return parallelogram; }
    void setRectangle_Flat(const Rectangle_Flat& r) { type = RECTANGLE_FLAT; rectangle_flat = r; }
    const Rectangle_Flat& getRectangle_Flat() const { return rectangle_flat; }
    void setCylinder_Y(const Cylinder_Y& c) { type = CYLINDER_Y; cylinder_y = c; }
    const Cylinder_Y& getCylinder_Y() const { return cylinder_y; }
    void setRectangleParabolic(const Rectangle_Parabolic& r) { type = RECTANGLE_PARABOLIC; rectangle_parabolic = r; }
    const Rectangle_Parabolic& getRectangleParabolic() const { return rectangle_parabolic; }

    Type type = UNKNOWN_TYPE;

private:
    Parallelogram parallelogram;
    Cylinder_Y cylinder_y;
    Rectangle_Parabolic rectangle_parabolic;
    Rectangle_Flat rectangle_flat;
};
