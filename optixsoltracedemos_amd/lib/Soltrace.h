// Soltrace.h - launch-parameter view of the drop-in API (cuda/Soltrace.h:8-58).
// The OptiX traversable handle becomes the tracer's scene handle; the device
// buffers are owned through the C ABI.
#pragma once

#include "GeometryDataST.h"
#include "soltrace_hip.h"
#include "vector_types.h"

namespace soltrace {

const unsigned int MAX_TRACE_DEPTH = 5u;

enum OpticalEntityType : unsigned int {
    RECTANGLE_FLAT_MIRROR = 0,
    RECTANGLE_PARABOLIC_MIRROR = 1,
    RECTANGLE_FLAT_RECEIVER = 2,
    CYLINDRICAL_RECEIVER = 3,
    NUM_OPTICAL_ENTITY_TYPES
};

struct LaunchParams {
    unsigned int width = 10;   // rays launched
    unsigned int height = 1;
    int max_depth = 5;
    float4* hit_point_buffer = nullptr;  // device, width*height*max_depth (tag, x, y, z)
    st_scene handle = nullptr;
    float3 sun_vector{0.f, 0.f, 10.f};
    float max_sun_angle = 0.f;
    float3 sun_v0{0, 0, 0}, sun_v1{0, 0, 0}, sun_v2{0, 0, 0}, sun_v3{0, 0, 0};
    GeometryDataST* geometry_data_array = nullptr;  // device copy of the AoS records
};

struct PerRayData {
    unsigned int ray_path_index;
    unsigned int depth;
};

}  // namespace soltrace
