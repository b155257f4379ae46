// vector_types.h - host-side float3 / float4 PODs used by the drop-in API
// (the reference takes them from CUDA's vector_types.h). Host-only headers:
// do not mix with HIP's own vector types in one translation unit.
#pragma once

struct float3 {
    float x, y, z;
};
struct float4 {
    float x, y, z, w;
};
inline float3 make_float3(float x, float y, float z) { return {x, y, z}; }
inline float4 make_float4(float x, float y, float z, float w) { return {x, y, z, w}; }
