// soltrace_type.h - aperture / surface enums of the drop-in API (lib/soltrace_type.h).
#pragma once

enum class ApertureType { RECTANGLE, CIRCLE };

enum class SurfaceType { FLAT, PARABOLIC, MESH, CYLINDER };

struct SurfaceApertureMap {
    SurfaceType surfaceType;
    ApertureType apertureType;
    bool operator==(const SurfaceApertureMap& o) const {
        return surfaceType == o.surfaceType && apertureType == o.apertureType;
    }
};
