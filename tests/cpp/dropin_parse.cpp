// CPU check of the C++ drop-in layer: STINPUT parsing, element frames and the
// host GeometryDataST view; prints one line per element for the Python test.
// usage: dropin_parse file.stinput
#include <cstdio>
#include <memory>

#include "lib/element.h"
#include "lib/soltrace_system.h"

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    SolTraceSystem sys(1000);
    if (!sys.read_st_input(argv[1])) {
        std::printf("parse failed\n");
        return 1;
    }
    const Vector3d s = sys.get_sun_vector();
    std::printf("sun %.9g %.9g %.9g angle %.9g shape %d heliostats %zu\n", s[0], s[1], s[2], sys.get_sun_angle(),
                sys.get_sun_shape(), sys.get_num_heliostats());
    for (const auto& e : sys.get_elements()) {
        e->compute_bounding_box();
        const Matrix33d R = e->get_rotation_matrix();
        const Vector3d lo = e->get_lower_bounding_box(), hi = e->get_upper_bounding_box();
        const GeometryDataST g = e->toDeviceGeometryData();
        std::printf("elem surf %d ap %d o %.9g %.9g %.9g aim %.9g %.9g %.9g zrot %.9g dims %.9g %.9g curv %.9g %.9g "
                    "z %.12g %.12g %.12g x %.12g %.12g %.12g lo %.9g %.9g %.9g hi %.9g %.9g %.9g gtype %d\n",
                    (int)e->get_surface_type(), (int)e->get_aperture_type(), e->get_origin()[0], e->get_origin()[1],
                    e->get_origin()[2], e->get_aim_point()[0], e->get_aim_point()[1], e->get_aim_point()[2],
                    e->get_zrot(), e->get_aperture()->get_width(), e->get_aperture()->get_height(),
                    e->get_surface()->get_curvature_1(), e->get_surface()->get_curvature_2(), R.get_z_basis()[0],
                    R.get_z_basis()[1], R.get_z_basis()[2], R.get_x_basis()[0], R.get_x_basis()[1], R.get_x_basis()[2],
                    lo[0], lo[1], lo[2], hi[0], hi[1], hi[2], (int)g.type);
    }
    return 0;
}
