// CPU check of the C++ drop-in layer: STINPUT parsing, element frames, the
// host GeometryDataST view and the element records (with optics) handed to the
// tracer; prints one line per element for the Python test.
// usage: dropin_parse file.stinput [optics]
#include <cstdio>
#include <cstring>
#include <memory>

#include "lib/element.h"
#include "lib/soltrace_system.h"

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    SolTraceSystem sys(1000);
    if (argc > 2 && std::strcmp(argv[2], "optics") == 0) sys.set_use_stinput_optics(true);
    if (!sys.read_st_input(argv[1])) {
        std::printf("parse failed\n");
        return 1;
    }
    const Vector3d s = sys.get_sun_vector();
    std::printf("sun %.9g %.9g %.9g angle %.9g shape %d heliostats %zu\n", s[0], s[1], s[2], sys.get_sun_angle(),
                sys.get_sun_shape(), sys.get_num_heliostats());
    const std::vector<st_element_f64> rec = sys.get_element_records();
    size_t i = 0;
    for (const auto& e : sys.get_elements()) {
        e->compute_bounding_box();
        const Matrix33d R = e->get_rotation_matrix();
        const Vector3d lo = e->get_lower_bounding_box(), hi = e->get_upper_bounding_box();
        const GeometryDataST g = e->toDeviceGeometryData();
        const st_element_f64& r = rec[i++];
        std::printf("elem surf %d ap %d o %.9g %.9g %.9g aim %.9g %.9g %.9g zrot %.9g dims %.9g %.9g curv %.9g %.9g "
                    "z %.12g %.12g %.12g x %.12g %.12g %.12g lo %.9g %.9g %.9g hi %.9g %.9g %.9g gtype %d "
                    "role %d rho %.9g slope %.9g spec %.9g\n",
                    (int)e->get_surface_type(), (int)e->get_aperture_type(), e->get_origin()[0], e->get_origin()[1],
                    e->get_origin()[2], e->get_aim_point()[0], e->get_aim_point()[1], e->get_aim_point()[2],
                    e->get_zrot(), e->get_aperture()->get_width(), e->get_aperture()->get_height(),
                    e->get_surface()->get_curvature_1(), e->get_surface()->get_curvature_2(), R.get_z_basis()[0],
                    R.get_z_basis()[1], R.get_z_basis()[2], R.get_x_basis()[0], R.get_x_basis()[1], R.get_x_basis()[2],
                    lo[0], lo[1], lo[2], hi[0], hi[1], hi[2], (int)g.type, r.role, r.reflectivity, r.slope_error,
                    r.spec_error);
    }
    return 0;
}
