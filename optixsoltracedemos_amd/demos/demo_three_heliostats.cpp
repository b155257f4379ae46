// Three flat heliostats and a flat receiver through the drop-in API, the scene
// of demos/demo_three_heliostats.cpp built the same way (Element + Surface +
// Aperture, add_element, initialize, run, write_output, clean_up).
// usage: demo_three_heliostats [num_rays] [out.csv] [seed]
#include <cmath>
#include <cstdlib>
#include <iostream>
#include <memory>
#include <string>

#include "lib/element.h"
#include "lib/soltrace_system.h"

int main(int argc, char** argv) {
    const int num_rays = argc > 1 ? std::atoi(argv[1]) : 1000000;
    const std::string out = argc > 2 ? argv[2] : "";
    const unsigned long long seed = argc > 3 ? std::strtoull(argv[3], nullptr, 10) : 0ull;
    SolTraceSystem system(num_rays);
    const double el = 60.0 * M_PI / 180.0, az = M_PI;
    const Vector3d sun(std::sin(az) * std::cos(el), std::cos(az) * std::cos(el), std::sin(el));
    const Vector3d receiver_center(0.0, 0.0, 10.0);
    for (double x : {-5.0, 0.0, 5.0}) {
        auto e = std::make_shared<Element>();
        const Vector3d o(x, 5.0, 0.0);
        const Vector3d n = (sun + (receiver_center - o).normalized()).normalized();  // bisector aim
        e->set_origin(o);
        e->set_aim_point(o + n * 100.0);
        e->set_zrot(0.0);
        e->set_surface(std::make_shared<SurfaceFlat>());
        e->set_aperture(std::make_shared<ApertureRectangle>(1.0, 1.95));
        system.add_element(e);
    }
    auto r = std::make_shared<Element>();
    r->set_origin(receiver_center);
    r->set_aim_point(Vector3d(0.0, 5.0, 0.0));
    r->set_zrot(0.0);
    r->set_surface(std::make_shared<SurfaceFlat>());
    r->set_aperture(std::make_shared<ApertureRectangle>(2.0, 2.0));
    system.add_element(r);
    system.set_sun_vector(sun);
    system.set_sun_angle(0.00465);
    system.set_position_sampling(ST_POS_PHILOX);
    system.set_seed(seed);
    system.set_flux_grid(64, 64);
    system.initialize();
    system.run();
    const auto st = system.get_stats();
    const auto hits = system.get_element_hits();
    std::cout << "num_rays, " << num_rays << ", timing_setup, " << system.get_time_setup() << ", timing_trace, "
              << system.get_time_trace() << ", launch_memory_mb, " << system.get_launch_memory_mb() << std::endl;
    std::cout << "receiver_hits, " << st[ST_STAT_RECEIVER_HITS] << ", mirror_hits, " << st[ST_STAT_MIRROR_HITS]
              << ", flux_binned, " << st[ST_STAT_FLUX_BINNED] << ", element_hits";
    for (auto h : hits) std::cout << ", " << h;
    std::cout << std::endl;
    if (!out.empty()) system.write_output_with_elements(out);
    system.clean_up();
    return 0;
}
