// Moving-sun sequence through run()/update(), the loop of
// demos/demo_dynamic_scene.cpp:82-97 (sun y component +1 per frame).
// usage: demo_dynamic_scene [num_rays] [frames]
#include <cstdlib>
#include <iostream>
#include <memory>

#include "lib/element.h"
#include "lib/soltrace_system.h"

int main(int argc, char** argv) {
    const int num_rays = argc > 1 ? std::atoi(argv[1]) : 100000;
    const int frames = argc > 2 ? std::atoi(argv[2]) : 40;
    SolTraceSystem system(num_rays);
    auto e = std::make_shared<Element>();
    e->set_origin(Vector3d(0, 5, 0));
    e->set_aim_point(Vector3d(0, -17.360680, 94.721360));
    e->set_zrot(0.0);
    auto s = std::make_shared<SurfaceParabolic>();
    s->set_curvature(0.0170679, 0.0370679);
    e->set_surface(s);
    e->set_aperture(std::make_shared<ApertureRectangle>(1.0, 1.95));
    system.add_element(e);
    auto r = std::make_shared<Element>();
    r->set_origin(Vector3d(0, 0, 10.0));
    r->set_aim_point(Vector3d(0, 5, 0.0));
    r->set_surface(std::make_shared<SurfaceFlat>());
    r->set_aperture(std::make_shared<ApertureRectangle>(2.0, 4.0));
    system.add_element(r);
    Vector3d sun(0.0, -20.0, 100.0);
    system.set_sun_vector(sun);
    system.set_sun_angle(0.0);
    system.set_position_sampling(ST_POS_PHILOX);
    system.initialize();
    for (int f = 0; f < frames; ++f) {
        system.run();
        const auto st = system.get_stats();
        std::cout << "frame, " << f << ", sun_y, " << sun[1] << ", mirror_hits, " << st[ST_STAT_MIRROR_HITS]
                  << ", receiver_hits, " << st[ST_STAT_RECEIVER_HITS] << ", direct, " << st[ST_STAT_DIRECT_SUN]
                  << std::endl;
        sun[1] += 1;
        system.set_sun_vector(sun);
        system.update();
    }
    std::cout << "timing_trace_total, " << system.get_time_trace() << std::endl;
    system.clean_up();
    return 0;
}
