// STINPUT-driven trace, like demos/demo_large_scene.cpp: read the scene file,
// set the sun angle, initialize, run, report timing, optionally dump the CSV.
// usage: demo_stinput file.stinput num_rays [out.csv] [use_sun_shape 0|1]
#include <cstdlib>
#include <iostream>
#include <string>

#include "lib/soltrace_system.h"

int main(int argc, char** argv) {
    if (argc < 3) {
        std::cout << "Usage: " << argv[0] << " <file.stinput> <num_rays> [out.csv] [use_sun_shape]" << std::endl;
        return 1;
    }
    const int num_rays = std::atoi(argv[2]);
    SolTraceSystem system(num_rays);
    if (!system.read_st_input(argv[1])) {
        std::cerr << "failed to read " << argv[1] << std::endl;
        return 2;
    }
    const bool sun_shape = argc > 4 ? std::atoi(argv[4]) != 0 : true;
    if (!sun_shape) system.set_sun_angle(0.0);
    system.initialize();
    system.run();
    const auto st = system.get_stats();
    std::cout << "elements, " << system.get_num_heliostats() + 1 << ", num_rays, " << num_rays << ", timing_setup, "
              << system.get_time_setup() << ", timing_trace, " << system.get_time_trace() << ", receiver_hits, "
              << st[ST_STAT_RECEIVER_HITS] << std::endl;
    if (argc > 3 && std::string(argv[3]) != "-") system.write_output(argv[3]);
    system.clean_up();
    return 0;
}
