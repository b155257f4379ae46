// STINPUT-driven trace, like demos/demo_large_scene.cpp: read the scene file,
// set the sun angle, initialize, run, report timing, optionally dump the CSV.
// usage: demo_stinput file.stinput num_rays [out.csv|-] [use_sun_shape 0|1]
//                     [optics] [seed=N] [flux=NXxNY:counts.txt]
//   optics: the optics named in the file reach the trace (set_use_stinput_optics)
//   flux:   bin the receiver into NX x NY and write the counts, one row of NX per line
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>

#include "lib/soltrace_system.h"

int main(int argc, char** argv) {
    if (argc < 3) {
        std::cout << "Usage: " << argv[0]
                  << " <file.stinput> <num_rays> [out.csv|-] [use_sun_shape] [optics] [seed=N] [flux=NXxNY:path]"
                  << std::endl;
        return 1;
    }
    const int num_rays = std::atoi(argv[2]);
    SolTraceSystem system(num_rays);
    int nx = 0, ny = 0;
    std::string flux_path;
    for (int i = 5; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "optics") {
            system.set_use_stinput_optics(true);
        } else if (a.rfind("seed=", 0) == 0) {
            system.set_seed(std::strtoull(a.c_str() + 5, nullptr, 10));
        } else if (a.rfind("flux=", 0) == 0) {
            const size_t colon = a.find(':');
            if (colon == std::string::npos || std::sscanf(a.c_str() + 5, "%dx%d", &nx, &ny) != 2) {
                std::cerr << "bad flux option " << a << std::endl;
                return 1;
            }
            flux_path = a.substr(colon + 1);
            system.set_flux_grid(nx, ny);
        } else {
            std::cerr << "unknown option " << a << std::endl;
            return 1;
        }
    }
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
              << st[ST_STAT_RECEIVER_HITS] << ", mirror_hits, " << st[ST_STAT_MIRROR_HITS] << ", absorbed, "
              << st[ST_STAT_ABSORBED] << std::endl;
    std::cout << "element_hits";
    for (uint64_t h : system.get_element_hits()) std::cout << ", " << h;
    std::cout << std::endl;
    if (!flux_path.empty()) {
        const std::vector<int64_t> f = system.get_flux_map();
        std::ofstream out(flux_path);
        for (int y = 0; y < ny; ++y) {
            for (int x = 0; x < nx; ++x) out << (x ? " " : "") << f[(size_t)y * nx + x];
            out << "\n";
        }
    }
    if (argc > 3 && std::string(argv[3]) != "-") system.write_output(argv[3]);
    system.clean_up();
    return 0;
}
