// A sweep of sun positions over one field through the drop-in API, in the
// reference's frame loop (demos/demo_dynamic_scene.cpp:82-97: move the sun,
// update(), run()): per step every heliostat is re-aimed on the host
// (Element::update_element with the FP64 bisector of the sun and the receiver
// centre, as lib/element.cpp:89-93 is driven), then update() re-collects,
// uploads and refits the scene and rebuilds its sun plane, run() traces, and
// the flux map is read back. Flux only (set_record_hits(false)), stratified
// sun positions, Philox stream (seed, job = step), like st_trace_batch.
//
// usage: demo_sweep elements.bin suns.bin budgets.bin out_flux.bin [nx ny]
//   elements.bin: st_element_f64[n] (last = receiver); suns.bin: K x 3 doubles;
//   budgets.bin: K uint64; out_flux.bin: K x ny x nx int64 written after the loop.
// Prints "steps, K, rays, R, loop_s, T" (T: wall time of the K steps).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <memory>
#include <vector>

#include "lib/element.h"
#include "lib/soltrace_system.h"

template <class T>
static std::vector<T> read_all(const char* path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error(std::string("cannot read ") + path);
    f.seekg(0, std::ios::end);
    const size_t bytes = (size_t)f.tellg();
    f.seekg(0);
    std::vector<T> v(bytes / sizeof(T));
    f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(v.size() * sizeof(T)));
    return v;
}

int main(int argc, char** argv) {
    if (argc < 5) {
        std::cerr << "usage: " << argv[0] << " elements.bin suns.bin budgets.bin out_flux.bin [nx ny]" << std::endl;
        return 1;
    }
    const int nx = argc > 6 ? std::atoi(argv[5]) : 256, ny = argc > 6 ? std::atoi(argv[6]) : 128;
    const std::vector<st_element_f64> el = read_all<st_element_f64>(argv[1]);
    const std::vector<double> suns = read_all<double>(argv[2]);
    const std::vector<uint64_t> budgets = read_all<uint64_t>(argv[3]);
    const size_t K = budgets.size();
    if (el.size() < 2 || suns.size() != 3 * K) {
        std::cerr << "inconsistent inputs" << std::endl;
        return 1;
    }
    SolTraceSystem sys((int)budgets[0]);
    std::vector<std::shared_ptr<Element>> helio;
    for (const st_element_f64& e : el) {
        auto x = std::make_shared<Element>();
        x->set_origin(Vector3d(e.origin[0], e.origin[1], e.origin[2]));
        x->set_aim_point(Vector3d(e.aim_point[0], e.aim_point[1], e.aim_point[2]));
        x->set_zrot(e.zrot_deg);
        x->set_aperture(std::make_shared<ApertureRectangle>(e.dim_x, e.dim_y));
        if (e.surface == ST_SURFACE_CYLINDER) {
            auto c = std::make_shared<SurfaceCylinder>();
            c->set_radius(0.5 * e.dim_x);
            c->set_half_height(0.5 * e.dim_y);
            x->set_surface(c);
        } else {
            x->set_surface(std::make_shared<SurfaceFlat>());
        }
        x->set_reflectivity(e.reflectivity);
        x->set_slope_error(e.slope_error);
        x->set_specularity_error(e.spec_error);
        sys.add_element(x);
        if (e.role == ST_ROLE_MIRROR) helio.push_back(x);
    }
    const st_element_f64& rec = el.back();
    const Vector3d target(rec.origin[0], rec.origin[1], rec.origin[2]);
    sys.set_sun_vector(Vector3d(suns[0], suns[1], suns[2]));
    sys.set_sun_angle(4.65e-3);
    sys.set_position_sampling(ST_POS_STRATIFIED);
    sys.set_record_hits(false);
    sys.set_flux_grid(nx, ny);
    sys.set_seed(0);
    sys.initialize();
    std::vector<int64_t> flux((size_t)K * nx * ny);
    uint64_t rays = 0;
    const auto t0 = std::chrono::steady_clock::now();
    for (size_t k = 0; k < K; ++k) {
        const Vector3d s = Vector3d(suns[3 * k], suns[3 * k + 1], suns[3 * k + 2]).normalized();
        for (const auto& h : helio) {
            // bisector of the sun and the direction to the receiver, 100 m out
            const Vector3d o = h->get_origin();
            Vector3d t = target - o;
            t = t.normalized();
            const Vector3d n = (s + t).normalized();
            h->update_element(o + n * 100.0, h->get_zrot());
        }
        sys.set_sun_vector(s);
        sys.set_sun_points((int)budgets[k]);
        sys.set_job_id((uint32_t)k);
        sys.update();
        sys.run();
        const std::vector<int64_t> f = sys.get_flux_map();
        std::copy(f.begin(), f.end(), flux.begin() + (ptrdiff_t)(k * (size_t)nx * ny));
        rays += budgets[k];
    }
    const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::ofstream out(argv[4], std::ios::binary);
    out.write(reinterpret_cast<const char*>(flux.data()), (std::streamsize)(flux.size() * sizeof(int64_t)));
    std::cout << "steps, " << K << ", rays, " << rays << ", loop_s, " << dt << std::endl;
    sys.clean_up();
    return 0;
}
