// soltrace_system.h - the system facade of the drop-in API
// (lib/soltrace_system.h:21-121): same constructor, lifecycle, setters and
// timers; the OptiX pipeline behind it is replaced by the HIP tracer's C ABI.
#pragma once

#include <cstdint>
#include <cstdio>
#include <memory>
#include <string>
#include <vector>

#include "soltrace_hip.h"
#include "timer.h"
#include "vector3d.h"

class dataManager;
class Element;
namespace stinput {
struct Document;
}

class SolTraceSystem {
public:
    explicit SolTraceSystem(int numSunPoints);
    ~SolTraceSystem();

    void initialize();   // elements -> device scene, sun plane, buffers
    void run();          // trace (timed like lib/soltrace_system.cpp:154-167)
    void update();       // re-read element poses and the sun, refit, new sun plane
    bool read_st_input(const char* filename);
    void write_output(const std::string& filename);
    void clean_up();

    void set_verbose(bool verbose) { m_verbose = verbose; }
    void set_sun_points(int num) { m_num_sunpoints = num; }
    void set_sun_vector(Vector3d vect);
    void set_sun_angle(double angle) { m_sun_angle = angle; }  // radians
    size_t get_num_heliostats() const { return m_element_list.size() - 1; }
    int get_num_receivers() const { return 1; }
    void add_element(std::shared_ptr<Element> element);
    double get_time_trace() { return m_timer_trace.get_time_sec(); }
    double get_time_setup() { return m_timer_setup.get_time_sec(); }
    void print_launch_params();
    // device memory taken between the start of initialize() and the start of
    // the last run() (free memory before minus after, in MiB), the figure the
    // reference prints in run() (lib/soltrace_system.cpp:77, 149-151)
    double get_launch_memory_mb() const { return m_launch_memory_mb; }

    // ---- extensions (no reference counterpart) -------------------------------
    void set_sun_shape(int st_sun_shape) { m_sun_shape = st_sun_shape; }        // ST_SUN_PILLBOX / GAUSSIAN
    void set_position_sampling(int mode) { m_position_sampling = mode; }        // ST_POS_HALTON (default) / PHILOX
    void set_seed(uint64_t seed) { m_seed = seed; }
    void set_job_id(uint32_t job) { m_job_id = job; }                           // Philox stream of run()
    void set_max_depth(int d) { m_max_depth = d; }
    void set_dni(double dni) { m_dni = dni; }
    void set_flux_grid(int nx, int ny) { m_flux_nx = nx; m_flux_ny = ny; }     // bins on the receiver
    void set_record_hits(bool on) { m_record_hits = on; }                     // hit_point_buffer on/off
    void set_include_direct_sun_in_flux(bool on) { m_flux_direct = on; }
    void set_use_stinput_optics(bool on) { m_use_stinput_optics = on; }
    void write_output_with_elements(const std::string& filename);               // CSV + element column
    std::vector<int64_t> get_flux_map() const;          // [ny][nx] hit counts
    std::vector<double> get_flux_map_watts() const;     // W/m^2
    std::vector<uint64_t> get_element_hits() const;
    std::vector<uint64_t> get_stats() const;            // ST_STAT_*
    double get_power_per_ray() const;
    st_sun_plane get_sun_plane() const;
    Vector3d get_sun_vector() const { return m_sun_vector; }
    double get_sun_angle() const { return m_sun_angle; }
    int get_sun_shape() const { return m_sun_shape; }
    st_scene get_scene() const { return m_scene; }
    const std::vector<std::shared_ptr<Element>>& get_elements() const { return m_element_list; }
    // the FP64 element records initialize() hands to the tracer (optics included)
    std::vector<st_element_f64> get_element_records() const { return collect_elements(); }

private:
    std::vector<st_element_f64> collect_elements() const;
    void write_csv(const std::string& filename, bool with_elements);

    bool apply_stinput(const stinput::Document& doc);

    std::shared_ptr<dataManager> data_manager;
    int m_num_sunpoints;
    bool m_verbose = false;
    Vector3d m_sun_vector{0.0, 0.0, 1.0};
    double m_sun_angle = 0.0;
    int m_sun_shape = ST_SUN_PILLBOX;
    int m_position_sampling = ST_POS_HALTON;
    uint64_t m_seed = 0;
    uint32_t m_job_id = 0;
    int m_max_depth = 5;
    double m_dni = 1000.0;
    int m_flux_nx = 0, m_flux_ny = 0;
    bool m_record_hits = true, m_flux_direct = false, m_use_stinput_optics = false;

    std::vector<std::shared_ptr<Element>> m_element_list;
    std::vector<std::string> m_element_optic;  // STINPUT optic name per element
    struct Optic {
        std::string name;
        double reflectivity, slope_mrad, spec_mrad;
    };
    std::vector<Optic> m_optics;

    st_context m_ctx = nullptr;
    st_scene m_scene = nullptr;
    uint32_t* m_record_elements = nullptr;  // device
    int64_t* m_flux = nullptr;              // device
    uint64_t* m_hits = nullptr;             // device
    uint64_t* m_stats = nullptr;            // device
    size_t m_traced_rays = 0;
    int m_traced_depth = 0;
    size_t m_mem_free_before = 0;
    double m_launch_memory_mb = 0.0;

    Timer m_timer_setup;
    Timer m_timer_trace;
};
