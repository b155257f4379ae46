// soltrace_system.cpp - SolTraceSystem over the HIP tracer's C ABI.
//
// Lifecycle mirrors lib/soltrace_system.cpp:49-297 (initialize / run / update /
// write_output / clean_up); the STINPUT reader follows the format handled at
// :449-846 and the CSV writer the one at :208-264.
#include "soltrace_system.h"

#include <cctype>
#include <cmath>
#include <cstring>
#include <fstream>
#include <iostream>
#include <stdexcept>

#include "data_manager.h"
#include "element.h"
#include "stinput.h"

namespace {
void check(int rc, const char* what) {
    if (rc != ST_OK) throw std::runtime_error(std::string(what) + ": " + st_last_error());
}
}  // namespace

SolTraceSystem::SolTraceSystem(int numSunPoints)
    : data_manager(std::make_shared<dataManager>()), m_num_sunpoints(numSunPoints) {}

SolTraceSystem::~SolTraceSystem() {
    try {
        clean_up();
    } catch (...) {
    }
}

void SolTraceSystem::add_element(std::shared_ptr<Element> e) {
    e->update_euler_angles();
    m_element_list.push_back(e);
    m_element_optic.emplace_back();
}

void SolTraceSystem::set_sun_vector(Vector3d vect) {
    m_sun_vector = vect;
    data_manager->launch_params_H.sun_vector = mathUtil::toFloat3(vect.normalized());
}

std::vector<st_element_f64> SolTraceSystem::collect_elements() const {
    std::vector<st_element_f64> out(m_element_list.size());
    for (size_t i = 0; i < m_element_list.size(); ++i) {
        const Element& e = *m_element_list[i];
        st_element_f64& d = out[i];
        std::memset(&d, 0, sizeof(d));
        for (int k = 0; k < 3; ++k) {
            d.origin[k] = e.get_origin()[k];
            d.aim_point[k] = e.get_aim_point()[k];
        }
        d.zrot_deg = e.get_zrot();
        if (!e.get_aperture() || !e.get_surface()) throw std::runtime_error("element without aperture or surface");
        d.aperture = e.get_aperture_type() == ApertureType::RECTANGLE ? ST_APERTURE_RECTANGLE : ST_APERTURE_CIRCLE;
        d.dim_x = e.get_aperture()->get_width();
        d.dim_y = e.get_aperture()->get_height();
        switch (e.get_surface_type()) {
            case SurfaceType::FLAT: d.surface = ST_SURFACE_FLAT; break;
            case SurfaceType::PARABOLIC:
                d.surface = ST_SURFACE_PARABOLIC;
                d.curv_x = e.get_surface()->get_curvature_1();
                d.curv_y = e.get_surface()->get_curvature_2();
                break;
            case SurfaceType::CYLINDER: d.surface = ST_SURFACE_CYLINDER; break;
            default: throw std::runtime_error("unsupported surface type for element " + std::to_string(i));
        }
        // the receiver is the last element (lib/geometry_manager.cpp:78-90)
        d.role = i + 1 == m_element_list.size() ? ST_ROLE_RECEIVER : ST_ROLE_MIRROR;
        d.reflectivity = e.get_reflectivity();
        d.slope_error = e.get_slope_error();
        d.spec_error = e.get_specularity_error();
        if (m_use_stinput_optics && !m_element_optic[i].empty()) {
            for (const Optic& o : m_optics)
                if (o.name == m_element_optic[i]) {
                    d.reflectivity = o.reflectivity;
                    d.slope_error = o.slope_mrad * 1e-3;
                    d.spec_error = o.spec_mrad * 1e-3;
                }
        }
    }
    return out;
}

void SolTraceSystem::initialize() {
    m_timer_setup.start();
    if (m_element_list.size() < 2) throw std::runtime_error("SolTraceSystem needs at least one heliostat and a receiver");
    if (m_num_sunpoints <= 0) throw std::runtime_error("number of sun points must be positive");
    if (!m_ctx) check(st_context_create(0, &m_ctx), "st_context_create");
    check(st_device_memory(0, &m_mem_free_before, nullptr), "st_device_memory");
    data_manager->set_context(m_ctx);
    const std::vector<st_element_f64> elems = collect_elements();
    if (m_scene) check(st_scene_destroy(m_scene), "st_scene_destroy");
    m_scene = nullptr;
    check(st_scene_create(m_ctx, elems.data(), (int32_t)elems.size(), &m_scene), "st_scene_create");

    st_sun_desc sun;
    std::memset(&sun, 0, sizeof(sun));
    const Vector3d s = m_sun_vector.normalized();
    for (int k = 0; k < 3; ++k) sun.vector[k] = s[k];
    sun.param = m_sun_angle;
    sun.dni = m_dni;
    sun.shape = m_sun_shape;
    sun.position_sampling = m_position_sampling;
    check(st_scene_set_sun(m_scene, &sun), "st_scene_set_sun");

    soltrace::LaunchParams& p = data_manager->launch_params_H;
    p.width = (unsigned)m_num_sunpoints;
    p.height = 1;
    p.max_depth = m_max_depth;
    p.handle = m_scene;
    p.sun_vector = mathUtil::toFloat3(s);
    p.max_sun_angle = (float)m_sun_angle;
    const st_sun_plane pl = get_sun_plane();
    p.sun_v0 = make_float3((float)pl.v0[0], (float)pl.v0[1], (float)pl.v0[2]);
    p.sun_v1 = make_float3((float)pl.v1[0], (float)pl.v1[1], (float)pl.v1[2]);
    p.sun_v2 = make_float3((float)pl.v2[0], (float)pl.v2[1], (float)pl.v2[2]);
    p.sun_v3 = make_float3((float)pl.v3[0], (float)pl.v3[1], (float)pl.v3[2]);

    // output buffers (hit_point_buffer: lib/soltrace_system.cpp:114-125)
    const size_t slots = (size_t)m_num_sunpoints * (size_t)m_max_depth;
    if (p.hit_point_buffer) check(st_buffer_free(m_ctx, p.hit_point_buffer), "free");
    if (m_record_elements) check(st_buffer_free(m_ctx, m_record_elements), "free");
    p.hit_point_buffer = nullptr;
    m_record_elements = nullptr;
    if (m_record_hits) {
        check(st_buffer_alloc(m_ctx, slots * sizeof(float4), (void**)&p.hit_point_buffer), "hit_point_buffer");
        check(st_buffer_zero(m_ctx, p.hit_point_buffer, slots * sizeof(float4)), "hit_point_buffer");
        check(st_buffer_alloc(m_ctx, slots * sizeof(uint32_t), (void**)&m_record_elements), "record elements");
    }
    for (void** b : {(void**)&m_flux, (void**)&m_hits, (void**)&m_stats})
        if (*b) {
            check(st_buffer_free(m_ctx, *b), "free");
            *b = nullptr;
        }
    if (m_flux_nx > 0 && m_flux_ny > 0)
        check(st_buffer_alloc(m_ctx, (size_t)m_flux_nx * m_flux_ny * sizeof(int64_t), (void**)&m_flux), "flux map");
    check(st_buffer_alloc(m_ctx, elems.size() * sizeof(uint64_t), (void**)&m_hits), "element hits");
    check(st_buffer_alloc(m_ctx, ST_STAT_COUNT * sizeof(uint64_t), (void**)&m_stats), "stats");

    // launch_params_H is kept current on the host. The reference uploads it and
    // a GeometryDataST array for its OptiX programs (lib/soltrace_system.cpp:
    // 130-143); no kernel here reads them, so no device copies are made
    // (dataManager still uploads them for a caller that uses it directly).
    if (m_verbose) print_launch_params();
    m_timer_setup.stop();
}

void SolTraceSystem::run() {
    if (!m_scene) throw std::runtime_error("SolTraceSystem::run before initialize");
    st_trace_desc d;
    std::memset(&d, 0, sizeof(d));
    d.num_rays = (uint64_t)m_num_sunpoints;
    d.seed = m_seed;
    d.job_id = m_job_id;
    d.max_depth = m_max_depth;
    d.flags = ST_TRACE_ZERO_OUTPUTS | (m_record_hits ? ST_TRACE_RECORDS : 0u) |
              (m_flux_direct ? ST_TRACE_FLUX_DIRECT_SUN : 0u);
    d.flux_element = -1;
    d.flux_nx = m_flux ? m_flux_nx : 0;
    d.flux_ny = m_flux ? m_flux_ny : 0;
    st_outputs o;
    std::memset(&o, 0, sizeof(o));
    o.flux = m_flux;
    o.element_hits = m_hits;
    o.stats = m_stats;
    o.records = (float*)data_manager->launch_params_H.hit_point_buffer;
    o.record_elements = m_record_elements;
    size_t free_now = 0;
    check(st_device_memory(0, &free_now, nullptr), "st_device_memory");
    m_launch_memory_mb = ((double)m_mem_free_before - (double)free_now) / (1024.0 * 1024.0);
    if (m_verbose) std::cout << "run: " << m_launch_memory_mb << " MiB of device memory taken since initialize()" << std::endl;
    m_timer_trace.start();
    check(st_trace(m_scene, &d, &o), "st_trace");
    m_timer_trace.stop();
    m_traced_rays = (size_t)m_num_sunpoints;
    m_traced_depth = m_max_depth;
}

void SolTraceSystem::update() {
    if (!m_scene) throw std::runtime_error("SolTraceSystem::update before initialize");
    const std::vector<st_element_f64> elems = collect_elements();
    st_sun_desc sun;
    std::memset(&sun, 0, sizeof(sun));
    const Vector3d s = m_sun_vector.normalized();
    for (int k = 0; k < 3; ++k) sun.vector[k] = s[k];
    sun.param = m_sun_angle;
    sun.dni = m_dni;
    sun.shape = m_sun_shape;
    sun.position_sampling = m_position_sampling;
    // new poses and the sun in one set-up (one sun plane and grid)
    check(st_scene_update(m_scene, elems.data(), (int32_t)elems.size(), &sun), "st_scene_update");
    soltrace::LaunchParams& p = data_manager->launch_params_H;
    p.sun_vector = mathUtil::toFloat3(s);
    p.max_sun_angle = (float)m_sun_angle;
    const st_sun_plane pl = get_sun_plane();
    p.sun_v0 = make_float3((float)pl.v0[0], (float)pl.v0[1], (float)pl.v0[2]);
    p.sun_v1 = make_float3((float)pl.v1[0], (float)pl.v1[1], (float)pl.v1[2]);
    p.sun_v2 = make_float3((float)pl.v2[0], (float)pl.v2[1], (float)pl.v2[2]);
    p.sun_v3 = make_float3((float)pl.v3[0], (float)pl.v3[1], (float)pl.v3[2]);
}

void SolTraceSystem::clean_up() {
    if (!m_ctx) return;
    soltrace::LaunchParams& p = data_manager->launch_params_H;
    if (p.hit_point_buffer) check(st_buffer_free(m_ctx, p.hit_point_buffer), "free");
    p.hit_point_buffer = nullptr;
    for (void** b : {(void**)&m_record_elements, (void**)&m_flux, (void**)&m_hits, (void**)&m_stats})
        if (*b) {
            check(st_buffer_free(m_ctx, *b), "free");
            *b = nullptr;
        }
    data_manager->cleanup();
    if (m_scene) check(st_scene_destroy(m_scene), "st_scene_destroy");
    m_scene = nullptr;
    check(st_context_destroy(m_ctx), "st_context_destroy");
    m_ctx = nullptr;
}

st_sun_plane SolTraceSystem::get_sun_plane() const {
    st_sun_plane pl;
    check(st_scene_get_sun_plane(m_scene, &pl), "st_scene_get_sun_plane");
    return pl;
}

std::vector<int64_t> SolTraceSystem::get_flux_map() const {
    std::vector<int64_t> out((size_t)m_flux_nx * m_flux_ny);
    if (m_flux) check(st_buffer_download(m_ctx, out.data(), m_flux, out.size() * sizeof(int64_t)), "download");
    return out;
}

double SolTraceSystem::get_power_per_ray() const {
    return m_dni * get_sun_plane().area / (double)m_num_sunpoints;
}

std::vector<double> SolTraceSystem::get_flux_map_watts() const {
    const std::vector<int64_t> c = get_flux_map();
    std::vector<double> out(c.size());
    if (c.empty()) return out;
    const Element& r = *m_element_list.back();
    const double w = r.get_aperture()->get_width(), h = r.get_aperture()->get_height();
    const double W = r.get_surface_type() == SurfaceType::CYLINDER ? M_PI * w : w;  // circumference 2 pi (w/2)
    const double bin_area = (W / m_flux_nx) * (h / m_flux_ny);
    const double ppr = get_power_per_ray();
    for (size_t i = 0; i < c.size(); ++i) out[i] = (double)c[i] * ppr / bin_area;
    return out;
}

std::vector<uint64_t> SolTraceSystem::get_element_hits() const {
    std::vector<uint64_t> out(m_element_list.size());
    if (m_hits) check(st_buffer_download(m_ctx, out.data(), m_hits, out.size() * sizeof(uint64_t)), "download");
    return out;
}

std::vector<uint64_t> SolTraceSystem::get_stats() const {
    std::vector<uint64_t> out(ST_STAT_COUNT);
    if (m_stats) check(st_buffer_download(m_ctx, out.data(), m_stats, out.size() * sizeof(uint64_t)), "download");
    return out;
}

void SolTraceSystem::print_launch_params() {
    const soltrace::LaunchParams& p = data_manager->launch_params_H;
    auto len = [](float3 a, float3 b) {
        const float x = a.x - b.x, y = a.y - b.y, z = a.z - b.z;
        return std::sqrt(x * x + y * y + z * z);
    };
    std::cout << "print launch params: " << std::endl;
    std::cout << "width: " << p.width << std::endl;
    std::cout << "height: " << p.height << std::endl;
    std::cout << "max_depth: " << p.max_depth << std::endl;
    std::cout << "hit_point_buffer: " << (void*)p.hit_point_buffer << std::endl;
    std::cout << "sun_vector: " << p.sun_vector.x << " " << p.sun_vector.y << " " << p.sun_vector.z << std::endl;
    std::cout << "max_sun_angle: " << p.max_sun_angle << std::endl;
    std::cout << "sun_v0: " << p.sun_v0.x << " " << p.sun_v0.y << " " << p.sun_v0.z << std::endl;
    std::cout << "sun_v1: " << p.sun_v1.x << " " << p.sun_v1.y << " " << p.sun_v1.z << std::endl;
    std::cout << "sun_v2: " << p.sun_v2.x << " " << p.sun_v2.y << " " << p.sun_v2.z << std::endl;
    std::cout << "sun_v3: " << p.sun_v3.x << " " << p.sun_v3.y << " " << p.sun_v3.z << std::endl;
    std::cout << "sun_box_edge_a: " << len(p.sun_v0, p.sun_v1) << std::endl;
    std::cout << "sun_box_edge_b: " << len(p.sun_v1, p.sun_v2) << std::endl;
}

// ---------------------------------------------------------------------------
// CSV dump, format of lib/soltrace_system.cpp:208-264: header
// "number,stage,loc_x,loc_y,loc_z"; slots walked in buffer order, an all-zero
// (x, y, z) slot closes the current ray, at most max_depth rows per ray.
void SolTraceSystem::write_csv(const std::string& filename, bool with_elements) {
    if (!m_record_hits || !data_manager->launch_params_H.hit_point_buffer) {
        std::cerr << "Error: hit records are off (set_record_hits(false)); nothing to write." << std::endl;
        return;
    }
    const int D = m_traced_depth ? m_traced_depth : m_max_depth;
    const size_t n = (size_t)m_num_sunpoints * (size_t)D;
    std::vector<float4> buf(n);
    std::vector<uint32_t> ids(with_elements ? n : 0);
    check(st_buffer_download(m_ctx, buf.data(), data_manager->launch_params_H.hit_point_buffer, n * sizeof(float4)),
          "download records");
    if (with_elements)
        check(st_buffer_download(m_ctx, ids.data(), m_record_elements, n * sizeof(uint32_t)), "download elements");
    std::ofstream out(filename);
    if (!out.is_open()) {
        std::cerr << "write_output: cannot create " << filename << std::endl;
        return;
    }
    out << "number,stage,loc_x,loc_y,loc_z" << (with_elements ? ",element" : "") << "\n";
    long long ray = 1;
    int stage = 0;
    for (size_t i = 0; i < n; ++i) {
        const float4& e = buf[i];
        if (e.y == 0 && e.z == 0 && e.w == 0) {
            if (stage > 0) {
                ++ray;
                stage = 0;
            }
            continue;
        }
        if (stage >= D) {
            ++ray;
            stage = 0;
        }
        out << ray << "," << e.x << "," << e.y << "," << e.z << "," << e.w;
        if (with_elements) out << "," << (ids[i] >= 0xFFFFFFFEu ? -1 : (long long)ids[i]);
        out << "\n";
        ++stage;
    }
    if (m_verbose) std::cout << "write_output: " << ray - (stage == 0 ? 1 : 0) << " rays written to " << filename << std::endl;
}

void SolTraceSystem::write_output(const std::string& filename) { write_csv(filename, false); }
void SolTraceSystem::write_output_with_elements(const std::string& filename) { write_csv(filename, true); }

// ---------------------------------------------------------------------------
// STINPUT: the file is parsed into a stinput::Document (stinput.cpp) and then
// turned into elements here, with the reference's reading of the format
// (lib/soltrace_system.cpp:449-846):
//  * SIGMA (mrad) becomes the sun's half-angle whatever SHAPE says (:513); the
//    shape itself is used only with set_use_stinput_optics(true);
//  * stage frames are taken as the global frame (element coordinates are global);
//  * a rectangle ('r') aperture gives width, height = parameters 0, 1; a
//    single-axis-curvature section ('l') with a cylindrical ('t') surface is a
//    cylinder of radius 1 / surface parameter 0 and length aperture parameter 2,
//    centred radius further along y than its row's origin; a circular ('c') flat
//    row is a cylinder end cap and is skipped;
//  * surfaces: 'f' flat, 'p' parabolic with curvatures = surface parameters 0, 1;
//  * the optic named in the row is kept per element; its front side's
//    reflectivity, RMS slope and RMS specularity (mrad) reach the trace with
//    set_use_stinput_optics(true) (the reference parses and drops them, :576-590).
bool SolTraceSystem::read_st_input(const char* filename) {
    stinput::Document doc;
    std::string err;
    if (!filename || !stinput::parse_file(filename, doc, err)) {
        std::cerr << "read_st_input: " << (filename ? filename : "(null)") << ": " << err << std::endl;
        return false;
    }
    return apply_stinput(doc);
}

bool SolTraceSystem::apply_stinput(const stinput::Document& doc) {
    std::vector<std::shared_ptr<Element>> elems;
    std::vector<std::string> optic_names;
    for (const stinput::Stage& st : doc.stages) {
        for (const stinput::ElementRow& r : st.elements) {
            const bool cylinder = r.aperture == 'l' && r.surface == 't';
            if (r.aperture == 'c' && r.surface == 'f') continue;  // cylinder end cap
            auto e = std::make_shared<Element>();
            Vector3d origin(r.origin[0], r.origin[1], r.origin[2]);
            if (cylinder) {
                if (!(r.surface_params[0] > 0)) {
                    std::cerr << "read_st_input: line " << r.line << ": cylinder curvature must be positive" << std::endl;
                    return false;
                }
                origin[1] += 1.0 / r.surface_params[0];
            }
            e->set_origin(origin);
            e->set_aim_point(Vector3d(r.aim[0], r.aim[1], r.aim[2]));
            e->set_zrot(r.zrot_deg);
            if (r.aperture == 'r') {
                e->set_aperture(std::make_shared<ApertureRectangle>(r.aperture_params[0], r.aperture_params[1]));
            } else if (cylinder) {
                e->set_aperture(std::make_shared<ApertureRectangle>(2.0 / r.surface_params[0], r.aperture_params[2]));
            } else {
                std::cerr << "read_st_input: line " << r.line << ": aperture code '" << r.aperture
                          << "' is not supported (rectangle 'r', or 'l' with a cylindrical surface)" << std::endl;
                return false;
            }
            if (r.surface == 'f') {
                e->set_surface(std::make_shared<SurfaceFlat>());
            } else if (r.surface == 'p') {
                auto s = std::make_shared<SurfaceParabolic>();
                s->set_curvature(r.surface_params[0], r.surface_params[1]);
                e->set_surface(s);
            } else if (cylinder) {
                auto s = std::make_shared<SurfaceCylinder>();
                s->set_radius(1.0 / r.surface_params[0]);
                s->set_half_height(0.5 * r.aperture_params[2]);
                e->set_surface(s);
            } else {
                std::cerr << "read_st_input: line " << r.line << ": surface code '" << r.surface
                          << "' is not supported (flat 'f', parabolic 'p', cylindrical 't')" << std::endl;
                return false;
            }
            elems.push_back(e);
            optic_names.push_back(r.optic);
        }
    }
    // the rows are all valid: only now does the system change
    set_sun_angle(doc.sun.sigma_mrad * 1e-3);
    if (m_use_stinput_optics) set_sun_shape(doc.sun.shape == 'g' ? ST_SUN_GAUSSIAN : ST_SUN_PILLBOX);
    set_sun_vector(Vector3d(doc.sun.xyz[0], doc.sun.xyz[1], doc.sun.xyz[2]));
    for (const stinput::OpticPair& p : doc.optics)
        m_optics.push_back({p.name, p.front.reflectivity, p.front.rms_slope_mrad, p.front.rms_specularity_mrad});
    for (size_t i = 0; i < elems.size(); ++i) {
        add_element(elems[i]);
        m_element_optic.back() = optic_names[i];
    }
    if (m_verbose)
        std::cout << "read_st_input: STINPUT " << doc.version[0] << "." << doc.version[1] << "." << doc.version[2]
                  << ", " << doc.stages.size() << " stages, " << elems.size() << " elements, " << doc.optics.size()
                  << " optics" << std::endl;
    return true;
}
