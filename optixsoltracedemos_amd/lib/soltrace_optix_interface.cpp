// soltrace_optix_interface.cpp (interface/soltrace_optix_interface.cpp:5-51)
#include "../interface/soltrace_optix_interface.hpp"

#include "soltrace_system.h"

struct soltrace_optix_interface::Impl {
    SolTraceSystem sys;
    explicit Impl(int n) : sys(n) {}
};

soltrace_optix_interface::soltrace_optix_interface(int num_rays) : impl(std::make_unique<Impl>(num_rays)) {}
soltrace_optix_interface::~soltrace_optix_interface() = default;
void soltrace_optix_interface::set_sun_vector(double x, double y, double z) { impl->sys.set_sun_vector(Vector3d(x, y, z)); }
void soltrace_optix_interface::set_sun_angle(double deg) { impl->sys.set_sun_angle(deg); }
void soltrace_optix_interface::set_sun_points(int n) { impl->sys.set_sun_points(n); }
void soltrace_optix_interface::read_st_input(const char* p) { impl->sys.read_st_input(p); }
void soltrace_optix_interface::initialize() { impl->sys.initialize(); }
void soltrace_optix_interface::run() { impl->sys.run(); }
void soltrace_optix_interface::write_output(const char* p) { impl->sys.write_output(p); }
