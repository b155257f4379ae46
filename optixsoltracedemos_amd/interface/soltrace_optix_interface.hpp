// soltrace_optix_interface.hpp - pimpl embedding wrapper
// (interface/soltrace_optix_interface.hpp:3-24). set_sun_angle takes radians,
// as the reference passes the value straight to SolTraceSystem::set_sun_angle.
#pragma once
#include <memory>

class soltrace_optix_interface {
public:
    explicit soltrace_optix_interface(int num_rays = 10000);
    ~soltrace_optix_interface();

    void set_sun_vector(double x, double y, double z);
    void set_sun_angle(double deg);
    void set_sun_points(int n);

    void read_st_input(const char* stinputPath);
    void initialize();
    void run();
    void write_output(const char* outPath);

private:
    struct Impl;
    std::unique_ptr<Impl> impl;
};
