// data_manager.cpp (lib/data_manager.cpp): H2D copies of the launch data.
#include "data_manager.h"

#include <stdexcept>
#include <string>

namespace {
void check(int rc, const char* what) {
    if (rc != ST_OK) throw std::runtime_error(std::string(what) + ": " + st_last_error());
}
}  // namespace

dataManager::dataManager() {
    launch_params_H.width = 10;
    launch_params_H.height = 1;
    launch_params_H.max_depth = 5;
}

dataManager::~dataManager() {
    try {
        cleanup();
    } catch (...) {
    }
}

void dataManager::allocateLaunchParams() {
    if (!ctx_) throw std::runtime_error("dataManager: no device context");
    if (!launch_params_D)
        check(st_buffer_alloc(ctx_, sizeof(soltrace::LaunchParams), (void**)&launch_params_D), "allocateLaunchParams");
}

void dataManager::updateLaunchParams() {
    if (!launch_params_D) throw std::runtime_error("Launch params are not allocated.");
    check(st_buffer_upload(ctx_, launch_params_D, &launch_params_H, sizeof(launch_params_H)), "updateLaunchParams");
}

void dataManager::allocateGeometryDataArray(std::vector<GeometryDataST> a) {
    if (!ctx_) throw std::runtime_error("dataManager: no device context");
    if (geometry_data_array_D) check(st_buffer_free(ctx_, geometry_data_array_D), "free geometry");
    geometry_data_array_D = nullptr;
    check(st_buffer_alloc(ctx_, a.size() * sizeof(GeometryDataST), (void**)&geometry_data_array_D),
          "allocateGeometryDataArray");
    geometry_count_ = a.size();
    check(st_buffer_upload(ctx_, geometry_data_array_D, a.data(), a.size() * sizeof(GeometryDataST)),
          "allocateGeometryDataArray");
    launch_params_H.geometry_data_array = geometry_data_array_D;
}

void dataManager::updateGeometryDataArray(std::vector<GeometryDataST> a) {
    if (!geometry_data_array_D) throw std::runtime_error("Geometry data array is not allocated.");
    if (a.size() != geometry_count_) throw std::runtime_error("Geometry data array changed size.");
    check(st_buffer_upload(ctx_, geometry_data_array_D, a.data(), a.size() * sizeof(GeometryDataST)),
          "updateGeometryDataArray");
}

void dataManager::cleanup() {
    if (!ctx_) return;
    if (launch_params_D) check(st_buffer_free(ctx_, launch_params_D), "free launch params");
    launch_params_D = nullptr;
    if (geometry_data_array_D) check(st_buffer_free(ctx_, geometry_data_array_D), "free geometry");
    geometry_data_array_D = nullptr;
}
