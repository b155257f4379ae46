// data_manager.h - host/device launch data of the drop-in API (lib/data_manager.h).
#pragma once
#ifndef DATAMANAGER_H
#define DATAMANAGER_H

#include <vector>

#include "Soltrace.h"
#include "element.h"

class dataManager {
public:
    soltrace::LaunchParams launch_params_H;
    soltrace::LaunchParams* launch_params_D = nullptr;
    GeometryDataST* geometry_data_array_D = nullptr;

    dataManager();
    ~dataManager();

    // device memory comes from this context (set by SolTraceSystem::initialize)
    void set_context(st_context ctx) { ctx_ = ctx; }

    void cleanup();
    soltrace::LaunchParams* getDeviceLaunchParams() const { return launch_params_D; }
    void allocateLaunchParams();
    void updateLaunchParams();
    void allocateGeometryDataArray(std::vector<GeometryDataST> geometry_data_array);
    void updateGeometryDataArray(std::vector<GeometryDataST> geometry_data_array_H);

private:
    st_context ctx_ = nullptr;
    size_t geometry_count_ = 0;
};

#endif
