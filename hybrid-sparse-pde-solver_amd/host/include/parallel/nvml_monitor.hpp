#pragma once
// =============================================================================
// nvml_monitor.hpp — GPU power / energy monitor with the reference's
// NVMLMonitor API (nvml_monitor.hpp, nvml_monitor.cpp:35-193), read from the
// MI355X's ROCm SMI counters through the C ABI (hsps_energy_*).  available()
// is false where the SMI library or the device's energy counter is missing;
// readings are then 0, as the reference's are without NVML.
// =============================================================================
#include <memory>
#include <string>
#include <vector>

#include "hsps_dev.h"

namespace hsps {

struct GPUPowerSample {
    double elapsed_s;
    double power_w;
    double energy_j_cumulative;
};

class GpuPowerMonitor {
public:
    explicit GpuPowerMonitor(int device_index = 0);
    ~GpuPowerMonitor();
    GpuPowerMonitor(const GpuPowerMonitor&) = delete;
    GpuPowerMonitor& operator=(const GpuPowerMonitor&) = delete;

    bool start();  ///< Begins a measurement; false when telemetry is unavailable.
    void stop();   ///< Takes a final sample.
    void reset();

    double current_power_w() const;
    double elapsed_energy_j() const;  ///< Energy since start(), joules
    double gpu_utilisation_pct() const;
    double memory_used_mb() const;
    double temperature_c() const;
    double power_limit_w() const;
    double power_headroom_w() const;
    bool available() const { return handle_ != nullptr; }
    int device_index() const { return device_index_; }
    std::string device_name() const { return device_name_; }

    void take_sample();
    const std::vector<GPUPowerSample>& history() const { return history_; }
    void dump_csv(const std::string& path) const;

    /// Cumulative device energy counter, joules (0 when unavailable).
    double read_energy_j() const;

private:
    int device_index_;
    hsps_energy handle_ = nullptr;
    std::string device_name_;
    double start_energy_j_ = 0.0;
    double start_time_s_ = 0.0;
    std::vector<GPUPowerSample> history_;
};

/// The reference's class name, for callers that use it directly.
using NVMLMonitor = GpuPowerMonitor;

}  // namespace hsps
