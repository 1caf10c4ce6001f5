// =============================================================================
// energy.cpp — EnergyMonitor (energy_monitor.cpp:12-164) and the GPU power
// monitor with the NVMLMonitor API (nvml_monitor.cpp:35-193), both reading the
// device's ROCm SMI energy counter through the C ABI.
// =============================================================================
#include <algorithm>
#include <cmath>
#include <fstream>
#include <iomanip>
#include <iostream>

#include "core/device.hpp"
#include "energy/energy_monitor.hpp"
#include "parallel/nvml_monitor.hpp"

namespace hsps {

static double wall_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// ---- EnergyMonitor ---------------------------------------------------------------------
EnergyMonitor::EnergyMonitor(const EnergyModelCoeffs& coeffs, int device) : coeffs_(coeffs) {
    if (hsps_energy_open(device, &hw_) != HSPS_OK) hw_ = nullptr;
}
EnergyMonitor::~EnergyMonitor() {
    if (hw_) hsps_energy_close(hw_);
}
double EnergyMonitor::read_hw_j() const {
    double j = 0.0;
    if (!hw_ || hsps_energy_read(hw_, &j, nullptr, nullptr) != HSPS_OK) return 0.0;
    return j;
}
void EnergyMonitor::start() {  // energy_monitor.cpp:59-67
    start_time_ = Clock::now();
    running_ = true;
    total_flops_ = 0;
    total_mem_bytes_ = 0;
    history_.clear();
    if (hw_) hw_start_j_ = read_hw_j();
}
void EnergyMonitor::stop() { running_ = false; }
void EnergyMonitor::reset() {
    running_ = false;
    total_flops_ = 0;
    total_mem_bytes_ = 0;
    history_.clear();
    hw_start_j_ = 0.0;
}
void EnergyMonitor::record(int iteration, double residual, long long flops, long long mem_bytes) {
    total_flops_ += flops;
    total_mem_bytes_ += mem_bytes;
    const double e = coeffs_.alpha_flop * (double)flops + coeffs_.alpha_mem * (double)mem_bytes;
    history_.push_back({iteration, residual, flops, mem_bytes, e, elapsed_seconds()});
}
void EnergyMonitor::record_with_hw(int iteration, double residual, long long flops_spmv, long long bytes_spmv,
                                   long long flops_blas1, long long bytes_blas1, double hw_start_j,
                                   double hw_end_j) {
    total_flops_ += flops_spmv + flops_blas1;
    total_mem_bytes_ += bytes_spmv + bytes_blas1;
    IterRecord r{};
    r.iteration = iteration;
    r.residual = residual;
    r.flops = flops_spmv + flops_blas1;
    r.mem_bytes = bytes_spmv + bytes_blas1;
    r.energy_spmv_j = coeffs_.alpha_flop_spmv * (double)flops_spmv + coeffs_.alpha_mem_dram * (double)bytes_spmv;
    r.energy_blas1_j = coeffs_.alpha_flop_blas1 * (double)flops_blas1 + coeffs_.alpha_mem_l3 * (double)bytes_blas1;
    r.energy_proxy_j = r.energy_spmv_j + r.energy_blas1_j;
    r.elapsed_s = elapsed_seconds();
    r.flops_spmv = flops_spmv;
    r.flops_blas1 = flops_blas1;
    r.bytes_spmv = bytes_spmv;
    r.bytes_blas1 = bytes_blas1;
    r.rapl_j_delta = hw_end_j > hw_start_j ? hw_end_j - hw_start_j : 0.0;
    history_.push_back(r);
}
double EnergyMonitor::snapshot_rapl() const { return hw_ ? read_hw_j() : 0.0; }
double EnergyMonitor::elapsed_seconds() const {
    return std::chrono::duration<double>(Clock::now() - start_time_).count();
}
double EnergyMonitor::total_energy_proxy_j() const {
    return coeffs_.alpha_flop * (double)total_flops_ + coeffs_.alpha_mem * (double)total_mem_bytes_;
}
double EnergyMonitor::total_energy_j() const {
    if (hw_) return std::max(0.0, read_hw_j() - hw_start_j_);  // the GPU counter is 64-bit: no wrap
    return total_energy_proxy_j();
}
EnergySample EnergyMonitor::sample() const {
    EnergySample s;
    s.elapsed_s = elapsed_seconds();
    s.flops = total_flops_;
    s.mem_bytes = total_mem_bytes_;
    s.hw_available = hw_ != nullptr;
    s.energy_joules = total_energy_j();
    s.avg_watts = s.elapsed_s > 1e-9 ? s.energy_joules / s.elapsed_s : 0.0;
    return s;
}
void EnergyMonitor::print_summary(std::ostream& os) const {
    const EnergySample s = sample();
    os << "\n=== Energy Monitor Summary ===\n"
       << "  Elapsed time    : " << std::fixed << std::setprecision(4) << s.elapsed_s << " s\n"
       << "  Total FLOPs     : " << s.flops << "\n"
       << "  Memory traffic  : " << s.mem_bytes / (1024 * 1024) << " MB\n"
       << "  Energy (proxy)  : " << std::setprecision(6) << total_energy_proxy_j() << " J\n";
    if (hw_) os << "  Energy (GPU)    : " << s.energy_joules << " J\n";
    os << "  Avg power       : " << std::setprecision(3) << s.avg_watts << " W\n"
       << "  HW energy avail : " << (hw_ ? "YES (GPU energy counter)" : "NO (proxy used)") << "\n"
       << "==============================\n";
}
void EnergyMonitor::dump_csv(const std::string& path) const {
    std::ofstream f(path);
    if (!f.good()) {
        std::cerr << "[EnergyMonitor] Cannot open " << path << "\n";
        return;
    }
    f << "iteration,residual,flops,mem_bytes,energy_proxy_j,elapsed_s\n";
    for (const auto& r : history_)
        f << r.iteration << "," << std::scientific << std::setprecision(8) << r.residual << "," << r.flops << ","
          << r.mem_bytes << "," << r.energy_proxy_j << "," << r.elapsed_s << "\n";
}
EnergyModelCoeffs EnergyMonitor::calibrate(int device) {
    EnergyModelCoeffs c;
    hsps_energy E = nullptr;
    if (hsps_energy_open(device, &E) != HSPS_OK) return c;
    hsps_ctx ctx = DeviceContext::current();
    const int n = 8 * 1024 * 1024;  // three 64 MiB vectors: beyond the 256 MB cache in one triad pass
    double *x = nullptr, *y = nullptr;
    dev_check(hsps_malloc(ctx, (size_t)n * 8, (void**)&x));
    dev_check(hsps_malloc(ctx, (size_t)n * 8, (void**)&y));
    dev_check(hsps_memset(ctx, x, 0, (size_t)n * 8));
    dev_check(hsps_memset(ctx, y, 0, (size_t)n * 8));
    const int reps = 400;  // ~20 GB of traffic: long enough for the counter's resolution
    dev_check(hsps_axpy_f64(ctx, n, 1.0, x, y));
    dev_check(hsps_ctx_sync(ctx));
    double j0 = 0.0, j1 = 0.0;
    hsps_energy_read(E, &j0, nullptr, nullptr);
    const double t0 = wall_s();
    for (int r = 0; r < reps; ++r) dev_check(hsps_axpy_f64(ctx, n, 1.0, x, y));
    dev_check(hsps_ctx_sync(ctx));
    const double dt = wall_s() - t0;
    hsps_energy_read(E, &j1, nullptr, nullptr);
    hsps_free(ctx, x);
    hsps_free(ctx, y);
    hsps_energy_close(E);
    const double bytes = (double)reps * 3.0 * n * sizeof(double);
    const double e = j1 - j0;
    if (e > 0 && dt > 0) {
        c.alpha_mem = e / bytes;
        c.alpha_mem_dram = c.alpha_mem;
        c.alpha_mem_l3 = c.alpha_mem * 0.1;
        c.alpha_flop = c.alpha_mem * (3.0 * sizeof(double)) / 2.0;  // the triad's energy per FLOP
        c.alpha_flop_spmv = c.alpha_flop * 1.3;
        c.alpha_flop_blas1 = c.alpha_flop;
        c.idle_watts = 0.0;
        c.calibrated = true;
        std::cout << "[EnergyMonitor::calibrate()] device triad: " << bytes / dt / 1e9 << " GB/s, " << e / dt
                  << " W, alpha_mem = " << c.alpha_mem << " J/B\n";
    }
    return c;
}

// ---- GPU power monitor (NVMLMonitor API) ----------------------------------------------
GpuPowerMonitor::GpuPowerMonitor(int device_index) : device_index_(device_index) {
    if (hsps_energy_open(device_index, &handle_) != HSPS_OK) {
        handle_ = nullptr;
        return;
    }
    char name[128] = {0};
    hsps_energy_info(handle_, nullptr, nullptr, nullptr, nullptr, name, (int)sizeof name);
    device_name_ = name;
}
GpuPowerMonitor::~GpuPowerMonitor() {
    if (handle_) hsps_energy_close(handle_);
}
double GpuPowerMonitor::read_energy_j() const {
    double j = 0.0;
    if (!handle_ || hsps_energy_read(handle_, &j, nullptr, nullptr) != HSPS_OK) return 0.0;
    return j;
}
bool GpuPowerMonitor::start() {
    if (!handle_) return false;
    start_time_s_ = wall_s();
    start_energy_j_ = read_energy_j();
    history_.clear();
    return true;
}
void GpuPowerMonitor::stop() { take_sample(); }
void GpuPowerMonitor::reset() {
    history_.clear();
    start_energy_j_ = read_energy_j();
}
double GpuPowerMonitor::current_power_w() const {
    double w = 0.0;
    if (!handle_ || hsps_energy_read(handle_, nullptr, &w, nullptr) != HSPS_OK || w < 0) return 0.0;
    return w;
}
double GpuPowerMonitor::elapsed_energy_j() const { return std::max(0.0, read_energy_j() - start_energy_j_); }
static double info_field(hsps_energy h, int which) {
    if (!h) return 0.0;
    double v[4] = {-1, -1, -1, -1};
    hsps_energy_info(h, &v[0], &v[1], &v[2], &v[3], nullptr, 0);
    return v[which] < 0 ? 0.0 : v[which];
}
double GpuPowerMonitor::power_limit_w() const { return info_field(handle_, 0); }
double GpuPowerMonitor::temperature_c() const { return info_field(handle_, 1); }
double GpuPowerMonitor::gpu_utilisation_pct() const { return info_field(handle_, 2); }
double GpuPowerMonitor::memory_used_mb() const { return info_field(handle_, 3); }
double GpuPowerMonitor::power_headroom_w() const { return power_limit_w() - current_power_w(); }
void GpuPowerMonitor::take_sample() {
    GPUPowerSample s;
    s.elapsed_s = wall_s() - start_time_s_;
    s.power_w = current_power_w();
    s.energy_j_cumulative = elapsed_energy_j();
    history_.push_back(s);
}
void GpuPowerMonitor::dump_csv(const std::string& path) const {
    std::ofstream f(path);
    if (!f.good()) return;
    f << "elapsed_s,power_w,energy_j_cumulative\n";
    for (const auto& s : history_)
        f << std::fixed << std::setprecision(6) << s.elapsed_s << "," << std::setprecision(3) << s.power_w << ","
          << std::setprecision(6) << s.energy_j_cumulative << "\n";
}

}  // namespace hsps
