#pragma once
// =============================================================================
// energy_monitor.hpp — EnergyMonitor with the reference's API
// (energy_monitor.hpp, energy_monitor.cpp:27-164): FLOP / byte proxy
// accounting per iteration, and a hardware energy source, which on MI355X is
// the GPU's own energy counter (ROCm SMI, through hsps_energy_*) in place of
// the CPU package's RAPL counter.  calibrate() measures J/byte with a device
// triad instead of a host STREAM loop.  Plus the flop_model byte/FLOP
// counters the solvers use to fill SolverStats (energy_monitor.hpp:163-177).
// =============================================================================
#include <chrono>
#include <iosfwd>
#include <string>
#include <vector>

#include "../core/types.hpp"
#include "hsps_dev.h"

namespace hsps {

struct EnergySample {
    double elapsed_s = 0.0;
    double energy_joules = 0.0;  ///< hardware (GPU energy counter) or proxy
    double avg_watts = 0.0;
    long long flops = 0;
    long long mem_bytes = 0;
    bool hw_available = false;
};

struct EnergyModelCoeffs {
    double alpha_flop = 2.0e-10;
    double alpha_flop_spmv = 2.0e-10;
    double alpha_flop_blas1 = 1.5e-10;
    double alpha_mem = 5.0e-9;
    double alpha_mem_l3 = 5.0e-10;
    double alpha_mem_dram = 5.0e-9;
    double alpha_comm = 1.0e-8;
    double idle_watts = 50.0;
    bool calibrated = false;
};

struct IterRecord {
    int iteration;
    double residual;
    long long flops;
    long long mem_bytes;
    double energy_proxy_j;
    double elapsed_s;
    long long flops_spmv = 0;
    long long flops_blas1 = 0;
    long long bytes_spmv = 0;
    long long bytes_blas1 = 0;
    double energy_spmv_j = 0.0;
    double energy_blas1_j = 0.0;
    double energy_precond_j = 0.0;
    double rapl_j_delta = 0.0;  ///< hardware energy for this iteration (GPU counter)
};

class EnergyMonitor {
public:
    explicit EnergyMonitor(const EnergyModelCoeffs& coeffs = {}, int device = 0);
    ~EnergyMonitor();
    EnergyMonitor(const EnergyMonitor&) = delete;
    EnergyMonitor& operator=(const EnergyMonitor&) = delete;

    void start();
    void stop();
    void reset();
    void record(int iteration, double residual, long long flops, long long mem_bytes);
    void record_with_hw(int iteration, double residual, long long flops_spmv, long long bytes_spmv,
                        long long flops_blas1, long long bytes_blas1, double hw_start_j, double hw_end_j);
    /// Current hardware energy counter, joules (0 when unavailable).
    double snapshot_rapl() const;
    EnergySample sample() const;
    double elapsed_seconds() const;
    double total_energy_j() const;  ///< hardware if available, else proxy
    double total_energy_proxy_j() const;
    long long total_flops() const { return total_flops_; }
    long long total_mem_bytes() const { return total_mem_bytes_; }
    bool hw_energy_available() const { return hw_ != nullptr; }
    const std::vector<IterRecord>& history() const { return history_; }
    void print_summary(std::ostream& os) const;
    void dump_csv(const std::string& path) const;
    /// J/byte and J/FLOP from a device triad (y += a x over 64 MiB) timed
    /// against the GPU energy counter; defaults when no counter is present.
    static EnergyModelCoeffs calibrate(int device = 0);

private:
    double read_hw_j() const;
    using Clock = std::chrono::steady_clock;
    EnergyModelCoeffs coeffs_;
    hsps_energy hw_ = nullptr;
    double hw_start_j_ = 0.0;
    Clock::time_point start_time_ = Clock::now();
    bool running_ = false;
    long long total_flops_ = 0;
    long long total_mem_bytes_ = 0;
    std::vector<IterRecord> history_;
};

namespace flop_model {
inline long long spmv_flops(int nnz) { return 2LL * nnz; }
inline long long spmv_bytes(int rows, int nnz) {
    return (long long)(rows + 1 + nnz) * sizeof(Index) + (long long)nnz * sizeof(Real) +
           (long long)rows * sizeof(Real) * 2;
}
inline long long dot_flops(int n) { return 2LL * n; }
inline long long dot_bytes(int n) { return 2LL * n * sizeof(Real); }
inline long long axpy_flops(int n) { return 2LL * n; }
inline long long axpy_bytes(int n) { return 3LL * n * sizeof(Real); }
}  // namespace flop_model

}  // namespace hsps
