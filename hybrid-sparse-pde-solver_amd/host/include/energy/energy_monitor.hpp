#pragma once
// energy_monitor.hpp — the flop_model byte/FLOP counters the solvers use to
// fill SolverStats (reference energy_monitor.hpp:163-177).  The RAPL reader and
// per-iteration history of the reference are not part of the device path.
#include "../core/types.hpp"

namespace hsps {

class EnergyMonitor {
public:
    virtual ~EnergyMonitor() = default;
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
