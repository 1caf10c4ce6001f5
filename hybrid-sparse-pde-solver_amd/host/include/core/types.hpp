#pragma once
// =============================================================================
// types.hpp — scalar aliases, enums, SolverParams / SolverStats of the solve
// path.  Same names, fields and defaults as the reference's
// include/core/types.hpp:16-143, plus the device-only controls at the end of
// SolverParams (precision, reliable_delta, check_every) and the FP64 true
// residual in SolverStats.
// =============================================================================
#include <cstddef>
#include <cstdint>
#include <limits>
#include <string>

namespace hsps {

using Real = double;
using Index = int;
using Size = std::size_t;

constexpr Real REAL_ZERO = Real(0);
constexpr Real REAL_ONE = Real(1);
constexpr Real REAL_EPS = std::numeric_limits<Real>::epsilon();
constexpr Real REAL_INF = std::numeric_limits<Real>::infinity();

enum class SolverType { CG, FGMRES, BICGSTAB, SSTEP_CG, ASYNC_FGMRES };
enum class PrecondType { NONE, JACOBI, ILU, AMG, BLOCK_JACOBI, CHEBYSHEV };
enum class AdaptiveState { EASY, MODERATE, HARD };

/// Inner-loop precision of the device path (no reference counterpart).
enum class Precision { FP64 = 0, MIXED = 1 };

inline const char* to_string(SolverType s) {
    switch (s) {
        case SolverType::CG: return "CG";
        case SolverType::FGMRES: return "FGMRES";
        case SolverType::BICGSTAB: return "BiCGSTAB";
        case SolverType::SSTEP_CG: return "s-step CG";
        case SolverType::ASYNC_FGMRES: return "Async-FGMRES";
    }
    return "Unknown";
}
inline const char* to_string(PrecondType p) {
    switch (p) {
        case PrecondType::NONE: return "None";
        case PrecondType::JACOBI: return "Jacobi";
        case PrecondType::ILU: return "ILU(0)/ILUT";
        case PrecondType::AMG: return "AMG";
        case PrecondType::BLOCK_JACOBI: return "Block-Jacobi";
        case PrecondType::CHEBYSHEV: return "Chebyshev";
    }
    return "Unknown";
}
inline const char* to_string(AdaptiveState st) {
    switch (st) {
        case AdaptiveState::EASY: return "EASY  [CG+Jacobi]";
        case AdaptiveState::MODERATE: return "MODERATE [FGMRES+ILU]";
        case AdaptiveState::HARD: return "HARD  [FGMRES+AMG]";
    }
    return "Unknown";
}

struct SolverStats {
    int iterations = 0;
    Real initial_residual = REAL_ZERO;
    Real final_residual = REAL_ZERO;
    bool converged = false;
    double solve_time_s = 0.0;
    double setup_time_s = 0.0;
    double energy_joules = 0.0;
    long long flop_count = 0;
    long long mem_bytes = 0;
    SolverType solver_used = SolverType::CG;
    PrecondType precond_used = PrecondType::JACOBI;
    AdaptiveState state_used = AdaptiveState::EASY;
    int restarts = 0;
    double energy_joules_hw = 0.0;
    double rapl_start_j = 0.0;
    double rapl_end_j = 0.0;
    int sstep_k = 0;
    bool async_mode = false;
    long long comm_bytes = 0;
    double comm_time_s = 0.0;
    // device path
    Real true_residual = REAL_ZERO;  ///< ||b - A x|| / ||b|| recomputed in FP64
    int reliable_updates = 0;        ///< mixed CG residual replacements
};

struct SolverParams {
    Real tol = 1e-8;
    int max_iter = 2000;
    int restart_size = 50;
    Real drop_tol = 1e-4;
    int amg_levels = 10;
    Real amg_strength = 0.25;
    Real amg_strong = 0.25;
    int amg_smooth_pre = 2;
    int amg_smooth_post = 2;
    int stall_window = 15;
    Real stall_threshold = 0.95;
    bool verbose = false;
    int print_every = 25;
    int sstep_k = 1;
    Real sstep_cond_tol = 1e12;
    bool async_enabled = false;
    int async_max_stale = 3;
    int fill_per_row = 0;
    int block_size = 1;
    bool record_hw_energy = false;
    // device path
    Precision precision = Precision::FP64;
    Real reliable_delta = 0.1;
    int check_every = 32;
    bool device_graphs = true;  ///< replay fused iteration chunks as hipGraphs (off when threads share the device)
    Real fgmres_inner_rel = 0.0;  ///< mixed FGMRES: FP32 cycle ends at this residual reduction (0: at tol)
};

}  // namespace hsps
