#pragma once
// =============================================================================
// solver_base.hpp — SolverBase with the reference's API (solver_base.hpp:20-62).
// solve() runs on the device; x is the initial guess when x.size()==b.size().
// =============================================================================
#include <memory>

#include "../core/sparse_matrix.hpp"
#include "../core/types.hpp"
#include "../core/vector.hpp"
#include "../preconditioners/preconditioner_base.hpp"

namespace hsps {

class EnergyMonitor;

class SolverBase {
public:
    explicit SolverBase(SolverType type) : type_(type) {}
    virtual ~SolverBase() = default;

    virtual bool solve(const SparseMatrix& A, const Vector& b, Vector& x, SolverStats& stats) = 0;

    void set_params(const SolverParams& p) { params_ = p; }
    void set_preconditioner(std::shared_ptr<PreconditionerBase> M) { precond_ = M; }
    void set_verbose(bool v) { params_.verbose = v; }
    void set_energy_monitor(std::shared_ptr<EnergyMonitor> mon) { energy_mon_ = mon; }

    const SolverParams& params() const { return params_; }
    SolverType type() const { return type_; }

protected:
    SolverType type_;
    SolverParams params_;
    std::shared_ptr<PreconditionerBase> precond_ = nullptr;
    std::shared_ptr<EnergyMonitor> energy_mon_ = nullptr;

    Real relative_residual(const SparseMatrix& A, const Vector& b, const Vector& x) const;
    /// Shared device call for the solvers below (C ABI entry per kind).
    enum DeviceSolver { kCG = 0, kFGMRES = 1, kPipelinedFGMRES = 2, kSStepCG = 3 };
    int sstep_k_ = 1;
    bool device_solve(int solver, const SparseMatrix& A, const Vector& b, Vector& x, SolverStats& stats);
};

hsps_params to_device_params(const SolverParams& p);

}  // namespace hsps
