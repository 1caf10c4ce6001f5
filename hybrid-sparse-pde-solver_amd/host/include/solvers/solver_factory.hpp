#pragma once
// solver_factory.hpp — SolverFactory with the reference's API
// (solver_factory.hpp:34-84).  Solver types without a device implementation
// (s-step CG, async FGMRES) map to the reference's own fallbacks.
#include <memory>
#include <string>
#include <utility>

#include "../core/sparse_matrix.hpp"
#include "../core/types.hpp"
#include "../preconditioners/preconditioner_base.hpp"
#include "solver_base.hpp"

namespace hsps {

using SolverPrecondPair = std::pair<std::unique_ptr<SolverBase>, std::shared_ptr<PreconditionerBase>>;

class SolverFactory {
public:
    static SolverPrecondPair make(SolverType solver_type, PrecondType precond_type, const SparseMatrix& A,
                                  const SolverParams& params, double* setup_time_s = nullptr);
    static SolverPrecondPair make_for_state(AdaptiveState state, const SparseMatrix& A, const SolverParams& params,
                                            double* setup_time_s = nullptr);
    static SolverPrecondPair make(const std::string& solver_name, const std::string& precond_name,
                                  const SparseMatrix& A, const SolverParams& params, double* setup_time_s = nullptr);
    static SolverType parse_solver(const std::string& s);
    static PrecondType parse_precond(const std::string& s);
    static std::shared_ptr<PreconditionerBase> make_precond(PrecondType type, const SparseMatrix& A,
                                                            const SolverParams& params,
                                                            double* setup_time_s = nullptr);
    static std::unique_ptr<SolverBase> make_solver(SolverType type, const SolverParams& params);

private:
    SolverFactory() = delete;
};

}  // namespace hsps
