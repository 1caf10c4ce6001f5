#pragma once
// cg_solver.hpp — CGSolver (reference cg_solver.hpp): device PCG.
#include "solver_base.hpp"

namespace hsps {

class CGSolver : public SolverBase {
public:
    CGSolver() : SolverBase(SolverType::CG) {}
    bool solve(const SparseMatrix& A, const Vector& b, Vector& x, SolverStats& stats) override;
};

}  // namespace hsps
