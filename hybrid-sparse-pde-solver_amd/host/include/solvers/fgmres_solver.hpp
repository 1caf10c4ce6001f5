#pragma once
// fgmres_solver.hpp — FGMRESSolver (reference fgmres_solver.hpp): device
// right-preconditioned flexible GMRES(m).
#include "solver_base.hpp"

namespace hsps {

class FGMRESSolver : public SolverBase {
public:
    FGMRESSolver() : SolverBase(SolverType::FGMRES) {}
    bool solve(const SparseMatrix& A, const Vector& b, Vector& x, SolverStats& stats) override;
};

}  // namespace hsps
