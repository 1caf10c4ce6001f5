#pragma once
// pipelined_fgmres_solver.hpp — PipelinedFGMRESSolver (reference
// pipelined_fgmres_solver.hpp): the FGMRES Krylov space with each iteration's
// orthogonalisation done as two fused classical Gram-Schmidt passes on the
// device (a fixed number of launches and reductions per iteration); a
// breakdown ends the cycle, and (its + 1) SpMVs per cycle are charged.
#include "solver_base.hpp"

namespace hsps {

class PipelinedFGMRESSolver : public SolverBase {
public:
    PipelinedFGMRESSolver() : SolverBase(SolverType::FGMRES) {}
    bool solve(const SparseMatrix& A, const Vector& b, Vector& x, SolverStats& stats) override;
};

}  // namespace hsps
