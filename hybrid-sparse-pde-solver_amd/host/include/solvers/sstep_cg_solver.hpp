#pragma once
// sstep_cg_solver.hpp — SStepCGSolver (reference sstep_cg_solver.hpp): s-step
// CG on the device.  Per batch of k steps one fused pass forms the Gram
// matrix of the normalised Newton-Chebyshev basis and B^T r, the host solves
// the k x k system, and x += B y.  As in the reference, a preconditioner set
// on it is not applied; FP64 only.
#include "solver_base.hpp"

namespace hsps {

class SStepCGSolver : public SolverBase {
public:
    explicit SStepCGSolver(int k = 4) : SolverBase(SolverType::SSTEP_CG), k_(k) {}
    bool solve(const SparseMatrix& A, const Vector& b, Vector& x, SolverStats& stats) override;

    /// Step size k (batch size); a batch halves it when the basis is ill-conditioned.
    void set_k(int k) { k_ = k; }
    int k() const { return k_; }

private:
    int k_;
};

}  // namespace hsps
