#pragma once
// ilu.hpp — ILUPreconditioner (reference ilu.hpp:19-54): ILU(0) when
// drop_tol == 0 and fill == 0, ILUT(tau, p) otherwise; device triangular solves.
#include "preconditioner_base.hpp"

namespace hsps {

class ILUPreconditioner : public PreconditionerBase {
public:
    ILUPreconditioner() : PreconditionerBase(PrecondType::ILU) {}
    double setup(const SparseMatrix& A) override;
    const char* energy_tier() const override { return "MODERATE"; }
    void set_drop_tol(Real tau) { drop_tol_ = tau; }
    void set_fill_per_row(int p) { fill_p_ = p; }

private:
    Real drop_tol_ = 0.0;
    int fill_p_ = 0;
};

}  // namespace hsps
