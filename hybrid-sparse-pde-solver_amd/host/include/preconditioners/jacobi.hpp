#pragma once
// jacobi.hpp — JacobiPreconditioner (reference jacobi.hpp:19-34): z = omega D^-1 r.
#include "preconditioner_base.hpp"

namespace hsps {

class JacobiPreconditioner : public PreconditionerBase {
public:
    JacobiPreconditioner() : PreconditionerBase(PrecondType::JACOBI) {}
    double setup(const SparseMatrix& A) override;
    const char* energy_tier() const override { return "CHEAP"; }
    void set_omega(Real omega) { omega_ = omega; }

private:
    Real omega_ = 1.0;
};

}  // namespace hsps
