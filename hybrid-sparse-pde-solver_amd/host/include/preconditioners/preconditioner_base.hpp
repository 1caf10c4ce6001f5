#pragma once
// =============================================================================
// preconditioner_base.hpp — PreconditionerBase with the reference's API
// (preconditioner_base.hpp:16-48).  Concrete preconditioners own an
// hsps_precond device object; apply() runs on the device.
// =============================================================================
#include "../core/sparse_matrix.hpp"
#include "../core/types.hpp"
#include "../core/vector.hpp"

namespace hsps {

class PreconditionerBase {
public:
    explicit PreconditionerBase(PrecondType type) : type_(type) {}
    virtual ~PreconditionerBase();

    /// Build from A; returns setup wall time in seconds.
    virtual double setup(const SparseMatrix& A) = 0;
    /// z = M^-1 r (z resized; device apply).
    virtual void apply(const Vector& r, Vector& z) const;

    PrecondType type() const { return type_; }
    bool is_ready() const { return ready_; }
    long long setup_flops() const { return setup_flops_; }
    virtual const char* energy_tier() const = 0;

    /// The device object the Krylov solvers call (null before setup()).
    hsps_precond device() const { return dev_; }

protected:
    /// Shared setup path: builds the device object for `kind` with params p.
    double setup_device(const SparseMatrix& A, int kind, const hsps_params& p);

    PrecondType type_;
    bool ready_ = false;
    long long setup_flops_ = 0;
    hsps_precond dev_ = nullptr;
    SparseMatrix A_;  ///< the operator (value copy, as the reference's AMG keeps; amg.cpp:190)
};

}  // namespace hsps
