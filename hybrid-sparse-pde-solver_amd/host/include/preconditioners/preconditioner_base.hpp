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

    /// For subclasses written by callers (no reference counterpart): z = M^-1 r
    /// on device vectors of n FP64 entries, the work enqueued on `hip_stream`.
    /// Return false (the default) to have the solvers call the host apply()
    /// instead, with r and z copied between host and device around it.
    virtual bool device_apply(const double* d_r, double* d_z, Index n, void* hip_stream) const {
        (void)d_r; (void)d_z; (void)n; (void)hip_stream;
        return false;
    }

    /// The object the Krylov solvers call for an n x n system: device() for
    /// the built-in preconditioners; for a subclass without one, a
    /// caller-written preconditioner (hsps_precond_create_user) that calls
    /// device_apply() or apply() each iteration (cg_solver.cpp:79, 147;
    /// fgmres_solver.cpp:112).
    hsps_precond solver_object(Index n) const;

protected:
    /// Shared setup path: builds the device object for `kind` with params p.
    double setup_device(const SparseMatrix& A, int kind, const hsps_params& p);

    PrecondType type_;
    bool ready_ = false;
    long long setup_flops_ = 0;
    hsps_precond dev_ = nullptr;
    SparseMatrix A_;  ///< the operator (value copy, as the reference's AMG keeps; amg.cpp:190)

private:
    // the caller-written object of solver_object(): its size and context
    mutable hsps_precond user_dev_ = nullptr;
    mutable Index user_n_ = -1;
    mutable hsps_ctx user_ctx_ = nullptr;
    static int user_apply(void* self, const double* d_r, double* d_z, void* hip_stream);
};

}  // namespace hsps
