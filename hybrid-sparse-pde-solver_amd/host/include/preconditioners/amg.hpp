#pragma once
// amg.hpp — AMGPreconditioner (reference amg.hpp:43-103): smoothed-aggregation
// AMG, host aggregation (bit-exact with the reference), device V-cycle.
#include <vector>

#include "preconditioner_base.hpp"

namespace hsps {

class AMGPreconditioner : public PreconditionerBase {
public:
    AMGPreconditioner() : PreconditionerBase(PrecondType::AMG) {}
    double setup(const SparseMatrix& A) override;
    const char* energy_tier() const override { return "EXPENSIVE"; }

    void set_max_levels(int n) { max_levels_ = n; }
    void set_strength_thresh(Real t) { strength_thresh_ = t; }
    void set_smooth_pre(int n) { smooth_pre_ = n; }
    void set_smooth_post(int n) { smooth_post_ = n; }
    /// Jacobi damping of the smoother and of the P smoothing (0.67, amg.hpp:101).
    void set_omega(Real w) { omega_ = w; }

    int levels() const;
    /// Aggregate map of level lv (0 <= lv < levels()-1).
    std::vector<Index> aggregates(int lv) const;
    /// rows / nnz of level lv's operator.
    std::pair<Index, Index> level_size(int lv) const;

private:
    int max_levels_ = 10;
    Real strength_thresh_ = 0.25;
    int smooth_pre_ = 2;
    int smooth_post_ = 2;
    Real omega_ = 0.67;
};

}  // namespace hsps
