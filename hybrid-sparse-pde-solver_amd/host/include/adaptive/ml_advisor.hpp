#pragma once
// =============================================================================
// ml_advisor.hpp — host advisors that pick the initial ladder state and tune
// SolverParams (reference ml_advisor.hpp:45-120): MLAdvisor interface,
// HeuristicAdvisor (the 3-branch rule) and FeatureAdvisor (the richer
// decision tree used by run_dataset).  The TorchScript GNN advisor is not
// part of the device path.
// =============================================================================
#include <memory>
#include <string>

#include "../core/types.hpp"
#include "adaptive_selector.hpp"

namespace hsps {

struct SolverAdvice {
    AdaptiveState initial_state = AdaptiveState::EASY;
    SolverParams params;
    std::string rationale;
};

class MLAdvisor {
public:
    virtual ~MLAdvisor() = default;
    virtual SolverAdvice advise(const AdaptiveSelector::MatrixFeatures& features) const = 0;
    virtual void record_outcome(const AdaptiveSelector::MatrixFeatures&, const SolverAdvice&, const SolverStats&) {}
    virtual const char* name() const = 0;
};

class HeuristicAdvisor : public MLAdvisor {
public:
    explicit HeuristicAdvisor(const SolverParams& base_params = {}) : base_(base_params) {}
    SolverAdvice advise(const AdaptiveSelector::MatrixFeatures& f) const override;
    const char* name() const override { return "HeuristicAdvisor"; }

private:
    SolverParams base_;
};

class FeatureAdvisor : public MLAdvisor {
public:
    explicit FeatureAdvisor(const SolverParams& base_params = {}) : base_(base_params) {}
    SolverAdvice advise(const AdaptiveSelector::MatrixFeatures& f) const override;
    const char* name() const override { return "FeatureAdvisor"; }

private:
    SolverParams base_;
    Real cond_moderate_thresh_ = 1e4;
    Real cond_hard_thresh_ = 1e7;
    Real dd_good_thresh_ = 0.8;
    Real dd_bad_thresh_ = 0.3;
    Real symm_thresh_ = 1e-4;
};

}  // namespace hsps
