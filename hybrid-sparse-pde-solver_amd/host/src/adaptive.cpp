// =============================================================================
// adaptive.cpp — AdaptiveSelector (adaptive_selector.cpp:38-394) and the
// host advisors (ml_advisor.cpp:17-100) of the host layer.
// =============================================================================
#include <algorithm>
#include <cmath>
#include <iomanip>
#include <sstream>

#include "adaptive/adaptive_selector.hpp"
#include "adaptive/ml_advisor.hpp"
#include "solvers/solver_factory.hpp"
#include "utils/logger.hpp"
#include "utils/timer.hpp"

namespace hsps {

AdaptiveSelector::AdaptiveSelector(const SolverParams& params) : params_(params) {}

AdaptiveSelector::MatrixFeatures AdaptiveSelector::extract_state_features(const SparseMatrix& A) const {
    MatrixFeatures f;
    double dev[8];
    dev_check(hsps_features(A.device(), dev));  // Lanczos on the device + structural scans
    f.n = A.rows();
    f.nnz = A.nnz();
    f.density = A.density();
    f.diag_dominance = dev[2];
    f.symmetry_residual = dev[3];
    f.off_diag_symmetry = dev[3];
    f.frobenius_norm = dev[4];
    f.estimated_cond = dev[5];
    f.spectral_radius_est = dev[6];
    f.lambda_min_est = dev[6] / dev[5];
    f.is_spd = dev[7] != 0.0;
    return f;
}

AdaptiveSelector::MatrixFeatures AdaptiveSelector::extract_features(const SparseMatrix& A) const {
    MatrixFeatures f = extract_state_features(A);
    // extended fields (adaptive_selector.cpp:130-196)
    f.bandwidth = A.bandwidth();
    f.avg_nnz_per_row = f.n > 0 ? static_cast<double>(f.nnz) / f.n : 0.0;
    f.nnz_variance = A.nnz_variance();
    const auto& rp = A.row_ptr();
    const auto& ci = A.col_idx();
    const auto& v = A.values();
    int pos = 0;
    Real gr = 0;
    for (Index i = 0; i < f.n; ++i) {
        Real s = 0;
        for (Index k = rp[i]; k < rp[i + 1]; ++k) s += std::abs(v[k]);
        gr = std::max(gr, s);
        for (Index k = rp[i]; k < rp[i + 1]; ++k)
            if (ci[k] == i) { if (v[k] > 0.0) ++pos; break; }
    }
    f.diag_sign_fraction = static_cast<double>(pos) / std::max(f.n, 1);
    f.gershgorin_radius = gr;
    f.diagonal_fraction = A.diagonal_fraction();
    f.lower_upper_ratio = A.lower_upper_ratio();
    f.n_zero_diag = A.count_zero_diagonal();
    f.fill_ratio_ilu0 = A.estimated_ilu_fill();
    f.coarsening_ratio = A.estimated_amg_coarsening_ratio(0.25);
    f.estimated_amg_depth = std::max(1, static_cast<int>(std::log2(static_cast<double>(f.n) / 50.0)));
    constexpr double af = 2e-10, am = 5e-9;
    const double ie = std::min(std::sqrt(f.estimated_cond) * 0.5, 5000.0);
    const double ih = std::min(std::log(f.estimated_cond + 1.0) * 5.0, 200.0);
    f.energy_proxy_easy = ie * (af * 2.0 * f.nnz + am * 2.0 * f.n * 8);
    f.energy_proxy_hard = ih * (af * 2.0 * f.nnz + am * 2.0 * f.n * 8);
    return f;
}

AdaptiveState AdaptiveSelector::select_initial_state(const MatrixFeatures& f) const {  // :225-233
    if (f.is_spd && f.estimated_cond < 1e6 && f.diag_dominance >= 0.8) return AdaptiveState::EASY;
    if (f.estimated_cond >= 1e6 || f.diag_dominance < 0.3) return AdaptiveState::HARD;
    return AdaptiveState::MODERATE;
}

bool AdaptiveSelector::should_escalate(AdaptiveState st, int iteration, Real now,
                                       const std::vector<Real>& hist) const {  // :238-267
    const int wnd = params_.stall_window;
    if (hist.size() >= 2 && now > 10.0 * hist[0]) return true;
    if ((int)hist.size() >= wnd) {
        const Real prev = hist[hist.size() - wnd];
        if (prev > REAL_EPS && now / prev > params_.stall_threshold) return true;
    }
    if (st == AdaptiveState::EASY && iteration > params_.max_iter * 4 / 10) return true;
    if (st == AdaptiveState::MODERATE && iteration > params_.max_iter * 7 / 10) return true;
    return false;
}

AdaptiveState AdaptiveSelector::next_state(AdaptiveState s) const {
    return s == AdaptiveState::EASY ? AdaptiveState::MODERATE : AdaptiveState::HARD;
}

bool AdaptiveSelector::solve(const SparseMatrix& A, const Vector& b, Vector& x, SolverStats& stats) {  // :298-394
    auto features = extract_features(A);
    AdaptiveState state;
    if (ml_advisor_) {
        auto adv = ml_advisor_->advise(features);
        state = adv.initial_state;
        const Precision keep = params_.precision;
        params_ = adv.params;
        params_.precision = keep;
        HSPS_LOG_INFO("MLAdvisor [", ml_advisor_->name(), "]: ", adv.rationale);
    } else {
        state = select_initial_state(features);
    }
    current_state_ = state;
    if (x.size() != A.rows()) x.resize(A.rows(), REAL_ZERO);
    bool solved = false;
    while (!solved) {
        double setup = 0.0;
        auto pair = SolverFactory::make_for_state(state, A, params_, &setup);
        SolverStats at;
        at.setup_time_s = setup;
        at.state_used = state;
        solved = pair.first->solve(A, b, x, at);  // x stays on the device between attempts
        at.state_used = state;
        stats = at;
        stats.setup_time_s += at.setup_time_s;  // as the reference reports it (:350-351)
        if (solved) break;
        if (state == AdaptiveState::HARD) {
            HSPS_LOG_WARN("AdaptiveSelector: FAILED to converge even at HARD state.");
            break;
        }
        const AdaptiveState ns = next_state(state);
        esc_log_.push_back({at.iterations, state, ns, at.final_residual, "convergence stall / budget exhausted"});
        state = ns;
        current_state_ = state;
    }
    stats.state_used = current_state_;
    if (ml_advisor_) ml_advisor_->record_outcome(features, SolverAdvice{current_state_, params_, ""}, stats);
    return solved;
}

void AdaptiveSelector::print_summary(std::ostream& os) const {
    os << "\n=== Adaptive Selector Summary ===\n  Final state     : " << to_string(current_state_)
       << "\n  Escalations     : " << esc_log_.size() << "\n";
    for (size_t i = 0; i < esc_log_.size(); ++i) {
        const auto& e = esc_log_[i];
        os << "  [" << i << "] iter " << e.at_iteration << "  " << to_string(e.from_state) << " → "
           << to_string(e.to_state) << "  rel_res=" << std::scientific << e.residual_at_escalation << "\n";
    }
    os << "=================================\n";
}

std::vector<double> AdaptiveSelector::MatrixFeatures::to_feature_vector() const {
    return {static_cast<double>(n), static_cast<double>(nnz), density, diag_dominance, symmetry_residual,
            frobenius_norm, estimated_cond, is_spd ? 1.0 : 0.0, bandwidth, avg_nnz_per_row, nnz_variance,
            diag_sign_fraction, off_diag_symmetry, spectral_radius_est, lambda_min_est, gershgorin_radius,
            diagonal_fraction, lower_upper_ratio, static_cast<double>(n_zero_diag), fill_ratio_ilu0,
            coarsening_ratio, static_cast<double>(estimated_amg_depth), energy_proxy_easy, energy_proxy_hard};
}

// ---- advisors --------------------------------------------------------------------------
SolverAdvice HeuristicAdvisor::advise(const AdaptiveSelector::MatrixFeatures& f) const {  // ml_advisor.cpp:17-32
    SolverAdvice a;
    a.params = base_;
    if (f.is_spd && f.estimated_cond < 1e6 && f.diag_dominance >= 0.8) {
        a.initial_state = AdaptiveState::EASY;
        a.rationale = "SPD + well-conditioned → CG+Jacobi";
    } else if (f.estimated_cond >= 1e6 || f.diag_dominance < 0.3) {
        a.initial_state = AdaptiveState::HARD;
        a.rationale = "High cond or poor diag-dominance → FGMRES+AMG";
    } else {
        a.initial_state = AdaptiveState::MODERATE;
        a.rationale = "Moderate → FGMRES+ILU";
    }
    return a;
}

SolverAdvice FeatureAdvisor::advise(const AdaptiveSelector::MatrixFeatures& f) const {  // ml_advisor.cpp:38-100
    SolverAdvice a;
    a.params = base_;
    const bool sym = f.symmetry_residual < symm_thresh_;
    AdaptiveState tier;
    if (f.estimated_cond < cond_moderate_thresh_ && f.diag_dominance >= dd_good_thresh_ && sym && f.is_spd)
        tier = AdaptiveState::EASY;
    else if (f.estimated_cond >= cond_hard_thresh_ || f.diag_dominance < dd_bad_thresh_ || !sym)
        tier = AdaptiveState::HARD;
    else
        tier = AdaptiveState::MODERATE;
    if (tier == AdaptiveState::HARD)
        a.params.restart_size = std::min(a.params.restart_size * 2, f.n);
    else if (tier == AdaptiveState::MODERATE)
        a.params.restart_size = std::min(static_cast<int>(a.params.restart_size * 1.5), f.n);
    if (f.estimated_cond > 1e5) a.params.drop_tol = 1e-5;
    else if (f.estimated_cond > 1e3) a.params.drop_tol = 1e-4;
    else a.params.drop_tol = 1e-3;
    if (f.diag_dominance > 0.9) a.params.amg_strong = 0.5;
    else if (f.diag_dominance > 0.5) a.params.amg_strong = 0.25;
    else a.params.amg_strong = 0.1;
    a.params.amg_levels = std::max(5, static_cast<int>(std::log2(static_cast<double>(f.n) / 50.0)) + 2);
    a.initial_state = tier;
    std::ostringstream o;
    o << to_string(tier) << "  cond=" << std::scientific << std::setprecision(2) << f.estimated_cond
      << "  dd=" << std::fixed << std::setprecision(3) << f.diag_dominance << "  sym=" << (sym ? "yes" : "no")
      << "  spd=" << (f.is_spd ? "yes" : "no") << "  restart=" << a.params.restart_size
      << "  drop_tol=" << std::scientific << a.params.drop_tol;
    a.rationale = o.str();
    return a;
}

}  // namespace hsps
