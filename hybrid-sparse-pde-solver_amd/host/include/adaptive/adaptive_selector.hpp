#pragma once
// =============================================================================
// adaptive_selector.hpp — the EASY -> MODERATE -> HARD ladder with the
// reference's API (adaptive_selector.hpp:49-161).  Features come from the
// device (Lanczos) plus host scans; each attempt runs on the device and the
// warm-start x stays on the device between attempts.
// =============================================================================
#include <memory>
#include <ostream>
#include <string>
#include <vector>

#include "../core/sparse_matrix.hpp"
#include "../core/types.hpp"
#include "../core/vector.hpp"
#include "../energy/energy_monitor.hpp"
#include "../preconditioners/preconditioner_base.hpp"
#include "../solvers/solver_base.hpp"

namespace hsps {

class MLAdvisor;
struct SolverAdvice;

struct EscalationRecord {
    int at_iteration;
    AdaptiveState from_state;
    AdaptiveState to_state;
    double residual_at_escalation;
    std::string reason;
};

class AdaptiveSelector {
public:
    explicit AdaptiveSelector(const SolverParams& params = {});

    bool solve(const SparseMatrix& A, const Vector& b, Vector& x, SolverStats& stats);

    struct MatrixFeatures {
        int n = 0;
        int nnz = 0;
        double density = 0;
        double diag_dominance = 0;
        double symmetry_residual = 0;
        double frobenius_norm = 0;
        double estimated_cond = 0;
        bool is_spd = false;
        double bandwidth = 0;
        double avg_nnz_per_row = 0;
        double nnz_variance = 0;
        double diag_sign_fraction = 0;
        double off_diag_symmetry = 0;
        double spectral_radius_est = 0;
        double lambda_min_est = 0;
        double gershgorin_radius = 0;
        double diagonal_fraction = 0;
        double lower_upper_ratio = 0;
        int n_zero_diag = 0;
        double fill_ratio_ilu0 = 0;
        double coarsening_ratio = 0;
        int estimated_amg_depth = 0;
        double energy_proxy_easy = 0;
        double energy_proxy_hard = 0;
        std::vector<double> to_feature_vector() const;
        static int feature_dim() { return 24; }
    };
    MatrixFeatures extract_features(const SparseMatrix& A) const;
    // The fields select_initial_state reads (n, nnz, density and the device
    // scan: diagonal dominance, symmetry, norms, the Lanczos condition
    // estimate, SPD), without the host scans of the extended fields: what a
    // server needs to pick a request's first ladder state.
    MatrixFeatures extract_state_features(const SparseMatrix& A) const;
    AdaptiveState select_initial_state(const MatrixFeatures& f) const;
    bool should_escalate(AdaptiveState current_state, int iteration, Real residual_now,
                         const std::vector<Real>& residual_history) const;

    const std::vector<EscalationRecord>& escalation_log() const { return esc_log_; }
    AdaptiveState current_state() const { return current_state_; }
    void print_summary(std::ostream& os) const;
    void set_ml_advisor(std::shared_ptr<MLAdvisor> advisor) { ml_advisor_ = advisor; }
    void set_energy_monitor(std::shared_ptr<EnergyMonitor> mon) { energy_mon_ = mon; }

private:
    AdaptiveState next_state(AdaptiveState s) const;
    SolverParams params_;
    AdaptiveState current_state_ = AdaptiveState::EASY;
    std::vector<EscalationRecord> esc_log_;
    std::shared_ptr<EnergyMonitor> energy_mon_ = nullptr;
    std::shared_ptr<MLAdvisor> ml_advisor_ = nullptr;
};

}  // namespace hsps
