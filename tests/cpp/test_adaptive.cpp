// Adaptive ladder + MatrixMarket loader through the host layer; cases follow
// the reference's tests/test_adaptive.cpp and the run_dataset fixture
// (dataset_results.csv) on the bundled matrices in tests/golden/data.
#include <fstream>

#include "adaptive/adaptive_selector.hpp"
#include "adaptive/ml_advisor.hpp"
#include "hsps_test.hpp"
#include "utils/matrix_market_io.hpp"

using namespace hsps;

static std::string data_dir() {
    const char* d = std::getenv("HSPS_DATA");
    return d ? d : "tests/golden/data";
}

TEST(easy_on_poisson_no_escalation) {
    auto A = SparseMatrix::poisson_2d(16);
    Vector b(256, 1.0), x;
    AdaptiveSelector sel;
    SolverStats s;
    expect_true(sel.solve(A, b, x, s), "converged");
    expect_true(s.state_used == AdaptiveState::EASY && sel.escalation_log().empty(), "EASY, no escalation");
    expect_true(s.iterations == 28, "28 iterations (reference ladder fixture)");
}
TEST(escalates_on_iteration_budget) {
    auto A = SparseMatrix::poisson_2d(24);
    Vector b(A.rows(), 1.0), x;
    SolverParams p;
    p.max_iter = 5;  // EASY and MODERATE exhaust their budget
    AdaptiveSelector sel(p);
    SolverStats s;
    sel.solve(A, b, x, s);
    expect_true(!sel.escalation_log().empty(), "escalated");
    expect_true(sel.escalation_log()[0].from_state == AdaptiveState::EASY, "from EASY");
}
TEST(convdiff_goes_moderate) {
    auto A = SparseMatrix::convection_diffusion_2d(16, 0.1, 5.0, 0.0);
    Vector b(A.rows(), 1.0), x;
    AdaptiveSelector sel;
    SolverStats s;
    expect_true(sel.solve(A, b, x, s), "converged");
    expect_true(s.state_used == AdaptiveState::MODERATE && s.iterations == 4, "MODERATE, 4 iterations (fixture)");
}
TEST(dataset_results_replay) {
    // dataset_results.csv: name, final_state, iterations (run_dataset.cpp:305-310 params)
    struct Row { const char* name; int state; int iters; };
    const Row rows[] = {{"convdiff_upwind_484", 2, 4}, {"helmholtz_400", 0, 37}, {"pde225", 0, 27},
                        {"pde2961", 0, 101}, {"pde900", 0, 55}, {"poisson2D", 2, 11}};
    for (const Row& r : rows) {
        const std::string mp = data_dir() + "/" + r.name + ".mtx";
        const std::string rp = data_dir() + "/" + r.name + "_rhs.mtx";
        std::ifstream chk(rp);
        auto prob = MatrixMarketIO::load_problem(mp, chk.good() ? rp : "");
        SolverParams p;
        p.tol = 1e-8;
        p.max_iter = 5000;
        p.restart_size = 60;
        p.stall_window = 20;
        AdaptiveSelector sel(p);
        sel.set_ml_advisor(std::make_shared<FeatureAdvisor>(p));
        Vector x(prob.first.rows(), 0.0);
        SolverStats s;
        expect_true(sel.solve(prob.first, prob.second, x, s), r.name);
        expect_true((int)s.state_used == r.state, std::string(r.name) + " state");
        expect_less(std::abs(s.iterations - r.iters), 0.05 * r.iters + 1.0, std::string(r.name) + " iterations");
        expect_less(s.true_residual, 1e-8, std::string(r.name) + " true residual");
    }
}
TEST(state_features_pick_the_same_state) {
    // the server's classification (extract_state_features) reads every field
    // select_initial_state reads, with the same values as extract_features
    const SparseMatrix mats[] = {SparseMatrix::poisson_2d(40), SparseMatrix::convection_diffusion_2d(30, 0.01, 1.0, 0.3),
                                 SparseMatrix::anisotropic_2d(36, 1e-3), SparseMatrix::convection_diffusion_2d(24, 1.0, 0.0, 0.0)};
    AdaptiveSelector sel;
    for (const auto& A : mats) {
        const auto f = sel.extract_features(A), g = sel.extract_state_features(A);
        expect_true(f.is_spd == g.is_spd && f.diag_dominance == g.diag_dominance, "state fields equal");
        expect_less(std::abs(f.estimated_cond - g.estimated_cond), 1e-12 * f.estimated_cond, "condition estimate");
        expect_true(sel.select_initial_state(f) == sel.select_initial_state(g), "same initial state");
    }
}
int main() { return run_all(); }
