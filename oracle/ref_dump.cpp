// =============================================================================
// ref_dump.cpp — fixture driver linked against a from-source build of the
// reference (oracle/_ref).  TEST INFRASTRUCTURE: it writes the golden JSON
// under tests/golden/ that pins the oracle and the device path.  Run through
// tests/golden/gen_golden.py; never shipped, never run on the GPU.
// =============================================================================
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "core/sparse_matrix.hpp"
#include "core/vector.hpp"
#include "solvers/cg_solver.hpp"
#include "solvers/fgmres_solver.hpp"
#include "preconditioners/jacobi.hpp"
#define private public  // read ILU factors and the AMG hierarchy (fixture only)
#include "preconditioners/ilu.hpp"
#include "preconditioners/amg.hpp"
#undef private
#include "adaptive/adaptive_selector.hpp"
#include "adaptive/ml_advisor.hpp"
#include "utils/matrix_market_io.hpp"
#include "utils/logger.hpp"

using namespace hsps;

static uint64_t sm64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static Vector seeded(uint64_t seed, int n) {
    Vector v(n);
    uint64_t s = seed;
    for (int i = 0; i < n; ++i) v[i] = 2.0 * ((double)(sm64(s) >> 11) * 0x1.0p-53) - 1.0;
    return v;
}
static uint64_t fnv(const std::vector<int>& a) {
    uint64_t h = 1469598103934665603ull;
    for (int x : a) {
        uint32_t u = (uint32_t)x;
        for (int b = 0; b < 4; ++b) { h ^= (u >> (8 * b)) & 0xff; h *= 1099511628211ull; }
    }
    return h;
}
static std::string num(double x) {
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.17g", x);
    return buf;
}
static std::string ivec(const std::vector<int>& a) {
    std::ostringstream o;
    o << "[";
    for (size_t i = 0; i < a.size(); ++i) o << (i ? "," : "") << a[i];
    o << "]";
    return o.str();
}
static std::string dvec(const double* a, size_t n) {
    std::ostringstream o;
    o << "[";
    for (size_t i = 0; i < n; ++i) o << (i ? "," : "") << num(a[i]);
    o << "]";
    return o.str();
}
static double relres(const SparseMatrix& A, const Vector& b, const Vector& x) {
    Vector r = A * x;
    for (int i = 0; i < b.size(); ++i) r[i] = b[i] - r[i];
    return r.norm2() / b.norm2();
}

struct Run {
    int iters;
    bool conv;
    double rel_true;
    double final_res;
    int restarts;
    Vector x;
    long long flops, bytes;  // SolverStats::flop_count / mem_bytes (cg_solver.cpp:178-179)
};

static Run run(const std::string& solver, const std::string& pre, const SparseMatrix& A, const Vector& b,
               SolverParams p, int amg_levels = 10) {
    std::shared_ptr<PreconditionerBase> M;
    if (pre == "jacobi") { auto j = std::make_shared<JacobiPreconditioner>(); j->setup(A); M = j; }
    if (pre == "ilu0") { auto j = std::make_shared<ILUPreconditioner>(); j->setup(A); M = j; }
    if (pre == "ilut") { auto j = std::make_shared<ILUPreconditioner>(); j->set_drop_tol(p.drop_tol); j->setup(A); M = j; }
    if (pre == "amg") { auto j = std::make_shared<AMGPreconditioner>(); j->set_max_levels(amg_levels); j->setup(A); M = j; }
    std::unique_ptr<SolverBase> s;
    if (solver == "cg") s = std::make_unique<CGSolver>(); else s = std::make_unique<FGMRESSolver>();
    s->set_params(p);
    if (M) s->set_preconditioner(M);
    Vector x(A.rows(), 0.0);
    SolverStats st;
    bool ok = s->solve(A, b, x, st);
    return {st.iterations, ok, relres(A, b, x), st.final_residual, st.restarts, x, st.flop_count, st.mem_bytes};
}

static std::string run_json(const Run& r) {
    double xn = r.x.norm2();
    std::ostringstream o;
    o << "{\"iters\":" << r.iters << ",\"converged\":" << (r.conv ? 1 : 0) << ",\"restarts\":" << r.restarts
      << ",\"rel_true\":" << num(r.rel_true) << ",\"final_residual\":" << num(r.final_res)
      << ",\"x_norm\":" << num(xn) << ",\"x_head\":" << dvec(r.x.data(), std::min(8, r.x.size()))
      << ",\"flop_count\":" << r.flops << ",\"mem_bytes\":" << r.bytes << "}";
    return o.str();
}

static std::string amg_json(const std::string& name, const SparseMatrix& A, double theta, int full_limit) {
    AMGPreconditioner amg;
    amg.set_strength_thresh(theta);
    amg.setup(A);
    std::ostringstream o;
    o << "{\"case\":\"" << name << "\",\"theta\":" << num(theta) << ",\"levels\":[";
    for (size_t lv = 0; lv < amg.hierarchy_.size(); ++lv) {
        const auto& L = amg.hierarchy_[lv];
        o << (lv ? "," : "") << "{\"rows\":" << L.A.rows() << ",\"nnz\":" << L.A.nnz();
        double asum = 0, aabs = 0;
        for (double v : L.A.values()) { asum += v; aabs += std::abs(v); }
        o << ",\"a_sum\":" << num(asum) << ",\"a_abs\":" << num(aabs);
        if (lv + 1 < amg.hierarchy_.size()) {
            std::vector<int> agg = amg.aggregate(L.A);
            o << ",\"agg_hash\":\"" << fnv(agg) << "\",\"n_agg\":" << (*std::max_element(agg.begin(), agg.end()) + 1);
            if ((int)agg.size() <= full_limit) o << ",\"agg\":" << ivec(agg);
            double psum = 0, pabs = 0;
            for (double v : L.P.values()) { psum += v; pabs += std::abs(v); }
            o << ",\"p_nnz\":" << L.P.nnz() << ",\"p_sum\":" << num(psum) << ",\"p_abs\":" << num(pabs);
            if (L.P.rows() <= 64) {
                o << ",\"p_rp\":" << ivec(L.P.row_ptr()) << ",\"p_ci\":" << ivec(L.P.col_idx())
                  << ",\"p_v\":" << dvec(L.P.values().data(), L.P.values().size());
            }
        }
        o << "}";
    }
    o << "]}";
    return o.str();
}

int main(int argc, char** argv) {
    if (argc < 3) { std::fprintf(stderr, "usage: ref_dump OUT.json DATA_DIR\n"); return 2; }
    Logger::instance().set_level(LogLevel::ERROR);
    std::ofstream out(argv[1]);
    const std::string data = argv[2];
    out << "{\n";

    // README table (README.md:195-199): Poisson 64^2, b = 1, tol 1e-8.
    {
        auto A = SparseMatrix::poisson_2d(64);
        Vector b(A.rows(), 1.0);
        SolverParams p;
        p.tol = 1e-8; p.max_iter = 2000; p.restart_size = 50;
        out << "\"readme_b1_64\":{";
        out << "\"cg_jacobi\":" << run_json(run("cg", "jacobi", A, b, p));
        out << ",\"cg_ilu0\":" << run_json(run("cg", "ilu0", A, b, p));
        out << ",\"fgmres50_jacobi\":" << run_json(run("fgmres", "jacobi", A, b, p));
        out << ",\"fgmres50_ilu0\":" << run_json(run("fgmres", "ilu0", A, b, p));
        out << ",\"fgmres50_amg\":" << run_json(run("fgmres", "amg", A, b, p));
        out << "},\n";
    }
    // Seeded cases: b = A x*, x* ~ U[-1,1] (splitmix64, seed).
    out << "\"seeded\":[";
    {
        struct C { std::string name; std::string kind; int n; double a, bx, by; uint64_t seed; };
        std::vector<C> cs = {
            {"poisson16_s1", "poisson", 16, 0, 0, 0, 1},
            {"poisson32_s7", "poisson", 32, 0, 0, 0, 7},
            {"poisson64_s42", "poisson", 64, 0, 0, 0, 42},
            {"poisson128_s3", "poisson", 128, 0, 0, 0, 3},
            {"convdiff32_s5", "convdiff", 32, 0.01, 0.7648421872844885, 0.644217687237691, 5},
            {"convdiff64_s9", "convdiff", 64, 0.01, 0.7648421872844885, 0.644217687237691, 9},
        };
        bool first = true;
        for (auto& c : cs) {
            SparseMatrix A = c.kind == "poisson" ? SparseMatrix::poisson_2d(c.n)
                                                 : SparseMatrix::convection_diffusion_2d(c.n, c.a, c.bx, c.by);
            Vector xs = seeded(c.seed, A.rows());
            Vector b = A * xs;
            SolverParams p;
            p.tol = 1e-8; p.max_iter = 20000; p.restart_size = 30;
            out << (first ? "" : ",") << "\n{\"case\":\"" << c.name << "\",\"kind\":\"" << c.kind << "\",\"n\":" << c.n
                << ",\"eps\":" << num(c.a) << ",\"bx\":" << num(c.bx) << ",\"by\":" << num(c.by) << ",\"seed\":" << c.seed;
            if (c.kind == "poisson") out << ",\"cg_jacobi\":" << run_json(run("cg", "jacobi", A, b, p));
            out << ",\"fgmres30_jacobi\":" << run_json(run("fgmres", "jacobi", A, b, p));
            out << ",\"fgmres30_ilu0\":" << run_json(run("fgmres", "ilu0", A, b, p));
            if (c.n <= 128) out << ",\"fgmres30_amg\":" << run_json(run("fgmres", "amg", A, b, p));
            out << "}";
            first = false;
        }
    }
    out << "],\n";
    // AMG hierarchies.
    out << "\"amg\":[";
    out << amg_json("poisson8", SparseMatrix::poisson_2d(8), 0.25, 100000) << ",\n";
    out << amg_json("poisson16", SparseMatrix::poisson_2d(16), 0.25, 100000) << ",\n";
    out << amg_json("poisson32", SparseMatrix::poisson_2d(32), 0.25, 100000) << ",\n";
    out << amg_json("poisson64", SparseMatrix::poisson_2d(64), 0.25, 5000) << ",\n";
    out << amg_json("poisson128", SparseMatrix::poisson_2d(128), 0.25, 0) << ",\n";
    out << amg_json("poisson180", SparseMatrix::poisson_2d(180), 0.25, 0) << ",\n";
    out << amg_json("poisson64_t05", SparseMatrix::poisson_2d(64), 0.5, 0) << ",\n";
    out << amg_json("convdiff64", SparseMatrix::convection_diffusion_2d(64, 0.01, 0.7648421872844885, 0.644217687237691), 0.25, 0) << ",\n";
    out << amg_json("convdiff48_b", SparseMatrix::convection_diffusion_2d(48, 0.05, 1.0, 0.5), 0.25, 0);
    out << "],\n";
    // ILU factors (small, full values).
    out << "\"ilu\":[";
    {
        auto A = SparseMatrix::poisson_2d(8);
        ILUPreconditioner f; f.setup(A);
        out << "{\"case\":\"poisson8_ilu0\",\"rp\":" << ivec(f.row_ptr_) << ",\"ci\":" << ivec(f.col_idx_)
            << ",\"lu\":" << dvec(f.lu_val_.data(), f.lu_val_.size()) << "},\n";
        auto B = SparseMatrix::convection_diffusion_2d(12, 0.1, 5.0, 0.0);
        ILUPreconditioner g; g.set_drop_tol(1e-4); g.setup(B);
        out << "{\"case\":\"convdiff12_ilut1e-4\",\"rp\":" << ivec(g.row_ptr_) << ",\"ci\":" << ivec(g.col_idx_)
            << ",\"lu\":" << dvec(g.lu_val_.data(), g.lu_val_.size()) << "}";
    }
    out << "],\n";
    // Adaptive ladder with FeatureAdvisor on the bundled matrices
    // (run_dataset.cpp:106-155, params :305-310).
    out << "\"dataset\":[";
    {
        const char* names[] = {"convdiff_upwind_484", "helmholtz_400", "pde225", "pde2961", "pde900", "poisson2D"};
        bool first = true;
        for (const char* nm : names) {
            std::string mp = data + "/" + nm + ".mtx", rp = data + "/" + nm + "_rhs.mtx";
            std::ifstream chk(rp);
            auto prob = MatrixMarketIO::load_problem(mp, chk.good() ? rp : "");
            SparseMatrix& A = prob.first;
            Vector& b = prob.second;
            SolverParams p;
            p.tol = 1e-8; p.max_iter = 5000; p.restart_size = 60; p.stall_window = 20;
            AdaptiveSelector sel(p);
            sel.set_ml_advisor(std::make_shared<FeatureAdvisor>(p));
            auto feats = sel.extract_features(A);
            Vector x(A.rows(), 0.0);
            SolverStats st;
            bool ok = sel.solve(A, b, x, st);
            out << (first ? "" : ",") << "\n{\"name\":\"" << nm << "\",\"n\":" << A.rows() << ",\"nnz\":" << A.nnz()
                << ",\"state\":" << (int)st.state_used << ",\"iters\":" << st.iterations << ",\"converged\":" << (ok ? 1 : 0)
                << ",\"escalations\":" << sel.escalation_log().size() << ",\"rel_res\":" << num(relres(A, b, x))
                << ",\"cond\":" << num(feats.estimated_cond) << ",\"dd\":" << num(feats.diag_dominance)
                << ",\"sym\":" << num(feats.symmetry_residual) << ",\"spd\":" << (feats.is_spd ? 1 : 0)
                << ",\"x_norm\":" << num(x.norm2()) << "}";
            first = false;
        }
    }
    out << "],\n";
    // Heuristic (no advisor) ladder on generated problems.
    out << "\"ladder\":[";
    {
        struct C { std::string name; std::string kind; int n; double a, bx, by; };
        std::vector<C> cs = {{"poisson16", "poisson", 16, 0, 0, 0},
                             {"poisson40", "poisson", 40, 0, 0, 0},
                             {"convdiff16", "convdiff", 16, 0.1, 5.0, 0.0},
                             {"convdiff24_hard", "convdiff", 24, 0.001, 1.0, 0.0}};
        bool first = true;
        for (auto& c : cs) {
            SparseMatrix A = c.kind == "poisson" ? SparseMatrix::poisson_2d(c.n)
                                                 : SparseMatrix::convection_diffusion_2d(c.n, c.a, c.bx, c.by);
            Vector b(A.rows(), 1.0);
            SolverParams p;
            AdaptiveSelector sel(p);
            auto feats = sel.extract_features(A);
            Vector x(A.rows(), 0.0);
            SolverStats st;
            bool ok = sel.solve(A, b, x, st);
            out << (first ? "" : ",") << "\n{\"name\":\"" << c.name << "\",\"kind\":\"" << c.kind << "\",\"n\":" << c.n
                << ",\"eps\":" << num(c.a) << ",\"bx\":" << num(c.bx) << ",\"by\":" << num(c.by)
                << ",\"state\":" << (int)st.state_used << ",\"iters\":" << st.iterations << ",\"converged\":" << (ok ? 1 : 0)
                << ",\"escalations\":" << sel.escalation_log().size() << ",\"rel_res\":" << num(relres(A, b, x))
                << ",\"cond\":" << num(feats.estimated_cond) << ",\"dd\":" << num(feats.diag_dominance)
                << ",\"spd\":" << (feats.is_spd ? 1 : 0) << "}";
            first = false;
        }
    }
    out << "]\n}\n";
    return 0;
}
