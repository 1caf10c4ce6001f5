// =============================================================================
// hsps_oracle.hpp — FP64 CPU oracle for the HAMLSS solve path.
//
// TEST INFRASTRUCTURE ONLY.  This is a plain C++17 restatement of the
// reference's solve path (CSR SpMV, BLAS-1, Jacobi, ILU(0)/ILUT, SA-AMG,
// PCG, FGMRES, the adaptive ladder and its feature advisor), written from the
// reference's behaviour.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it, and only as the checker / CPU baseline.  The
// product path never calls into it.
//
// Every function cites the reference file:line it follows (paths relative to
// the reference repository root).  Loops and OpenMP schedules follow the
// reference so that, built with -fopenmp, this is also the OpenMP CPU
// baseline timed beside the GPU.
//
// Pinning (what checks each part, beyond retyping the formula):
//   generators      closed forms (nnz = 5N-4n, row sums, entries), reference
//                   from-source fixture comparison
//   spmv / blas     dense brute force on small cases, Poisson A*1 closed form
//   jacobi          diagonal known answers (reference tests)
//   ilu0            tridiagonal == Thomas algorithm; ILUT(0,0) == ILU(0);
//                   reference fixture iteration counts
//   amg             aggregate maps / hierarchy shapes / P entries vs the
//                   reference build; dense GEPP residual
//   cg / fgmres     reference iteration counts (README 119/52/269/51, seeded
//                   config-1 173, FGMRES+AMG 8/10), textbook invariants
//   adaptive        dataset_results.csv (state, iterations, rel_res)
// =============================================================================
#pragma once
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

namespace orc {

using Index = int;
using Real = double;
constexpr Real EPS = 2.220446049250313e-16;  // numeric_limits<double>::epsilon

// ---- CSR container (sparse_matrix.hpp:34-170, members :162-166) -------------
struct Csr {
    Index rows = 0, cols = 0;
    std::vector<Index> rp{0};  // rows+1
    std::vector<Index> ci;     // nnz
    std::vector<Real> v;       // nnz
    Index nnz() const { return static_cast<Index>(v.size()); }
};

Csr from_coo(Index rows, Index cols, const std::vector<Index>& r,
             const std::vector<Index>& c, const std::vector<Real>& v);
Csr poisson_2d(Index n);
Csr convection_diffusion_2d(Index n, Real eps, Real bx, Real by);
// Anisotropic diffusion -eps*u_xx - u_yy, 5-point, no h^2 scaling (no
// reference counterpart: the batched workload's second family).
Csr anisotropic_2d(Index n, Real eps);

void spmv(const Csr& A, const Real* x, Real* y);
Real dot(Index n, const Real* a, const Real* b);
Real norm2(Index n, const Real* a);
Csr transpose(const Csr& A);
Csr mat_mat_product(const Csr& A, const Csr& B);
std::vector<Real> diagonal(const Csr& A);

// ---- seeded inputs: splitmix64 U[-1,1] -----------------------------------------
uint64_t splitmix64(uint64_t& state);
void uniform_pm1(uint64_t seed, Index n, Real* out);

// ---- parameters / stats (types.hpp:91-143) --------------------------------------
struct Params {
    Real tol = 1e-8;
    int max_iter = 2000;
    int restart_size = 50;
    Real drop_tol = 1e-4;
    int amg_levels = 10;
    Real amg_strong = 0.25;
    int amg_smooth_pre = 2;
    int amg_smooth_post = 2;
    int fill_per_row = 0;
    int sstep_k = 1;             // types.hpp:134
    Real sstep_cond_tol = 1e12;  // types.hpp:135
};
struct Stats {
    int iterations = 0;
    Real initial_residual = 0, final_residual = 0;
    bool converged = false;
    int restarts = 0;
    long long flop_count = 0, mem_bytes = 0;
    double energy_joules = 0;
    double setup_time_s = 0, solve_time_s = 0;
    int state = 0;  // 0 EASY 1 MODERATE 2 HARD
    int escalations = 0;
};

// ---- preconditioners (preconditioner_base.hpp:16-48) -----------------------------
struct Precond {
    virtual ~Precond() = default;
    virtual void apply(const Real* r, Real* z) const = 0;
    Index n = 0;
};

struct Jacobi : Precond {  // jacobi.cpp:13-51
    std::vector<Real> inv_diag;
    Real omega = 1.0;
    void setup(const Csr& A);
    void apply(const Real* r, Real* z) const override;
};

struct Ilu : Precond {  // ilu.cpp:31-271
    Real drop_tol = 0.0;
    int fill_p = 0;
    std::vector<Index> rp, ci;
    std::vector<Real> lu;
    void setup(const Csr& A);
    void setup_ilu0(const Csr& A);
    void setup_ilut(const Csr& A);
    void forward_sub(const Real* r, Real* y) const;
    void backward_sub(const Real* y, Real* z) const;
    void apply(const Real* r, Real* z) const override;
};

struct AmgLevel {
    Csr A, P, R;
    std::vector<Real> dinv;
    std::vector<Index> agg;  // aggregate map of this level (empty at coarsest)
};

struct Amg : Precond {  // amg.cpp:30-343, amg.hpp:96-103
    int max_levels = 10;
    Real theta = 0.25;
    int pre = 2, post = 2;
    int coarse_limit = 50;
    Real omega = 0.67;
    bool smooth_p = true;
    std::vector<AmgLevel> h;
    void setup(const Csr& A);
    void apply(const Real* r, Real* z) const override;
    void v_cycle(int lv, const Real* rhs, Real* x) const;
    void smooth(int lv, const Real* rhs, Real* x, int steps) const;
};

std::vector<Index> aggregate(const Csr& A, Real theta);
Csr build_prolongation(Index n_fine, const std::vector<Index>& agg, Index n_coarse);
Csr smooth_prolongation(const Csr& A, const Csr& P_raw, Real omega);
void dense_gepp_solve(const Csr& A, const Real* rhs, Real* x);

// ---- Krylov solvers ----------------------------------------------------------
bool cg_solve(const Csr& A, const Real* b, Real* x, const Precond* M,
              const Params& p, Stats& st);
// s-step CG (sstep_cg_solver.cpp:168-305; the preconditioner is not used) and
// pipelined FGMRES (pipelined_fgmres_solver.cpp:56-243).
bool sstep_cg_solve(const Csr& A, const Real* b, Real* x, const Params& prm, Stats& st);
bool pipelined_fgmres_solve(const Csr& A, const Real* b, Real* x, const Precond* M, const Params& prm, Stats& st);
// Lanczos estimate of [lambda_min, lambda_max] (sparse_matrix.cpp:451-479).
void spectral_interval(const Csr& A, int steps, Real* lo, Real* hi);

bool fgmres_solve(const Csr& A, const Real* b, Real* x, const Precond* M,
                  const Params& p, Stats& st);

// ---- adaptive ladder (adaptive_selector.cpp:38-394, ml_advisor.cpp:17-100) ----
struct Features {
    int n = 0, nnz = 0;
    double density = 0, diag_dominance = 0, symmetry_residual = 0, frobenius = 0;
    double estimated_cond = 0, lambda_max = 0, lambda_min = 0;
    bool is_spd = false;
};
Features extract_features(const Csr& A);
int select_initial_state(const Features& f);  // HeuristicAdvisor rule
// FeatureAdvisor: returns state, rewrites params (ml_advisor.cpp:38-100).
int feature_advise(const Features& f, Params& p);
// advisor: 0 = built-in heuristic, 1 = FeatureAdvisor
bool adaptive_solve(const Csr& A, const Real* b, Real* x, Params p, int advisor,
                    Stats& st);

// Factory: build preconditioner for a (state) the way solver_factory.cpp:55-133
// does.  kind: 0 none, 1 jacobi, 2 ilu (drop_tol from params), 3 amg.
std::unique_ptr<Precond> make_precond(int kind, const Csr& A, const Params& p);

}  // namespace orc
