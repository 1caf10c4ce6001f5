// =============================================================================
// oracle_capi.cpp — extern "C" face of the FP64 CPU oracle, for ctypes.
// TEST INFRASTRUCTURE ONLY: loaded by tests/, smoke() and bench.py's
// cpu_baseline leg, never by the product path.
// =============================================================================
#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <exception>
#include <string>

#include "hsps_oracle.hpp"

using namespace orc;

namespace {
thread_local std::string g_err;
}

extern "C" {

struct orc_params {
    double tol;
    int max_iter;
    int restart_size;
    double drop_tol;
    int amg_levels;
    double amg_strong;
    int amg_smooth_pre;
    int amg_smooth_post;
    int sstep_k;
    double sstep_cond_tol;
};

struct orc_stats {
    int iterations;
    int converged;
    int restarts;
    int state;
    int escalations;
    double initial_residual;
    double final_residual;
    double setup_time_s;
    double solve_time_s;
    long long flop_count;
    long long mem_bytes;
    double energy_joules;
};

static Params to_params(const orc_params* p) {
    Params q;
    if (!p) return q;
    q.tol = p->tol;
    q.max_iter = p->max_iter;
    q.restart_size = p->restart_size;
    q.drop_tol = p->drop_tol;
    q.amg_levels = p->amg_levels;
    q.amg_strong = p->amg_strong;
    q.amg_smooth_pre = p->amg_smooth_pre;
    q.amg_smooth_post = p->amg_smooth_post;
    q.sstep_k = p->sstep_k;
    q.sstep_cond_tol = p->sstep_cond_tol;
    return q;
}

static void from_stats(const Stats& s, orc_stats* o) {
    if (!o) return;
    o->iterations = s.iterations;
    o->converged = s.converged ? 1 : 0;
    o->restarts = s.restarts;
    o->state = s.state;
    o->escalations = s.escalations;
    o->initial_residual = s.initial_residual;
    o->final_residual = s.final_residual;
    o->setup_time_s = s.setup_time_s;
    o->solve_time_s = s.solve_time_s;
    o->flop_count = s.flop_count;
    o->mem_bytes = s.mem_bytes;
    o->energy_joules = s.energy_joules;
}

const char* orc_last_error(void) { return g_err.c_str(); }

#define ORC_TRY(body)                         \
    try {                                     \
        body;                                 \
    } catch (const std::exception& e) {       \
        g_err = e.what();                     \
        return -1;                            \
    }

void* orc_csr_poisson(int n) { return new Csr(poisson_2d(n)); }
void* orc_csr_convdiff(int n, double eps, double bx, double by) {
    return new Csr(convection_diffusion_2d(n, eps, bx, by));
}
void* orc_csr_aniso(int n, double eps) { return new Csr(anisotropic_2d(n, eps)); }
void* orc_csr_from_coo(int rows, int cols, int nz, const int* r, const int* c, const double* v) {
    return new Csr(from_coo(rows, cols, std::vector<Index>(r, r + nz), std::vector<Index>(c, c + nz),
                            std::vector<Real>(v, v + nz)));
}
void* orc_csr_from_csr(int rows, int cols, const int* rp, const int* ci, const double* v) {
    Csr* A = new Csr;
    A->rows = rows;
    A->cols = cols;
    A->rp.assign(rp, rp + rows + 1);
    A->ci.assign(ci, ci + rp[rows]);
    A->v.assign(v, v + rp[rows]);
    return A;
}
void orc_csr_free(void* h) { delete static_cast<Csr*>(h); }
void orc_csr_info(void* h, int* rows, int* cols, int* nnz) {
    const Csr* A = static_cast<Csr*>(h);
    *rows = A->rows;
    *cols = A->cols;
    *nnz = A->nnz();
}
void orc_csr_export(void* h, int* rp, int* ci, double* v) {
    const Csr* A = static_cast<Csr*>(h);
    std::memcpy(rp, A->rp.data(), sizeof(int) * (A->rows + 1));
    std::memcpy(ci, A->ci.data(), sizeof(int) * A->nnz());
    std::memcpy(v, A->v.data(), sizeof(double) * A->nnz());
}
void orc_spmv(void* h, const double* x, double* y) { spmv(*static_cast<Csr*>(h), x, y); }
void orc_uniform(unsigned long long seed, int n, double* out) { uniform_pm1(seed, n, out); }
void orc_spectral_interval(void* h, int steps, double* lo, double* hi) {
    spectral_interval(*static_cast<Csr*>(h), steps, lo, hi);
}

// solver: 0 CG, 1 FGMRES, 2 s-step CG (no preconditioner), 3 pipelined
// FGMRES.  precond: 0 none, 1 Jacobi, 2 ILU (params.drop_tol; 0 gives
// ILU(0)), 3 AMG.
int orc_solve(void* h, const double* b, double* x, int solver, int precond, const orc_params* p,
              orc_stats* out) {
    ORC_TRY({
        const Csr& A = *static_cast<Csr*>(h);
        Params q = to_params(p);
        Stats st;
        auto M = make_precond(precond, A, q);
        bool ok;
        switch (solver) {
            case 0: ok = cg_solve(A, b, x, M.get(), q, st); break;
            case 1: ok = fgmres_solve(A, b, x, M.get(), q, st); break;
            case 2: ok = sstep_cg_solve(A, b, x, q, st); break;
            case 3: ok = pipelined_fgmres_solve(A, b, x, M.get(), q, st); break;
            default: throw std::invalid_argument("orc_solve: unknown solver");
        }
        from_stats(st, out);
        return ok ? 1 : 0;
    });
}

int orc_precond_apply(void* h, int precond, const orc_params* p, const double* r, double* z) {
    ORC_TRY({
        const Csr& A = *static_cast<Csr*>(h);
        auto M = make_precond(precond, A, to_params(p));
        if (M) M->apply(r, z); else std::memcpy(z, r, sizeof(double) * A.rows);
        return 0;
    });
}

int orc_adaptive(void* h, const double* b, double* x, const orc_params* p, int advisor, orc_stats* out) {
    ORC_TRY({
        Stats st;
        const bool ok = adaptive_solve(*static_cast<Csr*>(h), b, x, to_params(p), advisor, st);
        from_stats(st, out);
        return ok ? 1 : 0;
    });
}

// Features: out[0..7] = n, nnz, diag_dominance, symmetry_residual, frobenius,
// estimated_cond, lambda_max, is_spd.
int orc_features(void* h, double* out) {
    ORC_TRY({
        Features f = extract_features(*static_cast<Csr*>(h));
        out[0] = f.n; out[1] = f.nnz; out[2] = f.diag_dominance; out[3] = f.symmetry_residual;
        out[4] = f.frobenius; out[5] = f.estimated_cond; out[6] = f.lambda_max; out[7] = f.is_spd ? 1 : 0;
        return 0;
    });
}

// AMG hierarchy handle.
void* orc_amg_setup(void* h, const orc_params* p) {
    try {
        Params q = to_params(p);
        auto* a = new Amg;
        a->max_levels = q.amg_levels;
        a->theta = q.amg_strong;
        a->pre = q.amg_smooth_pre;
        a->post = q.amg_smooth_post;
        a->setup(*static_cast<Csr*>(h));
        return a;
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}
void orc_amg_free(void* a) { delete static_cast<Amg*>(a); }
int orc_amg_levels(void* a) { return (int)static_cast<Amg*>(a)->h.size(); }
// which: 0 A, 1 P, 2 R.  Returns a new Csr handle (copy), or null.
void* orc_amg_level_matrix(void* a, int lv, int which) {
    const auto& L = static_cast<Amg*>(a)->h[lv];
    return new Csr(which == 0 ? L.A : which == 1 ? L.P : L.R);
}
int orc_amg_aggregates(void* a, int lv, int* out) {
    const auto& L = static_cast<Amg*>(a)->h[lv];
    std::memcpy(out, L.agg.data(), sizeof(int) * L.agg.size());
    return (int)L.agg.size();
}
void orc_amg_apply(void* a, const double* r, double* z) { static_cast<Amg*>(a)->apply(r, z); }
int orc_aggregate(void* h, double theta, int* out) {
    auto agg = aggregate(*static_cast<Csr*>(h), theta);
    std::memcpy(out, agg.data(), sizeof(int) * agg.size());
    return agg.empty() ? 0 : *std::max_element(agg.begin(), agg.end()) + 1;
}
int orc_gepp_solve(void* h, const double* rhs, double* x) {
    ORC_TRY({ dense_gepp_solve(*static_cast<Csr*>(h), rhs, x); return 0; });
}
void* orc_transpose(void* h) { return new Csr(transpose(*static_cast<Csr*>(h))); }
void* orc_matmat(void* a, void* b) {
    try {
        return new Csr(mat_mat_product(*static_cast<Csr*>(a), *static_cast<Csr*>(b)));
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

// ILU factor handle: returns L\U as a new Csr handle.
void* orc_ilu_factor(void* h, double drop_tol, int fill) {
    try {
        Ilu f;
        f.drop_tol = drop_tol;
        f.fill_p = fill;
        f.setup(*static_cast<Csr*>(h));
        Csr* LU = new Csr;
        LU->rows = LU->cols = f.n;
        LU->rp = f.rp;
        LU->ci = f.ci;
        LU->v = f.lu;
        return LU;
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

}  // extern "C"
