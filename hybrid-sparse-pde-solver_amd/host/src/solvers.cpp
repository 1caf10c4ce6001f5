// =============================================================================
// solvers.cpp — preconditioners, CG / FGMRES and SolverFactory of the host
// layer: each call is a thin wrapper over the C ABI (hsps_dev.h).
// =============================================================================
#include <algorithm>
#include <cctype>
#include <chrono>
#include <cmath>
#include <iomanip>
#include <iostream>

#include "core/device.hpp"
#include "energy/energy_monitor.hpp"
#include "preconditioners/amg.hpp"
#include "preconditioners/ilu.hpp"
#include "preconditioners/jacobi.hpp"
#include "solvers/cg_solver.hpp"
#include "solvers/fgmres_solver.hpp"
#include "solvers/pipelined_fgmres_solver.hpp"
#include "solvers/sstep_cg_solver.hpp"
#include "solvers/solver_factory.hpp"
#include "utils/logger.hpp"

namespace hsps {

hsps_params to_device_params(const SolverParams& p) {
    hsps_params d;
    hsps_params_default(&d);
    d.tol = p.tol;
    d.max_iter = p.max_iter;
    d.restart_size = p.restart_size;
    d.drop_tol = p.drop_tol;
    d.fill_per_row = p.fill_per_row;
    d.amg_levels = p.amg_levels;
    d.amg_strong = p.amg_strong;
    d.amg_smooth_pre = p.amg_smooth_pre;
    d.amg_smooth_post = p.amg_smooth_post;
    d.precision = p.precision == Precision::MIXED ? HSPS_MIXED : HSPS_FP64;
    d.reliable_delta = p.reliable_delta;
    d.check_every = p.check_every;
    d.batch_graphs = p.device_graphs ? 1 : 0;
    d.fgmres_inner_rel = p.fgmres_inner_rel;
    d.sstep_k = p.sstep_k;
    d.sstep_cond_tol = p.sstep_cond_tol;
    return d;
}

// ---- preconditioners ------------------------------------------------------------------
PreconditionerBase::~PreconditionerBase() {
    if (dev_) hsps_precond_destroy(dev_);
    if (user_dev_) hsps_precond_destroy(user_dev_);
}

hsps_precond PreconditionerBase::solver_object(Index n) const {
    if (dev_) return dev_;
    const hsps_ctx c = DeviceContext::current();
    if (!user_dev_ || user_n_ != n || user_ctx_ != c) {
        if (user_dev_) hsps_precond_destroy(user_dev_);
        user_dev_ = nullptr;
        dev_check(hsps_precond_create_user(c, n, &PreconditionerBase::user_apply, const_cast<PreconditionerBase*>(this),
                                           &user_dev_));
        user_n_ = n;
        user_ctx_ = c;
    }
    return user_dev_;
}

// The callback the device Krylov loops call for a caller-written subclass:
// its device_apply(), or its host apply() on host copies of r and z (in
// stream order: the copies synchronise the context's stream).  Exceptions
// become a status code that stops the solve.
int PreconditionerBase::user_apply(void* self, const double* d_r, double* d_z, void* hip_stream) {
    const auto* P = static_cast<const PreconditionerBase*>(self);
    try {
        const Index n = P->user_n_;
        if (P->device_apply(d_r, d_z, n, hip_stream)) return HSPS_OK;
        Vector r(n), z(n);
        dev_check(hsps_memcpy_d2h(P->user_ctx_, r.data(), d_r, sizeof(double) * (size_t)n));
        P->apply(r, z);
        if (z.size() != n) throw std::runtime_error("PreconditionerBase::apply: z has the wrong size");
        const Vector& zc = z;
        dev_check(hsps_memcpy_h2d(P->user_ctx_, d_z, zc.data(), sizeof(double) * (size_t)n));
        return HSPS_OK;
    } catch (const std::exception& e) {
        HSPS_LOG_WARN("user preconditioner apply failed: ", e.what());
        return HSPS_ERR_INVALID;
    }
}

double PreconditionerBase::setup_device(const SparseMatrix& A, int kind, const hsps_params& p) {
    if (dev_) {
        hsps_precond_destroy(dev_);
        dev_ = nullptr;
    }
    A_ = A;  // shares the device handle; keeps the operator alive for the precond
    double t = 0.0;
    dev_check(hsps_precond_create(A_.device(), kind, &p, &dev_, &t));
    ready_ = true;
    return t;
}

void PreconditionerBase::apply(const Vector& r, Vector& z) const {
    if (!dev_) throw std::runtime_error("preconditioner applied before setup()");
    if (z.size() != r.size()) z.resize(r.size());
    const double* rd = r.device_data();
    dev_check(hsps_precond_apply_f64(dev_, rd, z.device_data_mut()));
}

double JacobiPreconditioner::setup(const SparseMatrix& A) {  // jacobi.cpp:13-38
    hsps_params p;
    hsps_params_default(&p);
    p.jacobi_omega = omega_;
    setup_flops_ = A.rows();
    return setup_device(A, HSPS_PRECOND_JACOBI, p);
}

double ILUPreconditioner::setup(const SparseMatrix& A) {  // ilu.cpp:31-36
    hsps_params p;
    hsps_params_default(&p);
    p.drop_tol = drop_tol_;
    p.fill_per_row = fill_p_;
    return setup_device(A, HSPS_PRECOND_ILU, p);
}

double AMGPreconditioner::setup(const SparseMatrix& A) {  // amg.cpp:183-237
    hsps_params p;
    hsps_params_default(&p);
    p.amg_levels = max_levels_;
    p.amg_strong = strength_thresh_;
    p.amg_smooth_pre = smooth_pre_;
    p.amg_smooth_post = smooth_post_;
    p.amg_omega = omega_;
    return setup_device(A, HSPS_PRECOND_AMG, p);
}
int AMGPreconditioner::levels() const { return dev_ ? hsps_amg_levels(dev_) : 0; }
std::vector<Index> AMGPreconditioner::aggregates(int lv) const {
    auto sz = level_size(lv);
    std::vector<Index> a(sz.first);
    dev_check(hsps_amg_aggregates(dev_, lv, a.data()));
    return a;
}
std::pair<Index, Index> AMGPreconditioner::level_size(int lv) const {
    int r = 0, z = 0;
    dev_check(hsps_amg_level_info(dev_, lv, &r, &z));
    return {r, z};
}

// ---- solvers ------------------------------------------------------------------------------
Real SolverBase::relative_residual(const SparseMatrix& A, const Vector& b, const Vector& x) const {
    Vector r = A * x;  // solver_base.cpp:8-15
    const Real bn = b.norm2();
    r.axpby(1.0, b, -1.0);
    return bn > REAL_EPS ? r.norm2() / bn : r.norm2();
}

bool SolverBase::device_solve(int solver, const SparseMatrix& A, const Vector& b, Vector& x, SolverStats& stats) {
    const Index n = A.rows();
    if (b.size() != n) throw std::runtime_error("solve: b size != rows");
    if (x.size() != n) x.resize(n, REAL_ZERO);  // solver_base.hpp:29-31
    hsps_precond M = precond_ ? precond_->solver_object(n) : nullptr;
    hsps_params p = to_device_params(params_);
    if (solver == kSStepCG) p.sstep_k = sstep_k_;
    hsps_stats s{};
    const double* bd = b.device_data();
    double* xd = x.device_data_mut();
    int rc = 0;
    switch (solver) {
        case kCG: rc = hsps_cg_solve(A.device(), M, bd, xd, &p, &s); break;
        case kFGMRES: rc = hsps_fgmres_solve(A.device(), M, bd, xd, &p, &s); break;
        case kPipelinedFGMRES: rc = hsps_pipelined_fgmres_solve(A.device(), M, bd, xd, &p, &s); break;
        case kSStepCG: rc = hsps_sstep_cg_solve(A.device(), bd, xd, &p, &s); break;
        default: throw std::invalid_argument("solve: unknown solver");
    }
    dev_check(rc);
    stats.solver_used = solver == kCG ? SolverType::CG : solver == kSStepCG ? SolverType::SSTEP_CG : SolverType::FGMRES;
    if (solver == kSStepCG) {
        stats.sstep_k = s.sstep_k;
        stats.comm_bytes = s.comm_events;  // one global reduction per batch (sstep_cg_solver.cpp:229)
    }
    stats.precond_used = precond_ ? precond_->type() : PrecondType::NONE;
    stats.iterations = s.iterations;
    stats.initial_residual = s.initial_residual;
    stats.restarts = s.restarts;
    stats.solve_time_s = s.solve_time_s;
    stats.flop_count = s.flop_count;
    stats.mem_bytes = s.mem_bytes;
    stats.energy_joules = s.energy_joules;
    stats.true_residual = s.true_residual;
    stats.reliable_updates = s.reliable_updates;
    // (Pipelined) FGMRES leaves converged/final_residual untouched on max_iter
    // exhaustion in FP64 mode (fgmres_solver.cpp:197-226); CG and s-step CG
    // always write them.
    const bool write = solver == kCG || solver == kSStepCG || s.converged || params_.precision == Precision::MIXED;
    if (write) {
        stats.converged = s.converged != 0;
        stats.final_residual = s.final_residual;
    }
    if (energy_mon_) energy_mon_->record(s.iterations, s.final_residual, s.flop_count, s.mem_bytes);
    if (params_.verbose)
        std::cout << "  " << to_string(stats.solver_used) << " iters=" << s.iterations << " rel_res=" << std::scientific
                  << s.final_residual << " true=" << s.true_residual << "\n";
    return s.converged != 0;
}

bool CGSolver::solve(const SparseMatrix& A, const Vector& b, Vector& x, SolverStats& stats) {
    return device_solve(kCG, A, b, x, stats);
}
bool FGMRESSolver::solve(const SparseMatrix& A, const Vector& b, Vector& x, SolverStats& stats) {
    return device_solve(kFGMRES, A, b, x, stats);
}
bool PipelinedFGMRESSolver::solve(const SparseMatrix& A, const Vector& b, Vector& x, SolverStats& stats) {
    return device_solve(kPipelinedFGMRES, A, b, x, stats);
}
bool SStepCGSolver::solve(const SparseMatrix& A, const Vector& b, Vector& x, SolverStats& stats) {
    sstep_k_ = k_;
    return device_solve(kSStepCG, A, b, x, stats);
}

// ---- factory (solver_factory.cpp:25-175) --------------------------------------------------
std::unique_ptr<SolverBase> SolverFactory::make_solver(SolverType type, const SolverParams& params) {
    std::unique_ptr<SolverBase> s;
    switch (type) {
        case SolverType::CG:
            s = std::make_unique<CGSolver>();
            break;
        case SolverType::SSTEP_CG:
            s = std::make_unique<SStepCGSolver>(params.sstep_k);
            break;
        case SolverType::FGMRES:
        case SolverType::BICGSTAB:  // the reference falls back to FGMRES (solver_factory.cpp:34-37)
        case SolverType::ASYNC_FGMRES:
            s = std::make_unique<FGMRESSolver>();
            break;
        default:
            throw std::invalid_argument("SolverFactory: unknown SolverType");
    }
    s->set_params(params);
    return s;
}

std::shared_ptr<PreconditionerBase> SolverFactory::make_precond(PrecondType type, const SparseMatrix& A,
                                                                const SolverParams& params, double* setup_time_s) {
    std::shared_ptr<PreconditionerBase> M;
    double t = 0.0;
    switch (type) {
        case PrecondType::NONE:
            if (setup_time_s) *setup_time_s = 0.0;
            return nullptr;
        case PrecondType::JACOBI: {
            auto j = std::make_shared<JacobiPreconditioner>();
            t = j->setup(A);
            M = j;
            break;
        }
        case PrecondType::ILU: {
            auto i = std::make_shared<ILUPreconditioner>();
            i->set_drop_tol(params.drop_tol);  // ILUT(drop_tol) unless 0 (solver_factory.cpp:73-79)
            t = i->setup(A);
            M = i;
            break;
        }
        case PrecondType::AMG: {
            auto a = std::make_shared<AMGPreconditioner>();
            a->set_max_levels(params.amg_levels);
            a->set_strength_thresh(params.amg_strong);
            a->set_smooth_pre(params.amg_smooth_pre);
            a->set_smooth_post(params.amg_smooth_post);
            t = a->setup(A);
            M = a;
            break;
        }
        default:
            throw std::invalid_argument("SolverFactory: unknown PrecondType");
    }
    if (setup_time_s) *setup_time_s = t;
    return M;
}

SolverPrecondPair SolverFactory::make(SolverType st, PrecondType pt, const SparseMatrix& A, const SolverParams& params,
                                      double* setup_time_s) {
    double t = 0.0;
    auto M = make_precond(pt, A, params, &t);
    if (setup_time_s) *setup_time_s = t;
    auto s = make_solver(st, params);
    if (M) s->set_preconditioner(M);
    return {std::move(s), M};
}

SolverPrecondPair SolverFactory::make_for_state(AdaptiveState state, const SparseMatrix& A, const SolverParams& params,
                                                double* setup_time_s) {
    switch (state) {
        case AdaptiveState::EASY: return make(SolverType::CG, PrecondType::JACOBI, A, params, setup_time_s);
        case AdaptiveState::MODERATE: return make(SolverType::FGMRES, PrecondType::ILU, A, params, setup_time_s);
        case AdaptiveState::HARD: return make(SolverType::FGMRES, PrecondType::AMG, A, params, setup_time_s);
    }
    throw std::invalid_argument("SolverFactory: unknown AdaptiveState");
}

static std::string upper(std::string s) {
    std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return std::toupper(c); });
    return s;
}
SolverType SolverFactory::parse_solver(const std::string& s) {
    const auto u = upper(s);
    if (u == "CG") return SolverType::CG;
    if (u == "FGMRES") return SolverType::FGMRES;
    if (u == "BICGSTAB") return SolverType::BICGSTAB;
    if (u == "SSTEP_CG" || u == "SSTEP" || u == "S-STEP") return SolverType::SSTEP_CG;
    if (u == "ASYNC_FGMRES" || u == "ASYNC") return SolverType::ASYNC_FGMRES;
    throw std::invalid_argument("SolverFactory: unknown solver name '" + s + "'");
}
PrecondType SolverFactory::parse_precond(const std::string& s) {
    const auto u = upper(s);
    if (u == "NONE" || u == "IDENTITY") return PrecondType::NONE;
    if (u == "JACOBI" || u == "DIAG") return PrecondType::JACOBI;
    if (u == "ILU" || u == "ILU0") return PrecondType::ILU;
    if (u == "AMG") return PrecondType::AMG;
    throw std::invalid_argument("SolverFactory: unknown precond name '" + s + "'");
}
SolverPrecondPair SolverFactory::make(const std::string& sn, const std::string& pn, const SparseMatrix& A,
                                      const SolverParams& params, double* setup_time_s) {
    return make(parse_solver(sn), parse_precond(pn), A, params, setup_time_s);
}

}  // namespace hsps
