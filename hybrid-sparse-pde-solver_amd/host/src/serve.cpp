// =============================================================================
// serve.cpp — SolveServer (serving loop; no reference counterpart) and its C
// ABI (include/hsps_serve.h).
// =============================================================================
#include "serve/solve_server.hpp"

#include <algorithm>
#include <cstring>
#include <stdexcept>

#include "adaptive/adaptive_selector.hpp"
#include "core/device.hpp"
#include "core/vector.hpp"
#include "hsps_serve.h"
#include "solvers/solver_factory.hpp"
#include "utils/logger.hpp"

namespace hsps {

double SolveServer::estimate_bytes(Index n, long long nnz, const SolverParams& p) {
    const double m = std::max(1, std::min<int>(p.restart_size, n));
    const double matrix = 16.0 * (double)nnz + 16.0 * n;  // FP64 + FP32 values, columns, row blocks
    const double krylov = 16.0 * (m + 1) * n;             // FGMRES V and Z, FP64
    const double precond = 2.0 * matrix;                  // AMG hierarchy / ILU factors
    return matrix + krylov + precond + 64.0 * n;          // + b, x, r and scratch vectors
}

SolveServer::SolveServer(const ServeOptions& opt) : opt_(opt) {
    if (opt_.workers < 1) throw std::invalid_argument("SolveServer: workers must be >= 1");
    if (opt_.batch_max < 1) opt_.batch_max = 1;
    try {
        for (int w = 0; w < opt_.workers; ++w) {
            hsps_ctx c = nullptr;
            dev_check(hsps_ctx_create(opt_.device, &c));
            ctxs_.push_back(c);
        }
        for (int w = 0; w < opt_.workers; ++w) threads_.emplace_back([this, w] { worker(w); });
    } catch (...) {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_work_.notify_all();
        for (auto& t : threads_) t.join();
        for (auto c : ctxs_) hsps_ctx_destroy(c);
        throw;
    }
}

SolveServer::~SolveServer() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_work_.notify_all();
    for (auto& t : threads_) t.join();
    for (auto c : ctxs_) hsps_ctx_destroy(c);
}

SolveServer::Ticket SolveServer::submit(SparseMatrix A, std::vector<Real> b, const SolverParams& params) {
    if (A.rows() != A.cols() || (Index)b.size() != A.rows())
        throw std::invalid_argument("SolveServer::submit: A must be square and b of size A.rows()");
    auto r = std::make_shared<Request>();
    r->bytes = estimate_bytes(A.rows(), A.nnz(), params);
    r->A = std::move(A);
    r->b = std::move(b);
    r->p = params;
    r->p.device_graphs = false;  // workers share the device: no stream capture
    r->t_submit = Clock::now();
    Ticket id;
    {
        std::lock_guard<std::mutex> lk(mu_);
        if (stop_) throw std::runtime_error("SolveServer::submit: server is stopping");
        id = r->id = next_++;
        all_[id] = r;
        pending_.push_back(r);
    }
    cv_work_.notify_one();
    return id;
}

bool SolveServer::done(Ticket t) const {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = all_.find(t);
    if (it == all_.end()) throw std::invalid_argument("SolveServer::done: unknown ticket");
    return it->second->done;
}

ServeResult SolveServer::wait(Ticket t, std::vector<Real>* x) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = all_.find(t);
    if (it == all_.end()) throw std::invalid_argument("SolveServer::wait: unknown ticket");
    Req r = it->second;
    cv_done_.wait(lk, [&] { return r->done; });
    all_.erase(t);
    if (x) *x = std::move(r->x);
    return r->res;
}

bool SolveServer::batch_due(Clock::time_point now) const {
    if (pool_.empty()) return false;
    if ((int)pool_.size() >= opt_.batch_max || stop_) return true;
    return now >= pool_.front()->t_pool + std::chrono::duration_cast<Clock::duration>(
                                                  std::chrono::duration<double>(opt_.batch_window_s));
}

// Up to batch_max pooled requests with the oldest one's tolerance, iteration
// cap and precision (one batched launch shares them), in pool order.
std::vector<SolveServer::Req> SolveServer::take_batch() {
    std::vector<Req> out;
    const SolverParams& k = pool_.front()->p;
    for (auto it = pool_.begin(); it != pool_.end() && (int)out.size() < opt_.batch_max;) {
        const SolverParams& q = (*it)->p;
        if (q.tol == k.tol && q.max_iter == k.max_iter && q.precision == k.precision &&
            q.reliable_delta == k.reliable_delta && q.check_every == k.check_every) {
            out.push_back(*it);
            it = pool_.erase(it);
        } else {
            ++it;
        }
    }
    return out;
}

void SolveServer::worker(int w) {
    ThreadDevice bind(ctxs_[w]);
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
        const auto now = Clock::now();
        if (batch_due(now)) {
            auto batch = take_batch();
            lk.unlock();
            run_batch(w, std::move(batch));
            lk.lock();
            continue;
        }
        if (!pending_.empty()) {
            const Req& h = pending_.front();
            if (in_flight_ <= 0 || in_flight_ + h->bytes <= opt_.hbm_budget_bytes) {
                Req r = h;
                pending_.pop_front();
                in_flight_ += r->bytes;
                r->t_admit = now;
                lk.unlock();
                start(r);
                lk.lock();
                continue;
            }
        }
        if (stop_ && pending_.empty() && pool_.empty()) break;
        if (!pool_.empty())
            cv_work_.wait_until(lk, pool_.front()->t_pool + std::chrono::duration_cast<Clock::duration>(
                                                                 std::chrono::duration<double>(opt_.batch_window_s)));
        else
            cv_work_.wait(lk);
    }
}

// Classify (adaptive_selector.cpp:225-233 on the features of :38-220); EASY
// requests small enough go to the batch pool, the others start the ladder.
void SolveServer::start(const Req& r) {
    AdaptiveState state;
    try {
        AdaptiveSelector sel(r->p);
        state = sel.select_initial_state(sel.extract_features(r->A));
    } catch (const std::exception& e) {
        HSPS_LOG_WARN("SolveServer: request ", r->id, " features failed: ", e.what());
        r->res.converged = false;
        complete(r);
        return;
    }
    if (state == AdaptiveState::EASY && opt_.batch_max > 1 && r->A.rows() <= opt_.batch_rows_max) {
        r->A.release_device();
        {
            std::lock_guard<std::mutex> lk(mu_);
            r->t_pool = Clock::now();
            pool_.push_back(r);
        }
        cv_work_.notify_all();
        return;
    }
    run_ladder(r, state, {}, 0, false);
}

void SolveServer::run_batch(int w, std::vector<Req> batch) {
    const int S = (int)batch.size();
    std::vector<int> row_off(S + 1, 0);
    for (int k = 0; k < S; ++k) row_off[k + 1] = row_off[k] + batch[k]->A.rows();
    const int N = row_off[S];
    std::vector<int> rp(1, 0), ci;
    std::vector<double> v, b;
    for (int k = 0; k < S; ++k) {
        const SparseMatrix& A = batch[k]->A;
        const int base = rp.back();
        for (Index i = 0; i < A.rows(); ++i) rp.push_back(base + A.row_ptr()[i + 1]);
        for (Index c : A.col_idx()) ci.push_back(c + row_off[k]);
        v.insert(v.end(), A.values().begin(), A.values().end());
        b.insert(b.end(), batch[k]->b.begin(), batch[k]->b.end());
    }
    std::vector<int> it(S, 0), cv(S, 0);
    std::vector<double> tr(S, 0.0), x(N, 0.0);
    hsps_ctx c = ctxs_[w];
    hsps_batch B = nullptr;
    void *db = nullptr, *dx = nullptr;
    bool ok = true;
    try {
        dev_check(hsps_batch_create(c, S, row_off.data(), rp.data(), ci.data(), v.data(), &B));
        dev_check(hsps_malloc(c, sizeof(double) * std::max(N, 1), &db));
        dev_check(hsps_malloc(c, sizeof(double) * std::max(N, 1), &dx));
        dev_check(hsps_memcpy_h2d(c, db, b.data(), sizeof(double) * N));
        dev_check(hsps_memset(c, dx, 0, sizeof(double) * N));
        const SolverParams& p = batch[0]->p;
        hsps_params hp;
        hsps_params_default(&hp);
        hp.tol = p.tol;
        hp.max_iter = p.max_iter;
        hp.precision = p.precision == Precision::MIXED ? HSPS_MIXED : HSPS_FP64;
        hp.reliable_delta = p.reliable_delta;
        hp.check_every = p.check_every;
        hp.batch_tile_rows = 2048;  // a one-shot batch: no tile autotuning
        hp.batch_graphs = 0;        // other workers use the device concurrently: no stream capture
        dev_check(hsps_batch_solve(B, (const double*)db, (double*)dx, &hp, it.data(), cv.data(), tr.data()));
        dev_check(hsps_memcpy_d2h(c, x.data(), dx, sizeof(double) * N));
    } catch (const std::exception& e) {
        HSPS_LOG_WARN("SolveServer: batched launch of ", S, " requests failed: ", e.what());
        ok = false;
    }
    if (db) hsps_free(c, db);
    if (dx) hsps_free(c, dx);
    if (B) hsps_batch_destroy(B);
    for (int k = 0; k < S; ++k) {
        const Req& r = batch[k];
        if (!ok) {  // the batch could not run: each request takes the ladder alone
            run_ladder(r, AdaptiveState::EASY, {}, 0, false);
            continue;
        }
        std::vector<Real> xk(x.begin() + row_off[k], x.begin() + row_off[k + 1]);
        if (cv[k]) {
            r->res.converged = true;
            r->res.iterations = it[k];
            r->res.state = AdaptiveState::EASY;
            r->res.batched = true;
            r->res.true_residual = tr[k];
            r->x = std::move(xk);
            complete(r);
        } else {  // escalate as the ladder does, warm-started from the batch's x
            run_ladder(r, AdaptiveState::MODERATE, std::move(xk), 1, true);
        }
    }
}

void SolveServer::run_ladder(const Req& r, AdaptiveState state, std::vector<Real> x0, int escalations,
                             bool batched) {
    try {
        const Index n = r->A.rows();
        Vector b(r->b);
        Vector x = x0.empty() ? Vector(n, REAL_ZERO) : Vector(x0);
        for (;;) {
            double setup = 0.0;
            auto pair = SolverFactory::make_for_state(state, r->A, r->p, &setup);
            SolverStats st;
            const bool ok = pair.first->solve(r->A, b, x, st);
            r->res.converged = ok;
            r->res.iterations = st.iterations;
            r->res.true_residual = st.true_residual;
            r->res.state = state;
            if (ok || state == AdaptiveState::HARD) break;
            state = state == AdaptiveState::EASY ? AdaptiveState::MODERATE : AdaptiveState::HARD;
            ++escalations;
        }
        r->res.escalations = escalations;
        r->res.batched = batched;
        const Real* xd = x.data();
        r->x.assign(xd, xd + n);
    } catch (const std::exception& e) {
        HSPS_LOG_WARN("SolveServer: request ", r->id, " failed at state ", (int)state, ": ", e.what());
        r->res.converged = false;
    }
    r->A.release_device();
    complete(r);
}

void SolveServer::complete(const Req& r) {
    {
        std::lock_guard<std::mutex> lk(mu_);
        const auto now = Clock::now();
        r->res.queue_s = std::chrono::duration<double>(r->t_admit - r->t_submit).count();
        r->res.latency_s = std::chrono::duration<double>(now - r->t_submit).count();
        in_flight_ -= r->bytes;
        r->done = true;
    }
    cv_done_.notify_all();
    cv_work_.notify_all();  // budget freed: admission may proceed
}

}  // namespace hsps

// ---- C ABI ------------------------------------------------------------------------------
using namespace hsps;

struct hsps_server_s {
    SolveServer srv;
    explicit hsps_server_s(const ServeOptions& o) : srv(o) {}
};

static thread_local std::string g_serve_error;

template <class F>
static int serve_guarded(F&& f) {
    try {
        return f();
    } catch (const std::invalid_argument& e) {
        g_serve_error = e.what();
        return HSPS_ERR_INVALID;
    } catch (const std::exception& e) {
        g_serve_error = e.what();
        return HSPS_ERR_SERVE;
    }
}

extern "C" {

const char* hsps_server_last_error(void) { return g_serve_error.c_str(); }

int hsps_server_create(int device, int workers, double hbm_budget_bytes, int batch_max, int batch_rows_max,
                       double batch_window_s, hsps_server* out) {
    return serve_guarded([&] {
        if (!out) throw std::invalid_argument("hsps_server_create: null out");
        ServeOptions o;
        o.device = device;
        o.workers = workers;
        o.hbm_budget_bytes = hbm_budget_bytes;
        o.batch_max = batch_max;
        o.batch_rows_max = batch_rows_max;
        o.batch_window_s = batch_window_s;
        *out = new hsps_server_s(o);
        return HSPS_OK;
    });
}

int hsps_server_destroy(hsps_server S) {
    return serve_guarded([&] {
        delete S;
        return HSPS_OK;
    });
}

int hsps_server_submit(hsps_server S, int n, const int* row_ptr, const int* col_idx, const double* values,
                       const double* b, const hsps_params* p, long long* ticket) {
    return serve_guarded([&] {
        if (!S || !ticket || n < 0 || !row_ptr || (n > 0 && !b))
            throw std::invalid_argument("hsps_server_submit: bad argument");
        const int nnz = row_ptr[n];
        if (row_ptr[0] != 0 || nnz < 0 || (nnz > 0 && (!col_idx || !values)))
            throw std::invalid_argument("hsps_server_submit: bad row_ptr");
        for (int i = 0; i < n; ++i)
            if (row_ptr[i + 1] < row_ptr[i]) throw std::invalid_argument("hsps_server_submit: row_ptr not monotone");
        for (int k = 0; k < nnz; ++k)
            if (col_idx[k] < 0 || col_idx[k] >= n)
                throw std::invalid_argument("hsps_server_submit: column index out of range");
        SparseMatrix A(n, n, std::vector<Index>(row_ptr, row_ptr + n + 1), std::vector<Index>(col_idx, col_idx + nnz),
                       std::vector<Real>(values, values + nnz));
        hsps_params q;
        if (p) q = *p;
        else hsps_params_default(&q);
        SolverParams sp;
        sp.tol = q.tol;
        sp.max_iter = q.max_iter;
        sp.restart_size = q.restart_size;
        sp.drop_tol = q.drop_tol;
        sp.amg_levels = q.amg_levels;
        sp.amg_strong = q.amg_strong;
        sp.amg_smooth_pre = q.amg_smooth_pre;
        sp.amg_smooth_post = q.amg_smooth_post;
        sp.precision = q.precision == HSPS_MIXED ? Precision::MIXED : Precision::FP64;
        sp.reliable_delta = q.reliable_delta;
        sp.check_every = q.check_every;
        *ticket = S->srv.submit(std::move(A), std::vector<Real>(b, b + n), sp);
        return HSPS_OK;
    });
}

int hsps_server_poll(hsps_server S, long long ticket, int* done) {
    return serve_guarded([&] {
        if (!S || !done) throw std::invalid_argument("hsps_server_poll: bad argument");
        *done = S->srv.done(ticket) ? 1 : 0;
        return HSPS_OK;
    });
}

int hsps_server_wait(hsps_server S, long long ticket, double* x, hsps_serve_result* res) {
    return serve_guarded([&] {
        if (!S) throw std::invalid_argument("hsps_server_wait: null server");
        std::vector<Real> xv;
        const ServeResult r = S->srv.wait(ticket, x ? &xv : nullptr);
        if (x && !xv.empty()) std::memcpy(x, xv.data(), sizeof(double) * xv.size());
        if (res) {
            res->converged = r.converged ? 1 : 0;
            res->iterations = r.iterations;
            res->state = (int)r.state;
            res->batched = r.batched ? 1 : 0;
            res->escalations = r.escalations;
            res->true_residual = r.true_residual;
            res->queue_s = r.queue_s;
            res->latency_s = r.latency_s;
        }
        return r.converged ? 1 : 0;
    });
}

}  // extern "C"
