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

namespace {

double monotonic_s(std::chrono::steady_clock::time_point t) {
    return std::chrono::duration<double>(t.time_since_epoch()).count();
}

// Requests share a session when the parameters of its batched PCG agree.
bool same_session_params(const SolverParams& a, const SolverParams& b) {
    return a.tol == b.tol && a.max_iter == b.max_iter && a.precision == b.precision &&
           a.reliable_delta == b.reliable_delta && a.check_every == b.check_every;
}

hsps_params session_params(const SolverParams& p) {
    hsps_params hp;
    hsps_params_default(&hp);
    hp.tol = p.tol;
    hp.max_iter = p.max_iter;
    hp.precision = p.precision == Precision::MIXED ? HSPS_MIXED : HSPS_FP64;
    hp.reliable_delta = p.reliable_delta;
    hp.check_every = p.check_every;
    hp.batch_tile_rows = 0;  // the autotuned tile rows (chosen once per process and device)
    hp.batch_graphs = 1;     // the session thread alone captures, on its own stream
    return hp;
}

constexpr int kMaxSessions = 2;  // distinct parameter sets served by sessions at once

}  // namespace

SolveServer::SolveServer(const ServeOptions& opt) : opt_(opt) {
    if (opt_.workers < 1) throw std::invalid_argument("SolveServer: workers must be >= 1");
    if (opt_.batch_max < 1) opt_.batch_max = 1;
    const long long want = opt_.batch_rows_total > 0
                               ? opt_.batch_rows_total
                               : std::min<long long>((long long)opt_.batch_max * opt_.batch_rows_max, 1LL << 24);
    sess_rows_ = std::max<long long>(want, (long long)opt_.batch_rows_max + 64);
    sess_nnz_ = std::min<long long>(8 * sess_rows_, (1LL << 31) - 1);
    try {
        for (int w = 0; w <= opt_.workers; ++w) {  // the last context is the batch thread's
            hsps_ctx c = nullptr;
            dev_check(hsps_ctx_create(opt_.device, &c));
            ctxs_.push_back(c);
        }
        for (int w = 0; w < opt_.workers; ++w) threads_.emplace_back([this, w] { worker(w); });
        threads_.emplace_back([this] { batch_loop(); });
    } catch (...) {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
            batch_done_ = true;
        }
        cv_work_.notify_all();
        cv_batch_.notify_all();
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
    cv_batch_.notify_all();
    for (auto& t : threads_) t.join();
    for (auto& q : x0_pool_) hsps_free(ctxs_.back(), q.first);
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
    r->p.device_graphs = false;  // ladder solves share the device: no stream capture
    r->t_submit = Clock::now();
    r->res.submit_s = monotonic_s(r->t_submit);
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

// Workers: ladder work first (escalations from the sessions, requests the
// sessions cannot take), then admission in arrival order within the budget.
void SolveServer::worker(int w) {
    ThreadDevice bind(ctxs_[w]);
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
        if (!handoffs_.empty()) {
            Handoff h = std::move(handoffs_.front());
            handoffs_.pop_front();
            lk.unlock();
            run_ladder(h);
            lk.lock();
            continue;
        }
        if (!pending_.empty()) {
            const Req& h = pending_.front();
            if (in_flight_ <= 0 || in_flight_ + h->bytes <= opt_.hbm_budget_bytes) {
                Req r = h;
                pending_.pop_front();
                in_flight_ += r->bytes;
                r->t_admit = Clock::now();
                ++starting_;
                lk.unlock();
                start(r);
                lk.lock();
                --starting_;
                cv_batch_.notify_one();
                continue;
            }
        }
        if (stop_ && pending_.empty() && handoffs_.empty() && batch_done_) break;
        cv_work_.wait(lk);
    }
}

bool SolveServer::joins_session(const Req& r) const {
    const long long n = r->A.rows(), nnz = r->A.nnz();
    return opt_.batch_max > 1 && n >= 1 && n <= opt_.batch_rows_max && n + 64 <= sess_rows_ && nnz >= 1 &&
           nnz <= sess_nnz_;
}

// Classify (adaptive_selector.cpp:225-233 on the features of :38-220); EASY
// requests small enough join a session, the others start the ladder.
void SolveServer::start(const Req& r) {
    AdaptiveState state;
    try {
        AdaptiveSelector sel(r->p);
        state = sel.select_initial_state(sel.extract_state_features(r->A));
        r->res.classify_s = std::chrono::duration<double>(Clock::now() - r->t_admit).count();
    } catch (const std::exception& e) {
        HSPS_LOG_WARN("SolveServer: request ", r->id, " features failed: ", e.what());
        r->res.converged = false;
        complete(r);
        return;
    }
    if (state == AdaptiveState::EASY && joins_session(r)) {
        r->A.release_device();  // the session holds its own copy
        {
            std::lock_guard<std::mutex> lk(mu_);
            pool_.push_back(r);
        }
        cv_batch_.notify_one();
        return;
    }
    Handoff h;
    h.r = r;
    h.state = state;
    run_ladder(h);
}

// The batch thread: joins pooled requests to the sessions (first fit, in
// arrival order) and steps every session that holds systems, one chunk each.
void SolveServer::batch_loop() {
    ThreadDevice bind(ctxs_.back());
    std::vector<Sess> sessions;
    std::deque<Req> waiting;  // taken from the pool, no room yet
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
        while (!pool_.empty()) {
            waiting.push_back(std::move(pool_.front()));
            pool_.pop_front();
        }
        bool live = false;
        for (const Sess& S : sessions) live = live || !S.slots.empty();
        if (waiting.empty() && !live) {
            if (stop_ && pending_.empty() && starting_ == 0) break;
            cv_batch_.wait(lk);
            continue;
        }
        lk.unlock();
        for (auto it = waiting.begin(); it != waiting.end();) it = join(sessions, *it) ? waiting.erase(it) : it + 1;
        live = false;
        for (const Sess& S : sessions) live = live || !S.slots.empty();
        if (!live) {  // nothing running and still no room: those requests run alone
            for (Req& r : waiting) {
                Handoff h;
                h.r = r;
                std::lock_guard<std::mutex> g(mu_);
                handoffs_.push_back(std::move(h));
            }
            waiting.clear();
            cv_work_.notify_all();
        }
        for (Sess& S : sessions)
            if (!S.slots.empty()) step(S);
        lk.lock();
    }
    lk.unlock();
    for (Sess& S : sessions) hsps_session_destroy(S.h);
    {
        std::lock_guard<std::mutex> g(mu_);
        batch_done_ = true;
    }
    cv_work_.notify_all();
}

bool SolveServer::join(std::vector<Sess>& sessions, const Req& r) {
    auto to_ladder = [&](const char* why) {
        if (why) HSPS_LOG_WARN("SolveServer: request ", r->id, " runs alone: ", why);
        Handoff h;
        h.r = r;
        {
            std::lock_guard<std::mutex> g(mu_);
            handoffs_.push_back(std::move(h));
        }
        cv_work_.notify_one();
        return true;
    };
    Sess* S = nullptr;
    for (Sess& s : sessions)
        if (same_session_params(s.p, r->p)) S = &s;
    if (!S) {
        if ((int)sessions.size() >= kMaxSessions) {
            auto idle = std::find_if(sessions.begin(), sessions.end(), [](const Sess& s) { return s.slots.empty(); });
            if (idle == sessions.end()) return to_ladder(nullptr);  // every session busy with other parameters
            hsps_session_destroy(idle->h);
            sessions.erase(idle);
        }
        const hsps_params hp = session_params(r->p);
        hsps_session h = nullptr;
        if (hsps_session_create(ctxs_.back(), opt_.batch_max, sess_rows_, sess_nnz_, &hp, &h) < 0)
            return to_ladder(hsps_last_error());
        sessions.push_back(Sess{h, r->p, {}});
        S = &sessions.back();
    }
    const SparseMatrix& A = r->A;
    int slot = -1;
    const int rc = hsps_session_add(S->h, A.rows(), A.row_ptr().data(), A.col_idx().data(), A.values().data(),
                                    r->b.data(), nullptr, &slot);
    if (rc == HSPS_OK) {
        S->slots[slot] = r;
        r->t_join = Clock::now();
        return true;
    }
    if (rc == HSPS_ERR_NOMEM && !S->slots.empty()) return false;  // room frees as systems end
    // an error the ladder reports the reference's way (e.g. a zero diagonal)
    return to_ladder(rc == HSPS_ERR_NOMEM ? hsps_last_error() : nullptr);
}

// One chunk of a session; the systems that ended complete (converged) or
// continue at MODERATE on a worker from their x, handed over in device memory.
void SolveServer::step(Sess& S) {
    const int cap = opt_.batch_max;
    std::vector<int> done(cap), conv(cap);
    int nd = 0;
    if (hsps_session_step(S.h, done.data(), conv.data(), cap, &nd) < 0) {
        HSPS_LOG_WARN("SolveServer: session step failed: ", hsps_last_error());
        std::lock_guard<std::mutex> g(mu_);
        for (auto& kv : S.slots) {  // every system of the session runs alone
            Handoff h;
            h.r = kv.second;
            handoffs_.push_back(std::move(h));
        }
        S.slots.clear();
        cv_work_.notify_all();
        return;
    }
    for (int i = 0; i < nd; ++i) {
        auto it = S.slots.find(done[i]);
        if (it == S.slots.end()) continue;
        Req r = it->second;
        S.slots.erase(it);
        r->res.solve_s += std::chrono::duration<double>(Clock::now() - r->t_join).count();
        const Index n = r->A.rows();
        int iters = 0, cv = 0;
        double tr = 0.0;
        if (conv[i]) {
            r->x.assign((size_t)n, 0.0);
            if (hsps_session_result(S.h, done[i], r->x.data(), nullptr, &iters, &cv, &tr) < 0) {
                HSPS_LOG_WARN("SolveServer: request ", r->id, " result failed: ", hsps_last_error());
                cv = 0;
            }
            r->res.converged = cv != 0;
            r->res.iterations = iters;
            r->res.state = AdaptiveState::EASY;
            r->res.batched = true;
            r->res.true_residual = tr;
            complete(r);
            continue;
        }
        // escalate as the ladder does (adaptive_selector.cpp:350-380), warm
        // started from the session's x without a round trip through the host
        Handoff h;
        h.r = r;
        h.state = AdaptiveState::MODERATE;
        h.escalations = 1;
        h.batched = true;
        const size_t bytes = sizeof(double) * (size_t)n;
        {
            std::lock_guard<std::mutex> g(mu_);
            auto best = x0_pool_.end();  // the smallest idle buffer that holds x
            for (auto q = x0_pool_.begin(); q != x0_pool_.end(); ++q)
                if (q->second >= bytes && (best == x0_pool_.end() || q->second < best->second)) best = q;
            if (best != x0_pool_.end()) {
                h.x0 = best->first;
                h.x0_bytes = best->second;
                x0_pool_.erase(best);
            }
        }
        if (!h.x0) {
            if (hsps_malloc(ctxs_.back(), bytes, &h.x0) < 0) h.x0 = nullptr;
            h.x0_bytes = h.x0 ? bytes : 0;
        }
        if (hsps_session_result(S.h, done[i], h.x0 ? nullptr : (r->x.assign((size_t)n, 0.0), r->x.data()),
                                (double*)h.x0, &iters, &cv, &tr) < 0)
            HSPS_LOG_WARN("SolveServer: request ", r->id, " result failed: ", hsps_last_error());
        {
            std::lock_guard<std::mutex> g(mu_);
            handoffs_.push_back(std::move(h));
        }
        cv_work_.notify_one();
    }
}

void SolveServer::run_ladder(const Handoff& h) {
    const Req& r = h.r;
    AdaptiveState state = h.state;
    int escalations = h.escalations;
    void* x0 = h.x0;  // device x0 from a session, returned to the pool once copied
    auto give_back = [&] {
        std::lock_guard<std::mutex> g(mu_);
        x0_pool_.push_back({x0, h.x0_bytes});
        x0 = nullptr;
    };
    try {
        const Index n = r->A.rows();
        Vector b(r->b);
        Vector x(n, REAL_ZERO);
        if (x0) {  // copied on this worker's stream, which is drained before the buffer is reused
            x.assign_device(static_cast<const double*>(x0));
            dev_check(hsps_ctx_sync(DeviceContext::current()));
            give_back();
        } else if (!r->x.empty()) {
            x = Vector(r->x);
        }
        for (;;) {
            double setup = 0.0;
            auto pair = SolverFactory::make_for_state(state, r->A, r->p, &setup);
            SolverStats st;
            const bool ok = pair.first->solve(r->A, b, x, st);
            r->res.setup_s += setup;
            r->res.solve_s += st.solve_time_s;
            r->res.converged = ok;
            r->res.iterations = st.iterations;
            r->res.true_residual = st.true_residual;
            r->res.state = state;
            if (ok || state == AdaptiveState::HARD) break;
            state = state == AdaptiveState::EASY ? AdaptiveState::MODERATE : AdaptiveState::HARD;
            ++escalations;
        }
        r->res.escalations = escalations;
        r->res.batched = h.batched;
        const Real* xd = x.data();
        r->x.assign(xd, xd + n);
    } catch (const std::exception& e) {
        HSPS_LOG_WARN("SolveServer: request ", r->id, " failed at state ", (int)state, ": ", e.what());
        r->res.converged = false;
        if (x0 && hsps_ctx_sync(DeviceContext::current()) >= 0) give_back();  // else leave it: a copy may still read it
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
        r->res.done_s = monotonic_s(now);
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
                       long long batch_rows_total, hsps_server* out) {
    return serve_guarded([&] {
        if (!out) throw std::invalid_argument("hsps_server_create: null out");
        ServeOptions o;
        o.device = device;
        o.workers = workers;
        o.hbm_budget_bytes = hbm_budget_bytes;
        o.batch_max = batch_max;
        o.batch_rows_max = batch_rows_max;
        o.batch_rows_total = batch_rows_total;
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
            res->submit_s = r.submit_s;
            res->done_s = r.done_s;
            res->setup_s = r.setup_s;
            res->solve_s = r.solve_s;
            res->classify_s = r.classify_s;
        }
        return r.converged ? 1 : 0;
    });
}

}  // extern "C"
