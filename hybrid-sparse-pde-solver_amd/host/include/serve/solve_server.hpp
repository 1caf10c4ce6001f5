#pragma once
// =============================================================================
// solve_server.hpp — serving loop for independent solve requests (no
// reference counterpart; each request follows the reference's adaptive
// ladder, adaptive_selector.cpp:298-394: initial state from the matrix
// features, EASY -> MODERATE -> HARD on failure, x warm-started).
//
//  * Admission: requests queue in arrival order and are admitted while the
//    sum of their estimated device footprints stays within an HBM budget
//    (a request larger than the budget runs alone).
//  * Concurrency: worker threads, each bound to its own device context
//    (stream + reduction scratch), run admitted requests side by side.
//  * Continuous batching: requests classified EASY (PCG + Jacobi) up to
//    batch_rows_max rows wait at most batch_window_s for company and are
//    solved together in one batched launch (hsps_batch_*); a system the batch
//    does not converge continues alone at MODERATE from the batch's x.
// =============================================================================
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "core/sparse_matrix.hpp"
#include "core/types.hpp"
#include "hsps_dev.h"

namespace hsps {

struct ServeOptions {
    int device = 0;
    int workers = 4;                 ///< concurrent solves, one context and stream each
    double hbm_budget_bytes = 64e9;  ///< admitted requests' estimated device footprint
    int batch_max = 64;              ///< requests per batched launch (1 disables batching)
    Index batch_rows_max = 262144;   ///< EASY requests up to this many rows are batched
    double batch_window_s = 0.002;   ///< longest wait of an EASY request for company
};

struct ServeResult {
    bool converged = false;
    int iterations = 0;                         ///< of the attempt that produced x
    AdaptiveState state = AdaptiveState::EASY;  ///< ladder state that produced x
    bool batched = false;                       ///< the EASY attempt ran inside a batch
    int escalations = 0;
    Real true_residual = 0;  ///< FP64 ||b - A x|| / ||b|| of the returned x
    double queue_s = 0;      ///< submit -> admission
    double latency_s = 0;    ///< submit -> completion
};

class SolveServer {
public:
    using Ticket = long long;
    explicit SolveServer(const ServeOptions& opt = {});
    ~SolveServer();  ///< finishes every submitted request, then stops the workers
    SolveServer(const SolveServer&) = delete;
    SolveServer& operator=(const SolveServer&) = delete;

    /// Queues A x = b; returns at once.  params.precision selects FP64 or
    /// mixed (FP32 inner iterations with FP64 refinement).
    Ticket submit(SparseMatrix A, std::vector<Real> b, const SolverParams& params);
    bool done(Ticket t) const;
    /// Blocks until the request completes, hands back x and forgets the ticket.
    ServeResult wait(Ticket t, std::vector<Real>* x = nullptr);

    /// Device bytes a request may hold at once: the matrix (FP64 + FP32
    /// values, indices), FGMRES(restart) bases in FP64, an AMG / ILU
    /// preconditioner of about twice the matrix, and the work vectors.
    static double estimate_bytes(Index n, long long nnz, const SolverParams& p);

private:
    using Clock = std::chrono::steady_clock;
    struct Request {
        Ticket id = 0;
        SparseMatrix A;
        std::vector<Real> b, x;
        SolverParams p;
        double bytes = 0;
        Clock::time_point t_submit, t_admit, t_pool;
        ServeResult res;
        bool done = false;
    };
    using Req = std::shared_ptr<Request>;

    void worker(int w);
    void start(const Req& r);
    void run_batch(int w, std::vector<Req> batch);
    void run_ladder(const Req& r, AdaptiveState state, std::vector<Real> x0, int escalations, bool batched);
    void complete(const Req& r);
    bool batch_due(Clock::time_point now) const;  // callers hold mu_
    std::vector<Req> take_batch();                // callers hold mu_

    ServeOptions opt_;
    mutable std::mutex mu_;
    std::condition_variable cv_work_, cv_done_;
    std::deque<Req> pending_;  // submitted, not admitted (arrival order)
    std::deque<Req> pool_;     // admitted EASY requests waiting for a batch
    std::unordered_map<Ticket, Req> all_;
    double in_flight_ = 0;
    bool stop_ = false;
    Ticket next_ = 1;
    std::vector<hsps_ctx> ctxs_;
    std::vector<std::thread> threads_;
};

}  // namespace hsps
