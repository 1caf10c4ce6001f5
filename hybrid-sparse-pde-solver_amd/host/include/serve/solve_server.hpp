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
//    (stream + reduction scratch), classify admitted requests and run the
//    ladder side by side.
//  * Continuous batching: requests classified EASY (PCG + Jacobi) up to
//    batch_rows_max rows join a running session (hsps_session_*) owned by
//    one batch thread on its own context: they enter at the next chunk
//    boundary and leave when they end, so a request never waits for a batch
//    to form or to finish.  Sessions replay their iteration chunks as HIP
//    graphs and tile diagonal-stored systems at the autotuned size.  A system
//    the session does not converge continues at MODERATE on a worker, warm
//    started from the session's x through device memory.
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
    int workers = 4;                 ///< concurrent ladder solves, one context and stream each
    double hbm_budget_bytes = 64e9;  ///< admitted requests' estimated device footprint
    int batch_max = 64;              ///< systems a session holds at once (1 disables sessions)
    Index batch_rows_max = 262144;   ///< EASY requests up to this many rows join a session
    long long batch_rows_total = 0;  ///< rows a session holds (0: min(batch_max x batch_rows_max, 2^24));
                                     ///< nonzeros 8 x that
};

struct ServeResult {
    bool converged = false;
    int iterations = 0;                         ///< of the attempt that produced x
    AdaptiveState state = AdaptiveState::EASY;  ///< ladder state that produced x
    bool batched = false;                       ///< the EASY attempt ran inside a session
    int escalations = 0;
    Real true_residual = 0;  ///< FP64 ||b - A x|| / ||b|| of the returned x
    double queue_s = 0;      ///< submit -> admission
    double latency_s = 0;    ///< submit -> completion
    double submit_s = 0;     ///< submission time, CLOCK_MONOTONIC seconds
    double done_s = 0;       ///< completion time, CLOCK_MONOTONIC seconds
    double setup_s = 0;      ///< preconditioner setups of all attempts (host and device)
    double solve_s = 0;      ///< Krylov solves of all attempts (a session attempt: join to end)
    double classify_s = 0;   ///< admission -> ladder state chosen (features: device upload of A, Lanczos)
};

class SolveServer {
public:
    using Ticket = long long;
    explicit SolveServer(const ServeOptions& opt = {});
    ~SolveServer();  ///< finishes every submitted request, then stops the threads
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
        Clock::time_point t_submit, t_admit, t_join;
        ServeResult res;
        bool done = false;
    };
    using Req = std::shared_ptr<Request>;
    // A request handed to a worker's ladder at a given state, with its x0 on
    // the device (allocated on the batch context) or none.
    struct Handoff {
        Req r;
        AdaptiveState state = AdaptiveState::EASY;
        int escalations = 0;
        bool batched = false;
        void* x0 = nullptr;  // from x0_pool_, returned there once copied
        size_t x0_bytes = 0;
    };
    // One session per distinct (tol, max_iter, precision, reliable_delta,
    // check_every): the parameters its systems share.
    struct Sess {
        hsps_session h = nullptr;
        SolverParams p;
        std::unordered_map<int, Req> slots;
    };

    void worker(int w);
    void batch_loop();
    void start(const Req& r);
    void run_ladder(const Handoff& h);
    void complete(const Req& r);
    bool joins_session(const Req& r) const;
    bool join(std::vector<Sess>& sessions, const Req& r);  // false: no room now
    void step(Sess& S);

    ServeOptions opt_;
    long long sess_rows_ = 0, sess_nnz_ = 0;
    mutable std::mutex mu_;
    std::condition_variable cv_work_, cv_batch_, cv_done_;
    std::deque<Req> pending_;       // submitted, not admitted (arrival order)
    std::deque<Req> pool_;          // admitted EASY requests for the sessions
    std::deque<Handoff> handoffs_;  // ladder work for the workers
    // device buffers on the batch context that carry x0 to the workers, idle
    // ones here; the batch thread allocates them and they are freed at
    // shutdown, so no worker ever synchronises the batch thread's stream
    std::vector<std::pair<void*, size_t>> x0_pool_;
    std::unordered_map<Ticket, Req> all_;
    double in_flight_ = 0;
    int starting_ = 0;  // admitted requests being classified
    bool stop_ = false, batch_done_ = false;
    Ticket next_ = 1;
    std::vector<hsps_ctx> ctxs_;  // one per worker, then the batch thread's
    std::vector<std::thread> threads_;
};

}  // namespace hsps
