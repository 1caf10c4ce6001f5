/* =============================================================================
 * hsps_serve.h — C ABI of the solve server (lib/libhsps.so).  No reference
 * counterpart: the reference solves one system per call
 * (adaptive_selector.cpp:298-394); this serves a stream of independent
 * requests, each through that adaptive ladder, with HBM-budget admission,
 * concurrent workers (one device context and stream each) and continuous
 * batching: small EASY (PCG + Jacobi) requests join a running session
 * (hsps_session_*, hsps_dev.h) at its next chunk boundary and leave it when
 * they end.
 * Status codes and hsps_params are those of hsps_dev.h.
 * ========================================================================== */
#ifndef HSPS_SERVE_H
#define HSPS_SERVE_H

#include "hsps_dev.h"

#ifdef __cplusplus
extern "C" {
#endif

#define HSPS_ERR_SERVE (-8) /* runtime failure inside the server (std::runtime_error) */

typedef struct hsps_server_s* hsps_server;

typedef struct {
    int converged;        /* 1 if x meets tol in the FP64 true residual */
    int iterations;       /* of the attempt that produced x */
    int state;            /* ladder state of that attempt: 0 EASY, 1 MODERATE, 2 HARD */
    int batched;          /* 1 if the EASY attempt ran inside a session */
    int escalations;      /* ladder steps taken after the first attempt */
    double true_residual; /* ||b - A x|| / ||b|| in FP64 */
    double queue_s;       /* submit -> admission */
    double latency_s;     /* submit -> completion */
    double submit_s;      /* submission time, CLOCK_MONOTONIC seconds */
    double done_s;        /* completion time, CLOCK_MONOTONIC seconds */
    double setup_s;       /* preconditioner setups of all attempts */
    double solve_s;       /* Krylov solves of all attempts (a session attempt: join to end) */
    double classify_s;    /* admission -> ladder state chosen: the features (device upload of A, Lanczos) */
} hsps_serve_result;

/* workers: concurrent ladder solves; hbm_budget_bytes: admitted requests'
 * estimated device footprint; batch_max: systems a session holds at once (1
 * disables sessions); batch_rows_max: largest request a session takes;
 * batch_rows_total: rows a session holds (0 = min(batch_max x batch_rows_max,
 * 2^24)), nonzeros 8 x that.  Sessions tile systems at the size the batch
 * autotuner picks once per process and device, and replay their chunks as
 * HIP graphs. */
int hsps_server_create(int device, int workers, double hbm_budget_bytes, int batch_max, int batch_rows_max,
                       long long batch_rows_total, hsps_server* out);
/* Finishes every submitted request, then stops. */
int hsps_server_destroy(hsps_server S);
/* Copies the n x n CSR system and b; returns at once with a ticket.  p may be
 * NULL (defaults); p->precision selects FP64 or mixed. */
int hsps_server_submit(hsps_server S, int n, const int* h_row_ptr, const int* h_col_idx, const double* h_values,
                       const double* h_b, const hsps_params* p, long long* ticket);
/* *done = 1 once the request has completed. */
int hsps_server_poll(hsps_server S, long long ticket, int* done);
/* Blocks until the request completes; x (n doubles, may be NULL) receives the
 * solution; the ticket is then forgotten.  Returns 1 converged, 0 not. */
int hsps_server_wait(hsps_server S, long long ticket, double* h_x, hsps_serve_result* res);
const char* hsps_server_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* HSPS_SERVE_H */
