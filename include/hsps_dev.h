/*
 * hsps_dev.h — C ABI of the MI355X (gfx950) solve path for HAMLSS.
 *
 * Plain C: opaque handles, plain pointers and sizes, int status codes.  No
 * exception crosses this boundary; on a negative status, hsps_last_error()
 * returns a thread-local message.  The C++17 host layer (hsps::SparseMatrix,
 * SolverBase, PreconditionerBase, AdaptiveSelector) is written on top of it
 * and rethrows the reference's exception types.
 *
 * Memory: vector arguments named d_* are DEVICE pointers (hsps_malloc or any
 * HIP allocation on the context's device).  Matrix creation takes HOST CSR
 * arrays and uploads them.  All work is enqueued on the context's stream;
 * calls that return host scalars synchronise that stream.
 *
 * Each entry point cites the reference interface it replaces (paths relative
 * to the reference repository).
 */
#ifndef HSPS_DEV_H
#define HSPS_DEV_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (errors are negative) -------------------------------- */
enum {
    HSPS_OK = 0,
    HSPS_ERR_INVALID = -1,    /* bad argument / handle (std::invalid_argument) */
    HSPS_ERR_SIZE = -2,       /* size mismatch (vector.cpp:41-42, sparse_matrix.cpp:69-72) */
    HSPS_ERR_ZERO_DIAG = -3,  /* Jacobi zero diagonal (jacobi.cpp:28-30) */
    HSPS_ERR_PIVOT = -4,      /* ILU near-zero pivot (ilu.cpp:65-67, 257-259) */
    HSPS_ERR_HIP = -5,        /* HIP runtime failure */
    HSPS_ERR_NOMEM = -6,      /* device allocation failed / over budget */
    HSPS_ERR_UNSUPPORTED = -7 /* feature not available for this configuration */
};

/* precision of the Krylov inner loop */
enum { HSPS_FP64 = 0, HSPS_MIXED = 1 };
/* preconditioner kinds: PrecondType (types.hpp:40-47) */
enum { HSPS_PRECOND_NONE = 0, HSPS_PRECOND_JACOBI = 1, HSPS_PRECOND_ILU = 2, HSPS_PRECOND_AMG = 3,
       HSPS_PRECOND_USER = 4 /* caller-written (hsps_precond_create_user) */ };
/* solver kinds: SolverType (types.hpp:32-38) */
enum { HSPS_SOLVER_CG = 0, HSPS_SOLVER_FGMRES = 1 };

typedef struct hsps_ctx_s* hsps_ctx;
typedef struct hsps_csr_s* hsps_csr;
typedef struct hsps_precond_s* hsps_precond;
typedef struct hsps_batch_s* hsps_batch;

/* SolverParams, types.hpp:119-143 (fields on the solve path), plus the
 * device-only controls at the end. */
typedef struct hsps_params {
    double tol;            /* relative residual tolerance (1e-8) */
    int max_iter;          /* max total Krylov iterations (2000) */
    int restart_size;      /* FGMRES m (50) */
    double drop_tol;       /* ILUT tau; 0 selects ILU(0) (1e-4) */
    int fill_per_row;      /* ILUT extra fill p (0) */
    int amg_levels;        /* max AMG levels (10) */
    double amg_strong;     /* strength threshold theta (0.25) */
    int amg_smooth_pre;    /* (2) */
    int amg_smooth_post;   /* (2) */
    double amg_omega;      /* Jacobi damping used by SA smoothing and smoother (0.67, amg.hpp:101) */
    double jacobi_omega;   /* Jacobi preconditioner damping (1.0, jacobi.hpp:33) */
    int precision;         /* HSPS_FP64 or HSPS_MIXED */
    double reliable_delta; /* mixed CG: recompute the FP64 true residual when
                              |r| drops below delta * max|r| since the last one (0.1) */
    int check_every;       /* host polls the device convergence flag every k iterations (32) */
    int batch_tile_rows;   /* batched solves: rows per workgroup tile (CSR tiles also stop at 2048
                              nonzeros); 0 = autotune on first solve (0) */
    int batch_graphs;      /* batched solves: replay the iteration chunks as hipGraphs (1); 0 launches
                              them directly, for callers that solve from several threads at once */
    int cg_engine;         /* PCG + Jacobi (omega 1) on one system: 1 = run it as a one-system batch
                              (diagonal / stencil storage, fused kernels) when the matrix is
                              diagonal-stored or small (1); 0 = the plain kernel sequence */
    double fgmres_inner_rel; /* mixed FGMRES: an FP32 cycle also ends once its residual estimate is
                              below fgmres_inner_rel x the cycle's starting residual, and the FP64
                              restart refines (the inner tolerance of GMRES-based iterative
                              refinement; pays off for tolerances FP32 cannot reach in one cycle,
                              e.g. 1e-5 for tol 1e-10); 0 = cycles run to tol (0) */
    int sstep_k;           /* s-step CG batch size k (1, types.hpp:134) */
    double sstep_cond_tol; /* s-step CG: halve k when the Gram diagonal ratio exceeds this (1e12) */
} hsps_params;

/* SolverStats, types.hpp:91-114 (fields on the solve path) plus the FP64
 * true residual the device path always reports. */
typedef struct hsps_stats {
    int iterations;
    int converged;
    int restarts;
    int reliable_updates;  /* mixed CG: FP64 residual replacements */
    double initial_residual;
    double final_residual; /* the solver's own (recursive / Givens) estimate, relative */
    double true_residual;  /* ||b - A x||_2 / ||b||_2 recomputed in FP64 at exit */
    double solve_time_s;
    double setup_time_s;
    long long flop_count;  /* flop_model, energy_monitor.hpp:163-177 */
    long long mem_bytes;
    double energy_joules;  /* 2e-10 J/FLOP + 5e-9 J/B proxy (cg_solver.cpp:186-188) */
    int sstep_k;           /* s-step CG: the batch size used (types.hpp:110) */
    int comm_events;       /* s-step CG: global reductions, one per batch (types.hpp:112) */
} hsps_stats;

/* ---- errors / defaults ------------------------------------------------- */
const char* hsps_last_error(void);
/* Fills the reference defaults of SolverParams (types.hpp:119-143). */
void hsps_params_default(hsps_params* p);

/* ---- context (replaces ParallelContext/backend selection,
 *      parallel_context.cpp:31-50; one per device + stream) --------------- */
int hsps_ctx_create(int device, hsps_ctx* out);
int hsps_ctx_destroy(hsps_ctx ctx);
/* Use an existing hipStream_t (NULL = the context's own stream); the work
 * enqueued on the previous stream is drained first. */
int hsps_ctx_set_stream(hsps_ctx ctx, void* hip_stream);
void* hsps_ctx_stream(hsps_ctx ctx);
int hsps_ctx_sync(hsps_ctx ctx);
/* Bytes of device memory currently held by handles of this context. */
size_t hsps_ctx_bytes_in_use(hsps_ctx ctx);

/* ---- device memory (replaces DeviceVector, cuda_backend.cpp:26-83) ------ */
/* Device memory comes from a per-device cache of freed blocks (size classes
 * 2^k and 3 * 2^(k-1)), because hipFree waits for the work of every stream on
 * the device.  hsps_free drains the context's stream and returns the block to
 * the cache; work on other streams that uses it must be complete by then.
 * The library's own buffers go through the same cache, fenced by an event on
 * the context's stream.  HSPS_MEM_CACHE_MB bounds the cached bytes per device
 * (default: a quarter of its memory; 0 = plain hipMalloc / hipFree). */
int hsps_malloc(hsps_ctx ctx, size_t bytes, void** d_ptr);
int hsps_free(hsps_ctx ctx, void* d_ptr);
int hsps_memcpy_h2d(hsps_ctx ctx, void* d_dst, const void* h_src, size_t bytes);
int hsps_memcpy_d2h(hsps_ctx ctx, void* h_dst, const void* d_src, size_t bytes);
int hsps_memcpy_d2d(hsps_ctx ctx, void* d_dst, const void* d_src, size_t bytes);
int hsps_memset(hsps_ctx ctx, void* d_dst, int value, size_t bytes);
/* Diagnostic (no reference counterpart): the achievable HBM streaming rate
 * against which the HBM-bound kernels are judged.  Times reps launches of
 * y = a r + b p over n FP32 values (two reads and one write per value,
 * 16-byte nontemporal accesses; n rounded down to a multiple of 4096) on the
 * context's stream and returns the median rate in GB/s of moved bytes.  With
 * 12 n bytes well above the 256 MB Infinity Cache it measures HBM. */
int hsps_stream_probe(hsps_ctx ctx, long long n, int reps, double* gbs);

/* ---- CSR matrix (SparseMatrix, sparse_matrix.hpp:34-170; int32 indices,
 *      sorted columns).  Uploads FP64 values, an FP32 copy, the diagonal and
 *      the SpMV row-block schedule. -------------------------------------- */
int hsps_csr_create(hsps_ctx ctx, int rows, int cols, int nnz, const int* h_row_ptr,
                    const int* h_col_idx, const double* h_values, hsps_csr* out);
int hsps_csr_destroy(hsps_csr A);
int hsps_csr_info(hsps_csr A, int* rows, int* cols, int* nnz);
/* Replace the values (same pattern) — re-uploads FP64/FP32 values and diagonal. */
int hsps_csr_update_values(hsps_csr A, const double* h_values);

/* ---- SpMV (SparseMatrix::spmv, sparse_matrix.cpp:68-87;
 *      cusparseSpMV, cuda_backend.cpp:379-411) -------------------------- */
int hsps_spmv_f64(hsps_csr A, const double* d_x, double* d_y);
int hsps_spmv_f32(hsps_csr A, const float* d_x, float* d_y);
/* r = b - A x in FP64 (cg_solver.cpp:52-54, solver_base.cpp:8-15). */
int hsps_residual_f64(hsps_csr A, const double* d_b, const double* d_x, double* d_r);

/* ---- BLAS-1 (Vector, vector.cpp:39-119; cublasD*, cuda_backend.cpp:297-376).
 *      Reductions accumulate in FP64 and return to the host. -------------- */
int hsps_dot_f64(hsps_ctx ctx, int n, const double* d_x, const double* d_y, double* h_out);
int hsps_norm2_f64(hsps_ctx ctx, int n, const double* d_x, double* h_out);
int hsps_axpy_f64(hsps_ctx ctx, int n, double alpha, const double* d_x, double* d_y);   /* y += a x */
int hsps_axpby_f64(hsps_ctx ctx, int n, double a, const double* d_x, double b, double* d_y); /* y = a x + b y */
int hsps_scale_f64(hsps_ctx ctx, int n, double alpha, double* d_x);

/* ---- preconditioners (PreconditionerBase, preconditioner_base.hpp:16-48;
 *      JacobiPreconditioner jacobi.cpp:13-51; ILUPreconditioner ilu.cpp:31-271;
 *      AMGPreconditioner amg.cpp:183-343).  setup_seconds may be NULL. ---- */
int hsps_precond_create(hsps_csr A, int kind, const hsps_params* p, hsps_precond* out,
                        double* setup_seconds);
int hsps_precond_destroy(hsps_precond M);
/* A caller-written preconditioner: the virtual PreconditionerBase::apply
 * (preconditioner_base.hpp:32) that the reference's CG and FGMRES call each
 * iteration (cg_solver.cpp:79, 147; fgmres_solver.cpp:112).  The Krylov loops
 * call apply(user, d_r, d_z, hip_stream) with FP64 device vectors of n
 * entries (mixed precision converts its FP32 vectors around the call) and
 * expect z = M^-1 r, the work enqueued on hip_stream (or finished before the
 * return).  The loops enqueue iterations ahead of their convergence test, so
 * apply may also be called for iterations whose results are then discarded:
 * it must be a function of r alone.  A negative return stops the solve with
 * that status.  The callback runs on the thread that called the solve. */
typedef int (*hsps_user_apply_fn)(void* user, const double* d_r, double* d_z, void* hip_stream);
int hsps_precond_create_user(hsps_ctx ctx, int n, hsps_user_apply_fn apply, void* user, hsps_precond* out);
int hsps_precond_kind(hsps_precond M);
/* z = M^-1 r in FP64 (z is overwritten; AMG starts from z = 0, amg.cpp:340-343). */
int hsps_precond_apply_f64(hsps_precond M, const double* d_r, double* d_z);
/* The ILU factors hsps_precond_create(HSPS_PRECOND_ILU) uploads, computed on
 * the host only (no device call): ILU(0) on A's pattern (ilu.cpp:41-89) when
 * drop_tol == 0 and fill_per_row == 0, else ILUT(drop_tol, fill_per_row)
 * (ilu.cpp:94-220, the same factors as the reference's dense row scan, which
 * visits only the row's active columns) as one L\U CSR (unit L implicit).  *nnz_out receives the
 * factor's nonzeros; lu_rp (n+1 ints) and lu_ci / lu_v (nnz_out entries) are
 * filled where not NULL, so a first call with NULL arrays sizes them. */
int hsps_ilu_factor(int n, const int* h_row_ptr, const int* h_col_idx, const double* h_values, double drop_tol,
                    int fill_per_row, int* nnz_out, int* h_lu_row_ptr, int* h_lu_col_idx, double* h_lu_values);
/* AMG introspection: number of levels; rows/nnz of level lv; the host
 * aggregate map of level lv (length rows(lv); lv < levels-1). */
int hsps_amg_levels(hsps_precond M);
int hsps_amg_level_info(hsps_precond M, int lv, int* rows, int* nnz);
int hsps_amg_aggregates(hsps_precond M, int lv, int* h_out);
/* Copy level lv's operator (which=0 A, 1 P, 2 R) to host CSR arrays: the row
 * pointer has rows+1 entries (A, P: rows(lv); R: rows(lv+1)); with
 * h_col_idx or h_values NULL only the row pointer is written, whose last
 * entry sizes the other two. */
int hsps_amg_level_matrix(hsps_precond M, int lv, int which, int* h_row_ptr, int* h_col_idx,
                          double* h_values);
/* Wall seconds of the AMG setup stages (no reference counterpart; the
 * reference reports only the total): out[0..5] = strength pattern, host
 * aggregation, smoothed P, R = P^T, Galerkin products, coarsest operator. */
int hsps_amg_setup_profile(hsps_precond M, double* h_out6);
/* How an ILU preconditioner's two sweeps run (no reference counterpart;
 * diagnostic): 0 = level-scheduled, 1 = wavefront over a banded grid layout
 * (the factor's dependencies within a band of W rows, short histories),
 * 2 = nothing to sweep (*fwd only: no entry below the diagonal, y = r);
 * *w_fwd / *w_bwd get the band width W (0 otherwise).  All give the same
 * bits. */
int hsps_ilu_sweep_kinds(hsps_precond M, int* fwd, int* bwd, int* w_fwd, int* w_bwd);

/* ---- Krylov solves (SolverBase::solve, solver_base.hpp:32-35;
 *      CGSolver::solve cg_solver.cpp:37-191; FGMRESSolver::solve
 *      fgmres_solver.cpp:165-241).  d_x is the initial guess on entry and the
 *      solution on exit.  M may be NULL (unpreconditioned).  Returns 1 when
 *      converged, 0 when not, <0 on error. ------------------------------ */
int hsps_cg_solve(hsps_csr A, hsps_precond M, const double* d_b, double* d_x, const hsps_params* p,
                  hsps_stats* stats);
int hsps_fgmres_solve(hsps_csr A, hsps_precond M, const double* d_b, double* d_x,
                      const hsps_params* p, hsps_stats* stats);
/* PCG that allocates nothing.  No counterpart in the reference: for callers
 * that cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The same solve as hsps_cg_solve on its
 * plain kernel sequence (p->cg_engine is ignored; the same bits as
 * hsps_cg_solve with cg_engine = 0), with its workspace (r, p, q, z, the FP64
 * residual, the FP32 low-order x, reduction partials and the device state)
 * in one static 64 MiB array per device, reserved when the library loads and
 * held until the process ends; the host reads the state back synchronously
 * into stack memory.  Limit: unpreconditioned / Jacobi PCG needs 32 B per
 * row in FP64 (2.0 M rows) and 24 B in mixed precision (2.7 M rows); other
 * preconditioners add one vector (40 / 28 B per row).  A larger system
 * returns HSPS_ERR_NOMEM before any device work.  Every call on a device
 * shares that device's array: calls that overlap on one device race (run
 * them on one stream, or serialise them). */
int hsps_cg_solve_noalloc(hsps_csr A, hsps_precond M, const double* d_b, double* d_x, const hsps_params* p,
                          hsps_stats* stats);
/* PipelinedFGMRESSolver::solve (pipelined_fgmres_solver.cpp:165-243): the
 * same Krylov space as FGMRES; here each iteration orthogonalises with two
 * fused classical Gram-Schmidt passes (a fixed number of launches and
 * reductions whatever the iteration index), a breakdown ends the cycle, and
 * (its + 1) SpMVs per cycle are charged to flop_count / mem_bytes. */
int hsps_pipelined_fgmres_solve(hsps_csr A, hsps_precond M, const double* d_b, double* d_x,
                                const hsps_params* p, hsps_stats* stats);
/* SStepCGSolver::solve (sstep_cg_solver.cpp:168-305) with k = p->sstep_k:
 * per batch a normalised Newton-Chebyshev basis, its Gram matrix and B^T r in
 * one fused pass, a host Cholesky solve (steepest-descent step when it
 * fails), x += B y and the exact residual.  FP64; the reference applies no
 * preconditioner, so none is taken. */
int hsps_sstep_cg_solve(hsps_csr A, const double* d_b, double* d_x, const hsps_params* p, hsps_stats* stats);
/* SparseMatrix::spectral_interval (sparse_matrix.cpp:451-479): Lanczos from
 * q = 1/sqrt(n) for min(n, steps, 64) steps, Gershgorin bounds of T. */
int hsps_spectral_interval(hsps_csr A, int steps, double* lo, double* hi);

/* ---- GPU energy telemetry through ROCm SMI: the MI355X counterpart of
 *      NVMLMonitor (nvml_monitor.cpp:35-193) and of EnergyMonitor's RAPL
 *      reader (energy_monitor.cpp:27-164).  librocm_smi64 is loaded at run
 *      time; open returns HSPS_ERR_UNSUPPORTED without it or without a
 *      device energy counter. ---------------------------------------------- */
typedef struct hsps_energy_s* hsps_energy;
int hsps_energy_open(int device, hsps_energy* out);
/* Cumulative device energy in J (monotonic: difference two reads), current
 * socket power in W (-1 when unknown), counter timestamp in s; any may be NULL. */
int hsps_energy_read(hsps_energy E, double* joules, double* watts, double* timestamp_s);
/* Power cap (W), junction temperature (C), busy percent, VRAM used (MiB),
 * device name; -1 where unknown; any may be NULL. */
int hsps_energy_info(hsps_energy E, double* power_cap_w, double* temperature_c, double* busy_pct,
                     double* vram_used_mb, char* name, int name_len);
int hsps_energy_close(hsps_energy E);

/* ---- matrix features for the adaptive ladder
 *      (AdaptiveSelector::extract_features, adaptive_selector.cpp:38-220).
 *      out[0..7] = n, nnz, diag_dominance, symmetry_residual, frobenius,
 *      estimated_cond, lambda_max, is_spd.  The 30-step Lanczos runs on the
 *      device; the O(nnz) scans run on the host copy. ------------------- */
int hsps_features(hsps_csr A, double* h_out);

/* ---- batched solves (no reference counterpart: many independent systems in
 *      one launch).  Systems are concatenated: system s owns rows
 *      [row_off[s], row_off[s+1]) and its columns index inside that range
 *      (block-diagonal).  CG + Jacobi, FP64 or mixed with FP64 refinement. -- */
int hsps_batch_create(hsps_ctx ctx, int n_systems, const int* h_row_off, const int* h_row_ptr,
                      const int* h_col_idx, const double* h_values, hsps_batch* out);
int hsps_batch_destroy(hsps_batch B);
/* d_b, d_x: concatenated FP64 vectors (length row_off[n_systems]).
 * h_iters/h_converged/h_true_res: per-system results (may be NULL). */
int hsps_batch_solve(hsps_batch B, const double* d_b, double* d_x, const hsps_params* p,
                     int* h_iters, int* h_converged, double* h_true_res);
/* Storage chosen per system at creation: CSR, diagonal-major (nonzeros on
 * <= 8 diagonals) or stencil (one value per diagonal + a presence bit per
 * row).  HSPS_BATCH_STORAGE=csr|dia in the environment caps the choice. */
int hsps_batch_storage(hsps_batch B, int* n_csr, int* n_dia, int* n_stencil);

/* ---- sessions (no reference counterpart): systems join a running batched
 *      PCG + Jacobi and leave it at chunk boundaries — continuous batching.
 *      A session holds device tables for at most max_systems systems of
 *      max_rows rows and max_nnz nonzeros in total, all solved with the
 *      params given at creation (tol, max_iter, precision, batch_tile_rows:
 *      0 = the batch autotuner's choice on a probe batch, batch_graphs).  A
 *      system's iterates and result are the same bits whichever systems share
 *      the session and whenever it joins.  One thread at a time per session. -- */
typedef struct hsps_session_s* hsps_session;
int hsps_session_create(hsps_ctx ctx, int max_systems, long long max_rows, long long max_nnz, const hsps_params* p,
                        hsps_session* out);
int hsps_session_destroy(hsps_session S);
/* Uploads an n x n system (host CSR, host b; d_x0 = device initial guess or
 * NULL for zero) and starts it at the next step; *slot identifies it.
 * HSPS_ERR_NOMEM when the session has no room for it now (retry after a
 * result has been taken), HSPS_ERR_ZERO_DIAG as the Jacobi setup. */
int hsps_session_add(hsps_session S, int n, const int* h_row_ptr, const int* h_col_idx, const double* h_values,
                     const double* h_b, const double* d_x0, int* slot);
/* Runs one chunk (p->check_every iterations) of every live system; the
 * slots whose solves ended (reported once each, one chunk after the fact)
 * go to done[0 .. *n_done), at most cap of them per call (the rest follow),
 * and whether each converged to converged[] (may be NULL). */
int hsps_session_step(hsps_session S, int* done, int* converged, int cap, int* n_done);
/* The result of a finished slot: x to h_x (host) and / or d_x (device, n
 * doubles; either may be NULL), iterations, convergence (FP64 true residual
 * below tol) and the FP64 true relative residual; the slot is then free.
 * Returns 1 converged, 0 not. */
int hsps_session_result(hsps_session S, int slot, double* h_x, double* d_x, int* iterations, int* converged,
                        double* true_res);
/* Systems joined and still iterating. */
int hsps_session_live(hsps_session S);

#ifdef __cplusplus
}
#endif
#endif /* HSPS_DEV_H */
