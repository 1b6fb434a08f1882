/*
 * coolpdlp_hip.h -- C ABI of libcoolpdlp_hip.so, the MI355X (gfx950) PDLP/PDHG engine.
 *
 * Plain C types only (pointers, sizes, PODs); no C++ types or exceptions cross it.
 * Every function returns an int status (CPDLP_OK = 0); the message of the last
 * error on the calling thread is cpdlp_last_error().  Solver outcomes (OPTIMAL,
 * TIME_LIMIT, ...) are not errors: they are reported in cpdlp_stats.
 *
 * The reference (CoolPDLP.jl) is pure Julia and has no FFI of its own.  Each entry
 * point below names the reference function or plug-in method it replaces; the
 * Julia-side ccall binding a maintainer would add is
 * coolpdlp.jl_amd/julia/CoolPDLPHIP.jl.  Paths are relative to the reference repo.
 *
 * Threading: one handle owns one HIP stream and its device memory; calls on
 * different handles may run on different threads; a single handle is not
 * thread-safe (like the reference's single-task solve!).
 */
#ifndef COOLPDLP_HIP_H
#define COOLPDLP_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------- */
#define CPDLP_OK 0
#define CPDLP_ERR_DIMENSION 1   /* DimensionMismatch, src/problems/milp.jl:71-76 */
#define CPDLP_ERR_ARGUMENT 2    /* ArgumentError, milp.jl:83-92, utils/device.jl:6-12 */
#define CPDLP_ERR_MEMORY 3      /* device allocation failed */
#define CPDLP_ERR_HIP 4         /* HIP runtime error */
#define CPDLP_ERR_STATE 5       /* call out of order (e.g. solution before solve) */

/* ---- enums ------------------------------------------------------------------- */
/* TerminationStatus, src/components/termination.jl:13 (same order) */
#define CPDLP_OPTIMAL 0
#define CPDLP_TIME_LIMIT 1
#define CPDLP_ITERATION_LIMIT 2
#define CPDLP_STILL_RUNNING 3
/* Only with cpdlp_params.detect_infeasibility = 1 (no counterpart in the reference,
 * whose TerminationStatus stops at STILL_RUNNING): */
#define CPDLP_PRIMAL_INFEASIBLE 4  /* a dual ray (Farkas certificate) was found */
#define CPDLP_DUAL_INFEASIBLE 5    /* a primal ray: the LP is unbounded or infeasible */
#define CPDLP_NUMERICAL_ERROR 6    /* the KKT errors became NaN or infinite */

#define CPDLP_ALGO_PDLP 0       /* Algorithm{:PDLP}, src/algorithms/pdlp.jl:6-11 */
#define CPDLP_ALGO_PDHG 1       /* Algorithm{:PDHG}, src/algorithms/pdhg.jl:6-16 */

#define CPDLP_PREC_FP64 0       /* T = Float64 everywhere */
#define CPDLP_PREC_FP32 1       /* T = Float32 (the reference's float_type = Float32) */
#define CPDLP_PREC_MIXED 2      /* FP32 matrix/iterates/SpMV; FP64 checks; FP64 polish */

#define CPDLP_FMT_AUTO 0        /* autotuned per matrix */
#define CPDLP_FMT_CSR 1         /* GPUSparseMatrixCSR, src/utils/mat_csr.jl */
#define CPDLP_FMT_ELL 2         /* GPUSparseMatrixELL, src/utils/mat_ell.jl */
#define CPDLP_FMT_XSLICE 3      /* rows longer than 16 cut into 8 column slices, one per XCD's L2 */
#define CPDLP_FMT_XSLICE_ALL 4  /* every row cut into 8 column slices */

/* ---- problem input --------------------------------------------------------------
 * MILP(; c, lv, uv, A, lc, uc), src/problems/milp.jl:55-115, with A given as the
 * reference's SparseMatrixCSC (colptr n+1, rowval nnz, nzval nnz; rows sorted within
 * each column).  index_base 1 accepts Julia's arrays unchanged.  +-inf bounds allowed.
 * Host pointers; read only during cpdlp_create. */
typedef struct cpdlp_lp {
    int64_t m, n;
    int32_t index_base;      /* 0 or 1 */
    int32_t index_bits;      /* 32 or 64: width of colptr and rowval */
    const void* colptr;
    const void* rowval;
    const double* nzval;
    const double* c;
    const double* lv;
    const double* uv;
    const double* lc;
    const double* uc;
} cpdlp_lp;

/* ---- parameters -------------------------------------------------------------------
 * Algorithm{A}(T, Ti, M; kwargs...), src/algorithms/common.jl:53-116: same names,
 * same defaults (cpdlp_default_params), plus engine options. */
typedef struct cpdlp_params {
    int32_t algorithm;              /* CPDLP_ALGO_* */
    int32_t ruiz_iter;              /* 10 (PDHG: 0) */
    double chambolle_pock_alpha;    /* 1.0 */
    double invnorm_scaling;         /* 0.9 */
    double primal_weight_damping;   /* 0.5 (PDHG: NaN) */
    double zero_tol;                /* 1e-8 */
    double sufficient_decay;        /* 0.2 (PDHG: NaN) */
    double necessary_decay;         /* 0.8 (PDHG: NaN) */
    double artificial_decay;        /* 0.36 (PDHG: NaN) */
    int32_t show_progress;          /* 0; 1: a progress line on stderr every 2 s */
    int32_t check_every;            /* 100 */
    int32_t record_error_history;   /* 1 */
    int32_t precision;              /* CPDLP_PREC_* (T) */
    double termination_reltol;      /* 1e-4 */
    int64_t max_kkt_passes;         /* 100000 */
    double time_limit;              /* 100.0 s */
    /* engine options (no reference counterpart) */
    int32_t format;                 /* CPDLP_FMT_* (M) */
    int32_t device;                 /* HIP device ordinal */
    double powm_tol;                /* NaN: eps(T)*n^3 rule of powm!, linalg.jl:94 */
    int64_t powm_maxiter;           /* -1: n */
    uint64_t powm_seed;             /* power-method start-vector seed */
    double polish_reltol;           /* MIXED: FP32 phase stops at max(this, reltol) */
    int32_t use_graph;              /* capture the check_every-step chunk in a hipGraph */
    int32_t long_row;               /* rows longer than this use a wavefront per row (64) */
    int32_t batch_streams;          /* HIP streams of a batch (cpdlp_batch_*), 1..8 (4) */
    int32_t detect_infeasibility;   /* 0; 1: test ray certificates at every check that is
                                       not OPTIMAL, and stop on NaN errors */
    double infeasibility_tol;       /* 1e-8: a ray qualifies when its violation is at most
                                       this times its objective */
} cpdlp_params;

/* KKTErrors, src/components/errors.jl:8-21 */
typedef struct cpdlp_kkt {
    double primal, primal_scale, dual, dual_scale, gap, gap_scale;
} cpdlp_kkt;

/* ConvergenceStats, src/components/termination.jl:44-56, plus engine counters */
typedef struct cpdlp_stats {
    cpdlp_kkt err;
    double time_elapsed;            /* seconds since solve start (incl. preprocess) */
    int64_t kkt_passes;
    int32_t termination_status;     /* CPDLP_OPTIMAL ... */
    int32_t format_A, format_At;    /* chosen formats */
    int32_t line_tiles;             /* line-grouped short tiles in use: bit 0 A, bit 1 A' */
    double eta, omega;
    int64_t powm_iters;
    int64_t restarts;
    int64_t n_history;
    int64_t fp32_passes;            /* MIXED: passes done in FP32 before polishing */
    double setup_seconds;           /* scaling + step size on the device */
    double loop_seconds;            /* PDHG loop incl. checks */
    double step_kernel_ms;          /* device time of all step kernels (HIP events) */
    int64_t step_kernel_launches;
} cpdlp_stats;

/* The loop's state between two checks (no reference counterpart as a value; it is
 * PDLPState, src/algorithms/pdlp.jl:19-43, less its scratch and per-check statistics, which
 * the next check recomputes): for checkpoint and resume, and for continuing a batch
 * instance on a plain engine.  The vectors travel beside it in caller buffers:
 * CPDLP_STATE_VECTORS primal vectors of n entries (x) and as many dual vectors of m (y),
 * in the order sol, sol_last, sol_avg, sol_avg_last, sol_restart, SCALED (the solver's own
 * variables), in the LP's original order, with the diagonal scaling they belong to (D2,
 * n; D1, m; the unscaled point is x .* D2, y .* D1). */
#define CPDLP_STATE_VECTORS 5
typedef struct cpdlp_state {
    double eta;                     /* step size (StepSizes.η) */
    double omega;                   /* primal weight (StepSizes.ω) */
    double eta_sum;                 /* step sizes summed since the last restart (StepSizes.η_sum) */
    int64_t kkt_passes;             /* ConvergenceStats.kkt_passes */
    int64_t inner, total;           /* IterationCounter (iteration.jl:1-5) */
    int64_t restarts;               /* IterationCounter.outer */
    int64_t fp32_passes;            /* passes run in FP32 so far */
} cpdlp_state;

typedef struct cpdlp_handle cpdlp_handle;
typedef struct cpdlp_matrix cpdlp_matrix;

/* defaults of Algorithm{:PDLP}(...) / PDHG(...): common.jl:53-78, pdhg.jl:6-16 */
int cpdlp_default_params(cpdlp_params* p, int32_t algorithm);

/* preprocess + initialize inputs: copies the LP to HBM, builds CSR(A) and CSR(At)
 * (perform_conversion, src/components/conversion.jl:37-55; set_matrix_type,
 * src/problems/modify.jl:80-105).  Validates like MILP(...), milp.jl:71-92. */
int cpdlp_create(const cpdlp_lp* lp, const cpdlp_params* params, cpdlp_handle** out);

/* solve(milp, sol, algo) / solve(milp, algo), src/algorithms/common.jl:186-209:
 * scaling, step size, primal weight, the PDLP/PDHG loop to termination.
 * x0 (n) / y0 (m) are the unscaled start (NULL = zero, common.jl:207). */
int cpdlp_solve(cpdlp_handle* h, const double* x0, const double* y0);

/* initialize(milp, sol, algo), src/algorithms/common.jl:176 (pdlp.jl:45-67,
 * pdhg.jl:40-54): what cpdlp_solve does before its loop (preprocess, step size,
 * primal weight, the start point on the device), the loop then run in pieces by
 * cpdlp_run_checks.  Every precision: in mixed precision cpdlp_run_checks runs the FP32
 * phase, then the FP64 one from its state, as cpdlp_solve does. */
int cpdlp_initialize(cpdlp_handle* h, const double* x0, const double* y0);

/* solve!(state, milp, algo)'s loop, pdlp.jl:69-94 / pdhg.jl:58-74, for at most
 * max_checks chunks of check_every steps (the reference yields once per chunk,
 * pdlp.jl:77).  *status = the termination status after the last check
 * (CPDLP_STILL_RUNNING while the loop has not terminated).  Running the chunks one
 * call at a time gives the same iterates as one cpdlp_solve. */
int cpdlp_run_checks(cpdlp_handle* h, int64_t max_checks, int32_t* status);

/* The loop's state after the last check (cpdlp_run_checks, cpdlp_solve): *st, x
 * (CPDLP_STATE_VECTORS * n), y (CPDLP_STATE_VECTORS * m), D2 (n) and D1 (m), host buffers
 * (see cpdlp_state; FP32 iterates are widened to double).  No reference counterpart. */
int cpdlp_get_state(cpdlp_handle* h, cpdlp_state* st, double* x, double* y, double* D2, double* D1);

/* cpdlp_initialize, then the state of another run of the same LP in place of the start
 * point: cpdlp_run_checks continues that run.  Vectors whose scaling (D2, D1) equals this
 * handle's, bit for bit, are taken as they are, so a run stopped after some checks and
 * resumed here gives the same iterates as one that never stopped (checkpoint and
 * resume); others are carried over through the unscaled point (x .* D2 ./ D2_own), so a
 * batch instance continues on this engine (cpdlp_batch_take_state).  The step size is the
 * state's.  The time limit counts from the resume, and the error history restarts.
 * FP64 and FP32 handles (CPDLP_ERR_ARGUMENT for mixed: the state carries no phase).  No
 * reference counterpart. */
int cpdlp_resume(cpdlp_handle* h, const cpdlp_state* st, const double* x, const double* y,
                 const double* D2, const double* D1);

/* The solution kept on the device, as the reference's sol (common.jl:200): unscaled x
 * (n) and y (m), original order, into caller-owned DEVICE buffers (either may be NULL).
 * Valid after cpdlp_solve, cpdlp_initialize or cpdlp_run_checks. */
int cpdlp_get_solution_device(cpdlp_handle* h, double* x_dev, double* y_dev);

/* get_solution + Array(...), common.jl:233-235 and src/moi.jl:215-216:
 * unscaled x (n) and y (m) into host buffers. */
int cpdlp_get_solution(cpdlp_handle* h, double* x, double* y);

/* state.stats, src/components/termination.jl:44-76 */
int cpdlp_get_stats(cpdlp_handle* h, cpdlp_stats* out);

/* stats.error_history, termination.jl:56: up to cap entries; *count = total */
int cpdlp_get_history(cpdlp_handle* h, int64_t* kkt_passes, cpdlp_kkt* errs, int64_t cap,
                      int64_t* count);

/* kkt_errors!(scratch, sol, milp), src/components/errors.jl:65-109, on the
 * preconditioned problem for an UNSCALED host point (x, y); errors are in the
 * original space as in the reference.  Requires cpdlp_solve or cpdlp_prepare. */
int cpdlp_kkt_errors(cpdlp_handle* h, const double* x, const double* y, cpdlp_kkt* out);

/* No counterpart in the reference.  For callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
 * cpdlp_kkt_errors for an UNSCALED point given as DEVICE pointers x_dev (n) and y_dev (m),
 * original order, computed on the handle's stream with every intermediate in a fixed
 * per-device scratch array of 32 MiB (4,194,304 doubles) inside the library: no device
 * memory is allocated.  The LP must fit: n + m plus the block partials of both products
 * (a few per 256 rows) within that scratch, else CPDLP_ERR_ARGUMENT; the preprocessing
 * must have run (cpdlp_prepare or a solve, which allocate), else CPDLP_ERR_STATE.  Same
 * errors as cpdlp_kkt_errors on the same point.  Every call on one GPU shares that GPU's
 * scratch: calls that overlap on one GPU (from different threads or handles) race. */
int cpdlp_kkt_errors_device_noalloc(cpdlp_handle* h, const double* x_dev, const double* y_dev,
                                    cpdlp_kkt* out);

/* preprocess(...) + initialize's step sizes without running the loop:
 * pdlp_preconditioner / precondition (preconditioning.jl:146-155, :61-91),
 * fixed_stepsize (step_size.jl:22-27), primal_weight_init (:29-40). */
int cpdlp_prepare(cpdlp_handle* h);

/* Scaling results in the CSC(A) layout: D1 (m), D2 (n), scaled A values (nnz) and
 * scaled At values (nnz, transposing CSC(A) entry k).  Any pointer may be NULL. */
int cpdlp_get_scaling(cpdlp_handle* h, double* D1, double* D2, double* a_csc, double* at_csc);

/* CSR(A) structure built by cpdlp_create (0-based): rowptr (m+1), colval (nnz).
 * GPUSparseMatrixCSR(A), src/utils/mat_csr.jl:58-61. */
int cpdlp_get_csr(cpdlp_handle* h, int64_t* rowptr, int32_t* colval);

/* eta (after cpdlp_prepare or cpdlp_solve) and power-method iterations. */
int cpdlp_get_step_size(cpdlp_handle* h, double* eta, double* omega, int64_t* powm_iters);

/* Measurement hook (no reference counterpart): average device milliseconds per launch
 * of the primal half-step kernel (A'y + projection + extrapolation + averaging) and of
 * the dual half-step kernel (A xbar + projection + averaging), each timed alone with
 * HIP events over `reps` launches on a scratch copy of the iterates.  After a solve. */
int cpdlp_profile_steps(cpdlp_handle* h, int32_t reps, double* ms_primal, double* ms_dual);

void cpdlp_destroy(cpdlp_handle* h);

/* thread-local message of the last failing call on this thread ("" if none) */
const char* cpdlp_last_error(void);

/* library version string and the device architecture it was built for */
const char* cpdlp_version(void);

/* ---- batched / served solves (no reference counterpart: the reference solves one
 * LP per call to solve, src/algorithms/common.jl:186-209) ----------------------------
 * A batch holds many independent LPs ("instances") in HBM and runs the same PDLP
 * algorithm on all of them, one workgroup per instance, with per-instance scaling,
 * step size, restarts and termination (all on the device).  Instances can be added
 * while others are running (continuous batching): cpdlp_batch_run advances every
 * admitted instance by a bounded number of checks and returns; finished instances
 * are reported by cpdlp_batch_poll and their HBM is returned by cpdlp_batch_release.
 * Precision follows params->precision (MIXED: FP32 bulk, FP64 checks + FP64 polish).
 * Each instance must satisfy m + n <= 65536 (its gathered vectors live in LDS). */
typedef struct cpdlp_batch cpdlp_batch;

/* hbm_bytes: arena size for instance data (bytes of HBM reserved up front). */
int cpdlp_batch_create(const cpdlp_params* params, int64_t hbm_bytes, cpdlp_batch** out);

/* Admit one LP (host arrays, copied); *id identifies it.  Returns CPDLP_ERR_MEMORY
 * (and admits nothing) when the arena lacks room: the caller retries after a release.
 * Returns CPDLP_ERR_ARGUMENT when the LP's footprint exceeds the whole arena (it can
 * never be admitted).  A failed add leaves the arena as it was. */
int cpdlp_batch_add(cpdlp_batch* b, const cpdlp_lp* lp, int64_t* id);

/* cpdlp_batch_add with a start point: x0 (n) / y0 (m) are the unscaled start, as the
 * sol_init of solve(milp, sol, algo) (src/algorithms/common.jl:186-209); either may be
 * NULL (zero).  The instance starts from (x0 / D2, y0 / D1), with avg = restart = start,
 * as cpdlp_solve does. */
int cpdlp_batch_add_start(cpdlp_batch* b, const cpdlp_lp* lp, const double* x0, const double* y0,
                          int64_t* id);

/* HBM footprint an LP of this size would take in the arena. */
int64_t cpdlp_batch_footprint(int64_t m, int64_t n, int64_t nnz);

/* Advance every admitted, unfinished instance by up to max_checks checks (each check
 * = check_every PDHG steps + the KKT/restart check); blocks until done.  Returns the
 * number of still-running instances in *running. */
int cpdlp_batch_run(cpdlp_batch* b, int32_t max_checks, int64_t* running);

/* Ids of instances finished since the last poll (up to cap); *count = number written. */
int cpdlp_batch_poll(cpdlp_batch* b, int64_t* ids, int64_t cap, int64_t* count);

/* Unscaled solution and stats of an instance (finished or not). */
int cpdlp_batch_get(cpdlp_batch* b, int64_t id, double* x, double* y, cpdlp_stats* st);

/* Return an instance's HBM to the arena (after cpdlp_batch_get). */
int cpdlp_batch_release(cpdlp_batch* b, int64_t id);

/* Take a running instance out of the batch with its loop state (as cpdlp_get_state:
 * *st, x (CPDLP_STATE_VECTORS * n), y (CPDLP_STATE_VECTORS * m), D2 (n), D1 (m), host
 * buffers), for cpdlp_resume on a plain engine of the same LP: the instance's HBM returns
 * to the arena and its id becomes unknown.  Only between rounds, for an instance that is
 * set up and still running (else CPDLP_ERR_STATE).  No reference counterpart. */
int cpdlp_batch_take_state(cpdlp_batch* b, int64_t id, cpdlp_state* st, double* x, double* y,
                           double* D2, double* D1);

/* Bytes of the arena in use and its size. */
int cpdlp_batch_memory(cpdlp_batch* b, int64_t* used, int64_t* size);

void cpdlp_batch_destroy(cpdlp_batch* b);

/* ---- matrix plug-in: GPUSparseMatrixCSR / GPUSparseMatrixELL + mul! ----------------
 * GPUSparseMatrixCSR(A::SparseMatrixCSC), src/utils/mat_csr.jl:58-61;
 * GPUSparseMatrixELL(A), src/utils/mat_ell.jl:54-70; values FP64.  transpose != 0
 * builds the matrix of A' instead (the reference's MILP.At, src/problems/milp.jl:60). */
int cpdlp_matrix_create(int64_t m, int64_t n, int32_t index_base, int32_t index_bits,
                        const void* colptr, const void* rowval, const double* nzval,
                        int32_t format, int32_t transpose, int32_t device, cpdlp_matrix** out);

/* LinearAlgebra.mul!(c, A, b, alpha, beta), mat_csr.jl:81-92 / mat_ell.jl:87-98:
 * c = alpha*A*b + beta*c on DEVICE pointers, asynchronously on `stream` (a hipStream_t;
 * NULL = HIP's null stream), ordered with the caller's other work on that stream.
 * beta == 0 does not read c. */
int cpdlp_matrix_mul(cpdlp_matrix* A, double alpha, const double* b_dev, double beta,
                     double* c_dev, void* stream);

/* Same on host buffers (copies in and out; synchronous). */
int cpdlp_matrix_mul_host(cpdlp_matrix* A, double alpha, const double* b, double beta, double* c);

/* size(A), nnz(A), the format in use and, for ELL, the width d. */
int cpdlp_matrix_info(cpdlp_matrix* A, int64_t* m, int64_t* n, int64_t* nnz, int32_t* format,
                      int64_t* ell_width);

/* ELL arrays (column-major m x d, padding column -1, value 0), mat_ell.jl:54-70. */
int cpdlp_matrix_get_ell(cpdlp_matrix* A, int32_t* colval, double* nzval);

void cpdlp_matrix_destroy(cpdlp_matrix* A);

/* ---- serving: continuous batching of solve requests (no reference counterpart) ----
 * The served form of solve(milp, sol, algo) (src/algorithms/common.jl:186-209).
 * Requests are copied at submit (or borrowed).  Those with m + n <= batch_limit go to a
 * batch engine (as cpdlp_batch_*: admitted as arena room frees up, advanced by
 * cpdlp_server_pump), those with m + n <= batch2_limit to a second batch engine whose
 * rounds run apart from the first one's (on a thread of their own after
 * cpdlp_server_start), so that long rounds of medium LPs never hold up small ones;
 * the others to plain engines (as cpdlp_solve, CSR format, one HIP stream each) run by
 * worker threads in lanes by nonzero count (a request takes the first lane whose
 * lane_max_nnz admits it, else the last), shortest job first within a lane, under the
 * plain_bytes HBM budget.  Submit, poll and get may be called from any thread; pump
 * from one thread at a time, or by the library itself after cpdlp_server_start. */
#define CPDLP_SERVER_MAX_LANES 4
typedef struct cpdlp_server_config {
    int64_t batch_bytes;            /* batch route's HBM arena (16 GiB) */
    int64_t plain_bytes;            /* plain route's HBM budget (64 GiB) */
    int32_t batch_limit;            /* m + n limit of the batch route, <= 65536; 0: none (8192) */
    int32_t checks_per_round;       /* checks per pump for the smallest instances (20) */
    int32_t n_lanes;                /* plain lanes, 1..CPDLP_SERVER_MAX_LANES (2) */
    int32_t lane_workers[CPDLP_SERVER_MAX_LANES];   /* worker threads per lane (3, 1) */
    double lane_max_nnz[CPDLP_SERVER_MAX_LANES];    /* nonzeros a lane admits (1e6, inf) */
    int64_t batch2_bytes;           /* second batch route's HBM arena (2 GiB) */
    int32_t batch2_limit;           /* its m + n limit, <= 65536; none when <= batch_limit (0) */
    int32_t batch2_checks;          /* its checks per round; 0: checks_per_round (0) */
} cpdlp_server_config;
typedef struct cpdlp_server cpdlp_server;

int cpdlp_server_default_config(cpdlp_server_config* cfg);
/* params as for cpdlp_create (the format is fixed to CSR on the plain route).  Keep
 * params->batch_streams (per batch route) plus the lanes' workers within the device's
 * hardware queues (4 by default): streams beyond them share queues, and batch rounds then
 * wait behind plain solves' kernels. */
int cpdlp_server_create(const cpdlp_params* params, const cpdlp_server_config* cfg, cpdlp_server** out);
/* Queue one LP; x0 / y0 (NULL: zero) are an unscaled start as in cpdlp_solve. */
int cpdlp_server_submit(cpdlp_server* s, const cpdlp_lp* lp, const double* x0, const double* y0,
                        int64_t* id);
/* cpdlp_server_submit without the copy: every array the LP and the start point name must
 * stay valid and unchanged until cpdlp_server_get has returned for this id.  The arrays
 * are validated when an engine takes the request; an invalid one is reported by
 * cpdlp_server_get.  For callers that submit large LPs at a high rate. */
int cpdlp_server_submit_borrowed(cpdlp_server* s, const cpdlp_lp* lp, const double* x0,
                                 const double* y0, int64_t* id);
/* One scheduling round: admit, advance each batch route by its checks per round, harvest.
 * *pending = requests not yet finished. */
int cpdlp_server_pump(cpdlp_server* s, int64_t* pending);
/* Run the pump loop on a library thread from now on (idempotent): callers then only
 * submit, poll and get, and cpdlp_server_pump only reports the pending count.  A failure
 * of the loop is returned by every later pump and poll. */
int cpdlp_server_start(cpdlp_server* s);
/* Ids of finished requests (each reported once). */
int cpdlp_server_poll(cpdlp_server* s, int64_t* ids, int64_t cap, int64_t* count);
/* A finished request's unscaled x (n), y (m) and stats; *seconds = submit to finish;
 * *route = -1 (batch), -2 (second batch) or the plain lane.  Frees the request.  Returns the request's own
 * error code when its solve failed (the message in cpdlp_last_error).  A request that
 * has not finished gives CPDLP_ERR_STATE with the message "request still running" and
 * stays queued (a borrowed request's arrays are still read). */
int cpdlp_server_get(cpdlp_server* s, int64_t id, double* x, double* y, cpdlp_stats* st,
                     double* seconds, int32_t* route);
/* Waits for running plain solves; requests not yet admitted are dropped. */
void cpdlp_server_destroy(cpdlp_server* s);

#ifdef __cplusplus
}
#endif
#endif
