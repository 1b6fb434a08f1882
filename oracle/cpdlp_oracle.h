/*
 * cpdlp_oracle.h -- CPU FP64 restatement of CoolPDLP.jl's PDLP/PDHG path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP library is compared
 * against, and the CPU baseline that bench.py times beside the GPU.  Nothing in
 * the product (libcoolpdlp_hip.so, the coolpdlp_hip package) may call it.
 *
 * Index conventions: 0-based; row pointers int64, column indices int32.
 */
#ifndef CPDLP_ORACLE_H
#define CPDLP_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Mirrors Algorithm{A}(...) keyword arguments, src/algorithms/common.jl:53-78,
 * plus the power-method overrides (NaN / -1 select the reference's rule). */
typedef struct {
    int32_t algorithm;              /* 0 = PDLP, 1 = PDHG */
    int32_t ruiz_iter;              /* 10 (PDHG: 0, pdhg.jl:6-12) */
    double chambolle_pock_alpha;    /* 1.0 */
    double invnorm_scaling;         /* 0.9 */
    double primal_weight_damping;   /* 0.5 */
    double zero_tol;                /* 1e-8 */
    double sufficient_decay;        /* 0.2 */
    double necessary_decay;         /* 0.8 */
    double artificial_decay;        /* 0.36 */
    int32_t check_every;            /* 100 */
    int32_t record_error_history;   /* 1 */
    double termination_reltol;      /* 1e-4 */
    int64_t max_kkt_passes;         /* 100000 */
    double time_limit;              /* 100.0 seconds */
    double powm_tol;                /* NaN: eps(T) * n^3 (Int64 wrap) */
    int64_t powm_maxiter;           /* -1: n */
    uint64_t powm_seed;             /* seed of the power-method start vector */
    /* opt-in infeasibility detection (no counterpart in the reference; the engine's
     * detect_infeasibility / infeasibility_tol): status 4 PRIMAL_INFEASIBLE,
     * 5 DUAL_INFEASIBLE, 6 NUMERICAL_ERROR */
    int32_t detect_infeasibility;   /* 0 */
    int32_t ray_candidates;         /* bit 0: z - z_restart, bit 1: z, bit 2: z - z_last */
    double infeasibility_tol;       /* 1e-8 */
} orc_params;

/* KKTErrors, src/components/errors.jl:8-21 */
typedef struct {
    double primal, primal_scale, dual, dual_scale, gap, gap_scale;
} orc_kkt;

typedef struct {
    int64_t kkt_passes;
    orc_kkt err;
} orc_hist;

/* ConvergenceStats, src/components/termination.jl:44-56, plus solver scalars */
typedef struct {
    orc_kkt err;
    double time_elapsed;
    int64_t kkt_passes;
    int32_t termination_status;     /* 0 OPTIMAL 1 TIME_LIMIT 2 ITERATION_LIMIT 3 STILL_RUNNING (4-6 above) */
    int32_t pad;
    double eta, omega;
    int64_t powm_iters;
    int64_t restarts;
    int64_t n_history;              /* total entries produced (may exceed capacity) */
    double setup_seconds;           /* scaling + step size (before the first step) */
} orc_stats;

void orc_default_params(orc_params* p, int32_t algorithm);

/* The opt-in infeasibility test on one ray (dx, dy) of the scaled LP (A_s in CSR, A_s'
 * in CSR, scaled vectors, D1 / D2): 4 primal infeasible, 5 dual infeasible, 3 neither. */
int32_t orc_ray_test(int64_t m, int64_t n, const int64_t* a_rowptr, const int32_t* a_col,
                     const double* a_val, const int64_t* at_rowptr, const int32_t* at_col,
                     const double* at_val, const double* c, const double* lv, const double* uv,
                     const double* lc, const double* uc, const double* D1, const double* D2,
                     const double* dx, const double* dy, double tol);
void orc_set_threads(int32_t nthreads);

/* CSC (m x n) -> CSR (m x n) by stable counting sort: sparse(transpose(A)),
 * src/utils/mat_csr.jl:58-61.  rowptr has m+1 entries. */
void orc_csc_to_csr(int64_t m, int64_t n, const int64_t* colptr, const int32_t* rowval,
                    const double* nzval, int64_t* rowptr, int32_t* colval, double* csr_val);

/* c = alpha*A*b + beta*c, thread-per-row sequential sum; beta == 0 never reads c.
 * src/utils/mat_csr.jl:63-79 */
void orc_spmv_csr(int64_t m, const int64_t* rowptr, const int32_t* colval, const double* val,
                  double alpha, const double* b, double beta, double* c);

/* ELL (m x d, column-major, padding column index -1, value 0): mat_ell.jl:54-70 */
int64_t orc_ell_width(int64_t m, const int64_t* rowptr);
void orc_csr_to_ell(int64_t m, int64_t d, const int64_t* rowptr, const int32_t* colval,
                    const double* val, int32_t* ell_col, double* ell_val);

/* pdlp_preconditioner + precondition(milp), preconditioning.jl:95-155 and :209-239.
 * Inputs: CSC of A (m x n) with its values; At CSC values are derived internally.
 * Outputs: D1 (m), D2 (n), scaled A values in CSC order (a_csc_out, nnz) and scaled
 * At values in CSC-of-At order expressed on the CSC-of-A layout (at_csc_out, nnz):
 * at_csc_out[k] is the scaled At entry that transposes CSC(A) entry k. */
void orc_scaling(int64_t m, int64_t n, const int64_t* colptr, const int32_t* rowval,
                 const double* nzval, int32_t ruiz_iter, double cp_alpha,
                 double* D1, double* D2, double* a_csc_out, double* at_csc_out);

/* KKT errors of (x, y) on a scaled problem, errors.jl:65-109. */
void orc_kkt_errors(int64_t m, int64_t n,
                    const int64_t* a_rowptr, const int32_t* a_col, const double* a_val,
                    const int64_t* at_rowptr, const int32_t* at_col, const double* at_val,
                    const double* c, const double* lv, const double* uv,
                    const double* lc, const double* uc, const double* D1, const double* D2,
                    const double* x, const double* y, orc_kkt* out);

/* Power-method start vector: sum of 12 uniforms - 6, counter-based (exact on any device). */
void orc_powm_start(int64_t n, uint64_t seed, double* x0);

/* spectral_norm via power iteration on AtA (linalg.jl:86-96, powm! rule). */
double orc_spectral_norm(int64_t m, int64_t n,
                         const int64_t* a_rowptr, const int32_t* a_col, const double* a_val,
                         const int64_t* at_rowptr, const int32_t* at_col, const double* at_val,
                         double tol, int64_t maxiter, uint64_t seed, int64_t* iters_out);

/* restart_check!'s candidate choice (pdlp.jl:149-165) and should_restart
 * (src/components/restart.jl:40-57) on the five KKT errors of a check: the iterate,
 * the average, both one PDHG step earlier, and the restart point.  Returns 1 to
 * restart; *from_avg = 1 when the candidate is the average (ties go to the average). */
int32_t orc_restart_decision(const orc_kkt* e_sol, const orc_kkt* e_avg, const orc_kkt* e_last,
                             const orc_kkt* e_avg_last, const orc_kkt* e_restart, double omega,
                             int64_t inner, int64_t total, const orc_params* p, int32_t* from_avg);

/* primal_weight_update!, src/components/step_size.jl:58-75, given the distances
 * dx = ||x_cand - x_restart||, dy = ||y_cand - y_restart||. */
double orc_primal_weight_update(double omega, double dx, double dy, double damping, double zero_tol);

/* Per-check record of a solve (for tests): arrays of `cap` entries, filled for the
 * first min(count, cap) checks.  err holds 5 entries per check: the iterate, the
 * average, the iterate and the average one step earlier, and the restart point (PDHG:
 * the iterate only, the rest NaN).  x_avg / y_avg (may be NULL) receive the final
 * average, unscaled. */
typedef struct {
    int64_t cap;
    int64_t count;
    int64_t* passes;       /* kkt_passes at the check */
    orc_kkt* err;          /* 5 * cap */
    int32_t* from_avg;     /* the candidate was the average */
    int32_t* restarted;    /* the check restarted */
    double* omega;         /* primal weight in force at the check */
    double* omega_after;   /* primal weight after the check (updated on a restart) */
    int64_t* inner;        /* inner / total iteration counters at the check */
    int64_t* total;
    double* x_avg;
    double* y_avg;
} orc_trace;

/* Full solve: solve(milp, sol, algo), common.jl:186-201.  Input is the UNSCALED LP
 * as CSC; x0/y0 may be NULL (zero start).  Outputs unscaled x (n), y (m). */
int orc_solve_traced(int64_t m, int64_t n, const int64_t* colptr, const int32_t* rowval,
                     const double* nzval, const double* c, const double* lv, const double* uv,
                     const double* lc, const double* uc, const double* x0, const double* y0,
                     const orc_params* p, double* x_out, double* y_out, orc_stats* stats,
                     orc_hist* hist, int64_t hist_cap, orc_trace* trace);
int orc_solve(int64_t m, int64_t n, const int64_t* colptr, const int32_t* rowval,
              const double* nzval, const double* c, const double* lv, const double* uv,
              const double* lc, const double* uc, const double* x0, const double* y0,
              const orc_params* p, double* x_out, double* y_out, orc_stats* stats,
              orc_hist* hist, int64_t hist_cap);

#ifdef __cplusplus
}
#endif
#endif
