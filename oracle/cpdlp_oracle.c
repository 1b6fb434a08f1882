/*
 * cpdlp_oracle.c -- CPU FP64 restatement of CoolPDLP.jl's PDLP/PDHG path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker that the HIP library is compared against
 * and the CPU baseline timed beside it.  The product never links or calls it.
 *
 * Every routine follows the reference operation by operation (file:line cited at
 * each function).  Compiled with -ffp-contract=off so that a*b+c is two roundings,
 * as in Julia broadcasts.  Matrix-vector products use the reference's
 * thread-per-row sequential sum (src/utils/mat_csr.jl:63-79), parallelised over
 * rows with OpenMP.  Reductions (norm/dot/sum) are blocked: fixed 4096-element
 * blocks summed sequentially, then the block sums sequentially, so results do
 * not depend on the thread count.  (The reference uses BLAS nrm2/dot and Julia's
 * pairwise sum there; those bits are not reproducible without Julia, so parity
 * on those scalars is tolerance-based.)
 *
 * Pinning status of each part (what checks it) is listed in the oracle tests:
 * tests/test_oracle.py, tests/test_oracle_restart.py (restart decision, averaging, primal
 * weight update, the restart itself), and tests/test_infeasibility.py for the opt-in ray
 * test (which has no counterpart in the reference).  Parity unpinned: orc_powm_start (the reference draws
 * StableRNG(0) randn, linalg.jl:92, not reproducible here), and the last bits of
 * the BLAS-backed reductions.
 */
#include "cpdlp_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define RBLK 4096

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

void orc_set_threads(int32_t nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

/* Algorithm{A} defaults, common.jl:53-78; PDHG overrides, pdhg.jl:6-16. */
void orc_default_params(orc_params* p, int32_t algorithm) {
    memset(p, 0, sizeof(*p));
    p->algorithm = algorithm;
    p->ruiz_iter = algorithm == 0 ? 10 : 0;
    p->chambolle_pock_alpha = 1.0;
    p->invnorm_scaling = 0.9;
    p->primal_weight_damping = algorithm == 0 ? 0.5 : NAN;
    p->zero_tol = 1e-8;
    p->sufficient_decay = algorithm == 0 ? 0.2 : NAN;
    p->necessary_decay = algorithm == 0 ? 0.8 : NAN;
    p->artificial_decay = algorithm == 0 ? 0.36 : NAN;
    p->check_every = 100;
    p->record_error_history = 1;
    p->termination_reltol = 1e-4;
    p->max_kkt_passes = 100000;
    p->time_limit = 100.0;
    p->powm_tol = NAN;
    p->powm_maxiter = -1;
    p->powm_seed = 0;
    p->detect_infeasibility = 0;
    p->ray_candidates = 7;
    p->infeasibility_tol = 1e-8;
}

/* ---------------------------------------------------------------- helpers */

/* linalg.jl:4-5 */
static inline double pos_part(double a) { return a > 0.0 ? a : 0.0; }
static inline double neg_part(double a) { return -(a < 0.0 ? a : 0.0); }
/* Julia min/max on floats: NaN-propagating */
static inline double jmax(double a, double b) { return (a != a || b != b) ? NAN : (a > b ? a : b); }
static inline double jmin(double a, double b) { return (a != a || b != b) ? NAN : (a < b ? a : b); }
/* linalg.jl:19 proj_box(x, l, u) = min(u, max(l, x)) */
static inline double proj_box(double x, double l, double u) { return jmin(u, jmax(l, x)); }
/* linalg.jl:17 safeprod_left */
static inline double safeprod_left(double l, double r) { return isinf(l) ? r : l * r; }
/* linalg.jl:26-42 proj_multiplier */
static inline double proj_multiplier(double lam, double l, double u) {
    int lmin = (l == -INFINITY), umax = (u == INFINITY);
    if (lmin) return umax ? 0.0 : -neg_part(lam);
    return umax ? pos_part(lam) : lam;
}
/* linalg.jl:49-53 combine */
static inline double combine(double l, double u) {
    double ls = isfinite(l) ? fabs(l) : 0.0;
    double us = isfinite(u) ? fabs(u) : 0.0;
    return jmax(jmax(0.0, ls), us);
}

typedef double (*elem_fn)(int64_t i, const void* ctx);

/* blocked deterministic sum of f(i), i in [0, len) */
static double bsum(int64_t len, elem_fn f, const void* ctx) {
    if (len <= 0) return 0.0;
    int64_t nb = (len + RBLK - 1) / RBLK;
    double* part = (double*)malloc((size_t)nb * sizeof(double));
#pragma omp parallel for schedule(static)
    for (int64_t b = 0; b < nb; ++b) {
        int64_t lo = b * RBLK, hi = lo + RBLK < len ? lo + RBLK : len;
        double s = 0.0;
        for (int64_t i = lo; i < hi; ++i) s += f(i, ctx);
        part[b] = s;
    }
    double s = 0.0;
    for (int64_t b = 0; b < nb; ++b) s += part[b];
    free(part);
    return s;
}

typedef struct { const double* a; const double* b; } vv_ctx;
static double f_sq(int64_t i, const void* c) { const double* a = ((const vv_ctx*)c)->a; return a[i] * a[i]; }
static double f_dot(int64_t i, const void* c) { const vv_ctx* v = (const vv_ctx*)c; return v->a[i] * v->b[i]; }
static double f_sum(int64_t i, const void* c) { return ((const vv_ctx*)c)->a[i]; }
static double f_dsq(int64_t i, const void* c) {
    const vv_ctx* v = (const vv_ctx*)c; double d = v->a[i] - v->b[i]; return d * d;
}
static double nrm2(int64_t len, const double* a) { vv_ctx c = {a, 0}; return sqrt(bsum(len, f_sq, &c)); }
static double dotv(int64_t len, const double* a, const double* b) { vv_ctx c = {a, b}; return bsum(len, f_dot, &c); }
static double sumv(int64_t len, const double* a) { vv_ctx c = {a, 0}; return bsum(len, f_sum, &c); }
static double dist2(int64_t len, const double* a, const double* b) { vv_ctx c = {a, b}; return sqrt(bsum(len, f_dsq, &c)); }

/* ------------------------------------------------------------- matrices */

void orc_csc_to_csr(int64_t m, int64_t n, const int64_t* colptr, const int32_t* rowval,
                    const double* nzval, int64_t* rowptr, int32_t* colval, double* csr_val) {
    int64_t nnz = colptr[n];
    memset(rowptr, 0, (size_t)(m + 1) * sizeof(int64_t));
    for (int64_t k = 0; k < nnz; ++k) rowptr[rowval[k] + 1]++;
    for (int64_t i = 0; i < m; ++i) rowptr[i + 1] += rowptr[i];
    int64_t* next = (int64_t*)malloc((size_t)(m + 1) * sizeof(int64_t));
    memcpy(next, rowptr, (size_t)(m + 1) * sizeof(int64_t));
    for (int64_t j = 0; j < n; ++j)
        for (int64_t k = colptr[j]; k < colptr[j + 1]; ++k) {
            int64_t dst = next[rowval[k]]++;
            colval[dst] = (int32_t)j;
            if (csr_val) csr_val[dst] = nzval[k];
        }
    free(next);
}

/* perm[k_csr] = k_csc for the stable transpose above */
static void csc_to_csr_perm(int64_t m, int64_t n, const int64_t* colptr, const int32_t* rowval,
                            int64_t* rowptr, int32_t* colval, int64_t* perm) {
    int64_t nnz = colptr[n];
    memset(rowptr, 0, (size_t)(m + 1) * sizeof(int64_t));
    for (int64_t k = 0; k < nnz; ++k) rowptr[rowval[k] + 1]++;
    for (int64_t i = 0; i < m; ++i) rowptr[i + 1] += rowptr[i];
    int64_t* next = (int64_t*)malloc((size_t)(m + 1) * sizeof(int64_t));
    memcpy(next, rowptr, (size_t)(m + 1) * sizeof(int64_t));
    for (int64_t j = 0; j < n; ++j)
        for (int64_t k = colptr[j]; k < colptr[j + 1]; ++k) {
            int64_t dst = next[rowval[k]]++;
            colval[dst] = (int32_t)j;
            perm[dst] = k;
        }
    free(next);
}

/* mat_csr.jl:63-79: s = 0; s += nzval[k]*b[colval[k]] ascending; c = α s + β c.
 * With β == 0 (Julia's `false`, a strong zero) c is not read. */
void orc_spmv_csr(int64_t m, const int64_t* rowptr, const int32_t* colval, const double* val,
                  double alpha, const double* b, double beta, double* c) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < m; ++i) {
        double s = 0.0;
        for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k) s += val[k] * b[colval[k]];
        c[i] = (beta == 0.0) ? alpha * s : alpha * s + beta * c[i];
    }
}

/* mat_ell.jl:54-70: d = maximum row length */
int64_t orc_ell_width(int64_t m, const int64_t* rowptr) {
    int64_t d = 0;
    for (int64_t i = 0; i < m; ++i) {
        int64_t l = rowptr[i + 1] - rowptr[i];
        if (l > d) d = l;
    }
    return d;
}

/* mat_ell.jl:54-70: column-major m x d, left-packed in CSR order; the reference pads
 * with column 0 (1-based sentinel), here -1 (0-based sentinel), value 0. */
void orc_csr_to_ell(int64_t m, int64_t d, const int64_t* rowptr, const int32_t* colval,
                    const double* val, int32_t* ell_col, double* ell_val) {
    for (int64_t t = 0; t < m * d; ++t) { ell_col[t] = -1; ell_val[t] = 0.0; }
    for (int64_t i = 0; i < m; ++i)
        for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k) {
            int64_t s = k - rowptr[i];
            ell_col[s * m + i] = colval[k];
            ell_val[s * m + i] = val[k];
        }
}

/* ------------------------------------------------------------- scaling */

/* norm(view, p), linalg.jl:99 (LinearAlgebra norm on a vector view).
 * p = Inf: maximum of abs (order-free); p = 1: sequential sum of abs; p = 2: sqrt of
 * sequential sum of squares; otherwise (sum |v|^p)^(1/p). */
static double pnorm_acc(double acc, double v, double p) {
    double a = fabs(v);
    if (isinf(p)) return a > acc ? a : acc;
    if (p == 1.0) return acc + a;
    if (p == 2.0) return acc + a * a;
    return acc + pow(a, p);
}
static double pnorm_fin(double acc, double p) {
    if (isinf(p) || p == 1.0) return acc;
    if (p == 2.0) return sqrt(acc);
    return pow(acc, 1.0 / p);
}

/* diagonal_norm_preconditioner, preconditioning.jl:102-111:
 * col norms of A (CSC column j, rows ascending), row norms from At (row i of A,
 * columns ascending); d = iszero(norm) ? 1 : inv(sqrt(norm)). */
static void diag_norm_prec(int64_t m, int64_t n, const int64_t* colptr, const int64_t* rowptr,
                           const int64_t* perm, const double* a_val, const double* at_val,
                           double p_row, double p_col, double* d1, double* d2) {
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < n; ++j) {
        double acc = 0.0;
        for (int64_t k = colptr[j]; k < colptr[j + 1]; ++k) acc = pnorm_acc(acc, a_val[k], p_col);
        double cn = pnorm_fin(acc, p_col);
        d2[j] = cn == 0.0 ? 1.0 : 1.0 / sqrt(cn);
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < m; ++i) {
        double acc = 0.0;
        for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k) acc = pnorm_acc(acc, at_val[perm[k]], p_row);
        double rn = pnorm_fin(acc, p_row);
        d1[i] = rn == 0.0 ? 1.0 : 1.0 / sqrt(rn);
    }
}

/* precondition(cons, prec), preconditioning.jl:37-43: A <- D1*A*D2 evaluated as
 * (d1_i * a_ij) * d2_j; At <- D2*At*D1 evaluated as (d2_j * a_ji') * d1_i. */
static void apply_prec(int64_t n, const int64_t* colptr, const int32_t* rowval,
                       const double* a_in, const double* at_in, const double* d1,
                       const double* d2, double* a_out, double* at_out) {
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < n; ++j)
        for (int64_t k = colptr[j]; k < colptr[j + 1]; ++k) {
            int32_t i = rowval[k];
            a_out[k] = (d1[i] * a_in[k]) * d2[j];
            at_out[k] = (d2[j] * at_in[k]) * d1[i];
        }
}

/* pdlp_preconditioner, preconditioning.jl:146-155, with ruiz_preconditioner (:117-125)
 * and chambolle_pock_preconditioner (:113-115); then the matrix part of
 * precondition(milp, prec) (:61-91). */
void orc_scaling(int64_t m, int64_t n, const int64_t* colptr, const int32_t* rowval,
                 const double* nzval, int32_t ruiz_iter, double cp_alpha,
                 double* D1, double* D2, double* a_csc_out, double* at_csc_out) {
    int64_t nnz = colptr[n];
    int64_t* rowptr = (int64_t*)malloc((size_t)(m + 1) * sizeof(int64_t));
    int32_t* colval = (int32_t*)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(int32_t));
    int64_t* perm = (int64_t*)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(int64_t));
    csc_to_csr_perm(m, n, colptr, rowval, rowptr, colval, perm);
    double* a = (double*)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(double));
    double* at = (double*)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(double));
    double* d1 = (double*)malloc((size_t)(m > 0 ? m : 1) * sizeof(double));
    double* d2 = (double*)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    memcpy(a, nzval, (size_t)nnz * sizeof(double));
    memcpy(at, nzval, (size_t)nnz * sizeof(double));
    for (int64_t i = 0; i < m; ++i) D1[i] = 1.0;   /* identity_preconditioner, :95-100 */
    for (int64_t j = 0; j < n; ++j) D2[j] = 1.0;
    for (int32_t it = 0; it < ruiz_iter; ++it) {
        diag_norm_prec(m, n, colptr, rowptr, perm, a, at, INFINITY, INFINITY, d1, d2);
        apply_prec(n, colptr, rowval, a, at, d1, d2, a, at);
        for (int64_t i = 0; i < m; ++i) D1[i] = d1[i] * D1[i];   /* prec_next * prec */
        for (int64_t j = 0; j < n; ++j) D2[j] = D2[j] * d2[j];
    }
    /* cons_r = precondition(cons, prec_r) from the ORIGINAL matrices; cp_alpha = NaN
     * skips Chambolle-Pock (test hook: Ruiz alone, as ruiz_preconditioner) */
    if (!isnan(cp_alpha)) {
        apply_prec(n, colptr, rowval, nzval, nzval, D1, D2, a, at);
        diag_norm_prec(m, n, colptr, rowptr, perm, a, at, 2.0 - cp_alpha, cp_alpha, d1, d2);
        for (int64_t i = 0; i < m; ++i) D1[i] = D1[i] * d1[i];   /* prec_r * prec_cp */
        for (int64_t j = 0; j < n; ++j) D2[j] = d2[j] * D2[j];
    }
    /* precondition(milp, prec): A_p = D1*A*D2, At_p = D2*At*D1 from the originals */
    apply_prec(n, colptr, rowval, nzval, nzval, D1, D2, a_csc_out, at_csc_out);
    free(rowptr); free(colval); free(perm); free(a); free(at); free(d1); free(d2);
}

/* ------------------------------------------------------------- KKT */

typedef struct {
    int64_t m, n;
    const int64_t *a_rowptr; const int32_t* a_col; const double* a_val;
    const int64_t *at_rowptr; const int32_t* at_col; const double* at_val;
    const double *c, *lv, *uv, *lc, *uc, *D1, *D2;
    double *sx, *sy, *sr;   /* scratch n, m, n */
} orc_lp;

/* kkt_errors!, errors.jl:65-109 */
static void kkt_errors(const orc_lp* L, const double* x, const double* y, orc_kkt* out) {
    int64_t m = L->m, n = L->n;
    double *Ax = L->sy, *Aty = L->sx, *r = L->sr;
    orc_spmv_csr(m, L->a_rowptr, L->a_col, L->a_val, 1.0, x, 0.0, Ax);
    orc_spmv_csr(n, L->at_rowptr, L->at_col, L->at_val, 1.0, y, 0.0, Aty);
    double* tmp_m = (double*)malloc((size_t)(m > 0 ? m : 1) * sizeof(double));
    double* tmp_n = (double*)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < n; ++j) r[j] = proj_multiplier(L->c[j] - Aty[j], L->lv[j], L->uv[j]);
    /* primal */
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < m; ++i)
        tmp_m[i] = (1.0 / L->D1[i]) * (Ax[i] - proj_box(Ax[i], L->lc[i], L->uc[i]));
    double primal = nrm2(m, tmp_m);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < m; ++i) tmp_m[i] = (1.0 / L->D1[i]) * combine(L->lc[i], L->uc[i]);
    double primal_scale = 1.0 + nrm2(m, tmp_m);
    /* dual */
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < n; ++j) tmp_n[j] = (1.0 / L->D2[j]) * ((L->c[j] - Aty[j]) - r[j]);
    double dual = nrm2(n, tmp_n);
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < n; ++j) tmp_n[j] = (1.0 / L->D2[j]) * L->c[j];
    double dual_scale = 1.0 + nrm2(n, tmp_n);
    /* gap */
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < m; ++i)
        tmp_m[i] = safeprod_left(L->uc[i], pos_part(-y[i])) - safeprod_left(L->lc[i], neg_part(-y[i]));
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < n; ++j)
        tmp_n[j] = safeprod_left(L->uv[j], pos_part(-r[j])) - safeprod_left(L->lv[j], neg_part(-r[j]));
    double pc_sum = sumv(m, tmp_m), pv_sum = sumv(n, tmp_n), cx = dotv(n, L->c, x);
    out->primal = primal;
    out->primal_scale = primal_scale;
    out->dual = dual;
    out->dual_scale = dual_scale;
    out->gap = fabs(cx + pc_sum + pv_sum);
    out->gap_scale = 1.0 + fabs(pc_sum + pv_sum) + fabs(cx);
    free(tmp_m); free(tmp_n);
}

void orc_kkt_errors(int64_t m, int64_t n,
                    const int64_t* a_rowptr, const int32_t* a_col, const double* a_val,
                    const int64_t* at_rowptr, const int32_t* at_col, const double* at_val,
                    const double* c, const double* lv, const double* uv,
                    const double* lc, const double* uc, const double* D1, const double* D2,
                    const double* x, const double* y, orc_kkt* out) {
    orc_lp L = {m, n, a_rowptr, a_col, a_val, at_rowptr, at_col, at_val, c, lv, uv, lc, uc, D1, D2, 0, 0, 0};
    L.sx = (double*)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    L.sr = (double*)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    L.sy = (double*)malloc((size_t)(m > 0 ? m : 1) * sizeof(double));
    kkt_errors(&L, x, y, out);
    free(L.sx); free(L.sr); free(L.sy);
}

/* errors.jl:55-63 */
static double relative(const orc_kkt* e) {
    return jmax(jmax(e->primal / e->primal_scale, e->dual / e->dual_scale), e->gap / e->gap_scale);
}
static double absolute(const orc_kkt* e, double w) {
    return sqrt(w * w * e->primal * e->primal + (1.0 / (w * w)) * e->dual * e->dual + e->gap * e->gap);
}

/* ------------------------------------------------------------- power method */

static inline uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* Start vector of the power method.  The reference uses randn!(StableRNG(0), x0)
 * (linalg.jl:91-92); that stream needs Julia, so this is an approximately normal
 * start built from integer hashing and exact additions only: identical bits on
 * any IEEE device. */
void orc_powm_start(int64_t n, uint64_t seed, double* x0) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        double s = 0.0;
        for (int r = 0; r < 12; ++r) {
            uint64_t h = splitmix64(seed * 0x100000001B3ull ^ splitmix64((uint64_t)i * 16u + (uint64_t)r));
            s += (double)(h >> 11) * 0x1.0p-53;
        }
        x0[i] = s - 6.0;
    }
}

/* spectral_norm(K, Kt) = sqrt(λ) with powm!(KtK, x0): IterativeSolvers 0.9 power
 * method as called from linalg.jl:86-96: x normalised; per iteration
 * Ax = KtK x; θ = x·Ax; r = Ax - θx; residual = ‖r‖; x = Ax * (1/‖Ax‖);
 * stop before an iteration when iteration > maxiter or residual <= tol*|θ|. */
double orc_spectral_norm(int64_t m, int64_t n,
                         const int64_t* a_rowptr, const int32_t* a_col, const double* a_val,
                         const int64_t* at_rowptr, const int32_t* at_col, const double* at_val,
                         double tol, int64_t maxiter, uint64_t seed, int64_t* iters_out) {
    double* x = (double*)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    double* Ax = (double*)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    double* t = (double*)malloc((size_t)(m > 0 ? m : 1) * sizeof(double));
    orc_powm_start(n, seed, x);
    double inv0 = 1.0 / nrm2(n, x);
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < n; ++j) x[j] = x[j] * inv0;
    double theta = 0.0, residual = DBL_MAX;
    int64_t it = 0;
    while (!(it > maxiter || residual <= tol * fabs(theta))) {
        orc_spmv_csr(m, a_rowptr, a_col, a_val, 1.0, x, 0.0, t);
        orc_spmv_csr(n, at_rowptr, at_col, at_val, 1.0, t, 0.0, Ax);
        theta = dotv(n, x, Ax);
        double acc = 0.0;
        {
            int64_t nb = (n + RBLK - 1) / RBLK;
            double* part = (double*)malloc((size_t)(nb > 0 ? nb : 1) * sizeof(double));
#pragma omp parallel for schedule(static)
            for (int64_t b = 0; b < nb; ++b) {
                int64_t lo = b * RBLK, hi = lo + RBLK < n ? lo + RBLK : n;
                double s = 0.0;
                for (int64_t j = lo; j < hi; ++j) {
                    double rj = Ax[j] + (-theta) * x[j];
                    s += rj * rj;
                }
                part[b] = s;
            }
            for (int64_t b = 0; b < nb; ++b) acc += part[b];
            free(part);
        }
        residual = sqrt(acc);
        double inv = 1.0 / nrm2(n, Ax);
#pragma omp parallel for schedule(static)
        for (int64_t j = 0; j < n; ++j) x[j] = Ax[j] * inv;
        ++it;
    }
    if (iters_out) *iters_out = it;
    free(x); free(Ax); free(t);
    return sqrt(theta);
}

/* tol = eps(Float64) * n^3 with Int64 wraparound; maxiter = n (linalg.jl:72, :94) */
static double powm_default_tol(int64_t n) {
    uint64_t u = (uint64_t)n;
    int64_t n3 = (int64_t)(u * u * u);
    return DBL_EPSILON * (double)n3;
}

/* ------------------------------------------------------------- solve */

typedef struct { double *x, *y; } pd;

static void record(orc_stats* st, orc_hist* hist, int64_t cap, int rec) {
    if (!rec) return;
    if (st->n_history < cap) {
        hist[st->n_history].kkt_passes = st->kkt_passes;
        hist[st->n_history].err = st->err;
    }
    st->n_history++;
}

/* Opt-in infeasibility test (no counterpart in the reference), on the ray
 * (dx, dy) in scaled space.  In original units x = D2 x_s, y = D1 y_s, A x = A_s x_s / D1,
 * A'y = A_s'y_s / D2; a bound times a ray component is the same in either space.
 * Dual ray e = dy, r = -A'e: Farkas objective lc e+ - uc e- + lv r+ - uv r- over finite
 * bounds, with sign violations e+ (lc = -inf), e- (uc = +inf), r+ (lv = -inf),
 * r- (uv = +inf); primal infeasible when objective > 0 and |violations| <= tol * objective.
 * Primal ray d = dx: c'd < 0 and |violation of A d in the recession cone of [lc, uc] and
 * of d in that of [lv, uv]| <= tol * |c'd| gives dual infeasible.  Returns 4, 5 or 3. */
static int32_t ray_test(const orc_lp* L, const double* dx, const double* dy, double tol) {
    int64_t m = L->m, n = L->n;
    double* Ad = (double*)malloc((size_t)(m > 0 ? m : 1) * sizeof(double));
    double* Ate = (double*)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    orc_spmv_csr(m, L->a_rowptr, L->a_col, L->a_val, 1.0, dx, 0.0, Ad);
    orc_spmv_csr(n, L->at_rowptr, L->at_col, L->at_val, 1.0, dy, 0.0, Ate);
    double obj = 0.0, sviol = 0.0, cd = 0.0, cviol = 0.0;
    for (int64_t i = 0; i < m; ++i) {
        int lf = isfinite(L->lc[i]), uf = isfinite(L->uc[i]);
        double e = dy[i], ep = pos_part(e), en = neg_part(e);
        obj += (lf ? L->lc[i] * ep : 0.0) - (uf ? L->uc[i] * en : 0.0);
        double sv = ((lf ? 0.0 : ep) + (uf ? 0.0 : en)) * L->D1[i];
        sviol += sv * sv;
        double v = Ad[i] / L->D1[i];
        double cv = lf && uf ? v : lf ? fmin(v, 0.0) : uf ? fmax(v, 0.0) : 0.0;
        cviol += cv * cv;
    }
    for (int64_t j = 0; j < n; ++j) {
        int lf = isfinite(L->lv[j]), uf = isfinite(L->uv[j]);
        double r = -Ate[j], rp = pos_part(r), rn = neg_part(r);
        obj += (lf ? L->lv[j] * rp : 0.0) - (uf ? L->uv[j] * rn : 0.0);
        double sv = ((lf ? 0.0 : rp) + (uf ? 0.0 : rn)) / L->D2[j];
        sviol += sv * sv;
        double d = dx[j];
        cd += L->c[j] * d;
        double cv = (lf && uf ? d : lf ? fmin(d, 0.0) : uf ? fmax(d, 0.0) : 0.0) * L->D2[j];
        cviol += cv * cv;
    }
    free(Ad);
    free(Ate);
    if (obj > 0.0 && sqrt(sviol) <= tol * obj) return 4;
    if (cd < 0.0 && sqrt(cviol) <= tol * -cd) return 5;
    return 3;
}

int32_t orc_ray_test(int64_t m, int64_t n, const int64_t* a_rowptr, const int32_t* a_col,
                     const double* a_val, const int64_t* at_rowptr, const int32_t* at_col,
                     const double* at_val, const double* c, const double* lv, const double* uv,
                     const double* lc, const double* uc, const double* D1, const double* D2,
                     const double* dx, const double* dy, double tol) {
    orc_lp L = {m, n, a_rowptr, a_col, a_val, at_rowptr, at_col, at_val, c, lv, uv, lc, uc, D1, D2, 0, 0, 0};
    return ray_test(&L, dx, dy, tol);
}

static int32_t infeasibility_status(const orc_lp* L, const orc_params* p, const orc_kkt* err,
                                    const pd* sol, const pd* rst, const pd* last) {
    if (!isfinite(relative(err))) return 6;
    int64_t m = L->m, n = L->n;
    double* dx = (double*)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    double* dy = (double*)malloc((size_t)(m > 0 ? m : 1) * sizeof(double));
    int32_t st = 3;
    for (int k = 0; k < 3 && st == 3; ++k) {
        if (!(p->ray_candidates >> k & 1)) continue;
        const pd* o = k == 0 ? rst : last;
        for (int64_t j = 0; j < n; ++j) dx[j] = k == 1 ? sol->x[j] : sol->x[j] - o->x[j];
        for (int64_t i = 0; i < m; ++i) dy[i] = k == 1 ? sol->y[i] : sol->y[i] - o->y[i];
        st = ray_test(L, dx, dy, p->infeasibility_tol);
    }
    free(dx);
    free(dy);
    return st;
}

/* termination_status, termination.jl:89-101 */
static int32_t term_status(const orc_stats* st, const orc_params* p) {
    if (relative(&st->err) <= p->termination_reltol) return 0;
    if (st->time_elapsed >= p->time_limit) return 1;
    if (st->kkt_passes >= p->max_kkt_passes) return 2;
    return 3;
}

/* restart_check!, pdlp.jl:149-165 (candidate: the iterate if strictly better than the
 * average, else the average; the same choice one step earlier, F10), then should_restart,
 * restart.jl:40-57 */
int32_t orc_restart_decision(const orc_kkt* e_sol, const orc_kkt* e_avg, const orc_kkt* e_last,
                             const orc_kkt* e_avg_last, const orc_kkt* e_restart, double omega,
                             int64_t inner, int64_t total, const orc_params* p, int32_t* from_avg) {
    int32_t fa = !(absolute(e_sol, omega) < absolute(e_avg, omega));
    const orc_kkt* cand = fa ? e_avg : e_sol;
    const orc_kkt* cand_last = absolute(e_last, omega) < absolute(e_avg_last, omega) ? e_last : e_avg_last;
    double ac = absolute(cand, omega), ar = absolute(e_restart, omega), acl = absolute(cand_last, omega);
    int sufficient = ac <= p->sufficient_decay * ar;
    int necessary = ac <= p->necessary_decay * ar;
    int no_local = ac > acl;
    int long_inner = (double)inner >= p->artificial_decay * (double)total;
    if (from_avg) *from_avg = fa;
    return (sufficient || (necessary && no_local) || long_inner) ? 1 : 0;
}

/* primal_weight_update!, step_size.jl:58-75 */
double orc_primal_weight_update(double omega, double dx, double dy, double damping, double zero_tol) {
    if (dx > zero_tol && dy > zero_tol)
        return exp(damping * log(dy / dx) + (1.0 - damping) * log(omega));
    return omega;
}

int orc_solve(int64_t m, int64_t n, const int64_t* colptr, const int32_t* rowval,
              const double* nzval, const double* c_in, const double* lv_in, const double* uv_in,
              const double* lc_in, const double* uc_in, const double* x0, const double* y0,
              const orc_params* p, double* x_out, double* y_out, orc_stats* st,
              orc_hist* hist, int64_t hist_cap) {
    return orc_solve_traced(m, n, colptr, rowval, nzval, c_in, lv_in, uv_in, lc_in, uc_in, x0, y0, p,
                            x_out, y_out, st, hist, hist_cap, NULL);
}

int orc_solve_traced(int64_t m, int64_t n, const int64_t* colptr, const int32_t* rowval,
                     const double* nzval, const double* c_in, const double* lv_in, const double* uv_in,
                     const double* lc_in, const double* uc_in, const double* x0, const double* y0,
                     const orc_params* p, double* x_out, double* y_out, orc_stats* st,
                     orc_hist* hist, int64_t hist_cap, orc_trace* tr) {
    double t_start = now_s();                         /* common.jl:191 */
    int64_t nnz = colptr[n];
    size_t nz1 = (size_t)(nnz > 0 ? nnz : 1), m1 = (size_t)(m > 0 ? m : 1), n1 = (size_t)(n > 0 ? n : 1);
    memset(st, 0, sizeof(*st));
    st->err.primal = st->err.primal_scale = st->err.dual = NAN;
    st->err.dual_scale = st->err.gap = st->err.gap_scale = NAN;
    st->termination_status = 3;

    /* ---- preprocess (common.jl:154-169): scaling on the host in FP64 */
    double* D1 = (double*)malloc(m1 * sizeof(double));
    double* D2 = (double*)malloc(n1 * sizeof(double));
    double* a_csc = (double*)malloc(nz1 * sizeof(double));
    double* at_csc = (double*)malloc(nz1 * sizeof(double));
    orc_scaling(m, n, colptr, rowval, nzval, p->ruiz_iter, p->chambolle_pock_alpha, D1, D2, a_csc, at_csc);
    double* c = (double*)malloc(n1 * sizeof(double));
    double* lv = (double*)malloc(n1 * sizeof(double));
    double* uv = (double*)malloc(n1 * sizeof(double));
    double* lc = (double*)malloc(m1 * sizeof(double));
    double* uc = (double*)malloc(m1 * sizeof(double));
    for (int64_t j = 0; j < n; ++j) { c[j] = D2[j] * c_in[j]; lv[j] = lv_in[j] / D2[j]; uv[j] = uv_in[j] / D2[j]; }
    for (int64_t i = 0; i < m; ++i) { lc[i] = D1[i] * lc_in[i]; uc[i] = D1[i] * uc_in[i]; }
    /* CSR(A_p) = stable transpose of CSC(A_p); CSR(At_p) = CSC(A) arrays with At_p values */
    int64_t* a_rowptr = (int64_t*)malloc((size_t)(m + 1) * sizeof(int64_t));
    int32_t* a_col = (int32_t*)malloc(nz1 * sizeof(int32_t));
    double* a_val = (double*)malloc(nz1 * sizeof(double));
    orc_csc_to_csr(m, n, colptr, rowval, a_csc, a_rowptr, a_col, a_val);
    orc_lp L = {m, n, a_rowptr, a_col, a_val, colptr, rowval, at_csc, c, lv, uv, lc, uc, D1, D2, 0, 0, 0};
    L.sx = (double*)malloc(n1 * sizeof(double));
    L.sr = (double*)malloc(n1 * sizeof(double));
    L.sy = (double*)malloc(m1 * sizeof(double));

    /* solution buffers: sol, sol_last, avg, avg_last, restart */
    double* buf = (double*)calloc(5 * (n1 + m1), sizeof(double));
    pd sol = {buf, buf + n1}, sol_last = {buf + n1 + m1, buf + 2 * n1 + m1};
    pd avg = {buf + 2 * (n1 + m1), buf + 3 * n1 + 2 * m1};
    pd avg_last = {buf + 3 * (n1 + m1), buf + 4 * n1 + 3 * m1};
    pd rst = {buf + 4 * (n1 + m1), buf + 5 * n1 + 4 * m1};
    for (int64_t j = 0; j < n; ++j) sol.x[j] = x0 ? x0[j] / D2[j] : 0.0;   /* precondition(sol) */
    for (int64_t i = 0; i < m; ++i) sol.y[i] = y0 ? y0[i] / D1[i] : 0.0;
    int pdlp = p->algorithm == 0;
    if (pdlp) {                                       /* pdlp.jl:51-54 */
        memcpy(avg.x, sol.x, (size_t)n * sizeof(double)); memcpy(avg.y, sol.y, (size_t)m * sizeof(double));
    }
    /* restart = sol (PDHG keeps it as the start point: the infeasibility test's origin) */
    memcpy(rst.x, sol.x, (size_t)n * sizeof(double)); memcpy(rst.y, sol.y, (size_t)m * sizeof(double));

    /* ---- initialize: η (step_size.jl:22-27) and ω (:29-40; PDHG ω = 1, pdhg.jl:48) */
    double tol = isnan(p->powm_tol) ? powm_default_tol(n) : p->powm_tol;
    int64_t maxiter = p->powm_maxiter < 0 ? n : p->powm_maxiter;
    double snorm = orc_spectral_norm(m, n, a_rowptr, a_col, a_val, colptr, rowval, at_csc,
                                     tol, maxiter, p->powm_seed, &st->powm_iters);
    double eta = p->invnorm_scaling * (1.0 / snorm);
    double omega = 1.0;
    if (pdlp) {
        double cn = nrm2(n, c);
        double* cb = (double*)malloc(m1 * sizeof(double));
        for (int64_t i = 0; i < m; ++i) cb[i] = combine(lc[i], uc[i]);
        double bn = nrm2(m, cb);
        free(cb);
        omega = (cn > p->zero_tol && bn > p->zero_tol) ? cn / bn : 1.0;
    }
    double eta_sum = 0.0;
    int64_t inner = 0, total = 0;
    st->setup_seconds = now_s() - t_start;

    /* early exit (common.jl:194-198) */
    int all_zero_c = 1;
    for (int64_t j = 0; j < n; ++j) if (c[j] != 0.0) { all_zero_c = 0; break; }
    if (m == 0 && all_zero_c) {
        for (int64_t j = 0; j < n; ++j) sol.x[j] = proj_box(0.0, lv[j], uv[j]);
        st->termination_status = 0;
        goto finish;
    }

    for (;;) {                                         /* pdlp.jl:75-91 / pdhg.jl:62-73 */
        int terminated = 0;
        for (;;) {
            for (int32_t s = 0; s < p->check_every; ++s) {
                /* step!, pdlp.jl:96-124 */
                pd tmp = sol; sol = sol_last; sol_last = tmp;
                const double *x = sol_last.x, *y = sol_last.y;
                double tau = eta / omega, sigma = eta * omega, isig = 1.0 / sigma;
                orc_spmv_csr(n, colptr, rowval, at_csc, 1.0, y, 0.0, L.sx);
#pragma omp parallel for schedule(static)
                for (int64_t j = 0; j < n; ++j) {
                    double xp = proj_box(x[j] - tau * (c[j] - L.sx[j]), lv[j], uv[j]);
                    sol.x[j] = xp;
                    L.sx[j] = 2.0 * xp - x[j];
                }
                orc_spmv_csr(m, a_rowptr, a_col, a_val, 1.0, L.sx, 0.0, L.sy);
#pragma omp parallel for schedule(static)
                for (int64_t i = 0; i < m; ++i) {
                    double ax = L.sy[i];
                    sol.y[i] = (y[i] - sigma * ax) - sigma * proj_box(isig * y[i] - ax, -uc[i], -lc[i]);
                }
                st->kkt_passes += 1;
                if (pdlp) {
                    /* update_average!, pdlp.jl:126-136 */
                    memcpy(avg_last.x, avg.x, (size_t)n * sizeof(double));
                    memcpy(avg_last.y, avg.y, (size_t)m * sizeof(double));
                    double wa = eta / (eta + eta_sum), wb = eta_sum / (eta + eta_sum);
#pragma omp parallel for schedule(static)
                    for (int64_t j = 0; j < n; ++j) avg.x[j] = wa * sol.x[j] + wb * avg.x[j];
#pragma omp parallel for schedule(static)
                    for (int64_t i = 0; i < m; ++i) avg.y[i] = wa * sol.y[i] + wb * avg.y[i];
                    eta_sum += eta;
                    inner += 1; total += 1;             /* add_inner!, iteration.jl:7-11 */
                }
            }
            /* termination_check!, common.jl:218-231 */
            st->time_elapsed = now_s() - t_start;
            kkt_errors(&L, sol.x, sol.y, &st->err);
            record(st, hist, hist_cap, p->record_error_history);
            st->termination_status = term_status(st, p);
            if (p->detect_infeasibility && st->termination_status != 0) {
                int32_t inf = infeasibility_status(&L, p, &st->err, &sol, &rst, &sol_last);
                if (inf != 3) { st->termination_status = inf; goto finish; }
            }
            int64_t ti = tr ? tr->count : 0;
            int trace_this = tr && ti < tr->cap;
            if (tr) tr->count += 1;
            if (trace_this) {
                const orc_kkt nan_err = {NAN, NAN, NAN, NAN, NAN, NAN};
                tr->passes[ti] = st->kkt_passes;
                tr->err[5 * ti] = st->err;
                for (int q = 1; q < 5; ++q) tr->err[5 * ti + q] = nan_err;
                tr->from_avg[ti] = 0;
                tr->restarted[ti] = 0;
                tr->omega[ti] = tr->omega_after[ti] = omega;
                tr->inner[ti] = inner;
                tr->total[ti] = total;
            }
            if (st->termination_status != 3) { terminated = 1; break; }
            if (!pdlp) continue;
            /* restart_check!, pdlp.jl:138-170 */
            orc_kkt e_s = st->err, e_a, e_l, e_al, e_r;
            kkt_errors(&L, avg.x, avg.y, &e_a);
            kkt_errors(&L, sol_last.x, sol_last.y, &e_l);
            kkt_errors(&L, avg_last.x, avg_last.y, &e_al);
            kkt_errors(&L, rst.x, rst.y, &e_r);
            int32_t from_avg = 0;
            int32_t restart = orc_restart_decision(&e_s, &e_a, &e_l, &e_al, &e_r, omega, inner, total, p,
                                                   &from_avg);
            if (trace_this) {
                tr->err[5 * ti + 1] = e_a;
                tr->err[5 * ti + 2] = e_l;
                tr->err[5 * ti + 3] = e_al;
                tr->err[5 * ti + 4] = e_r;
                tr->from_avg[ti] = from_avg;
                tr->restarted[ti] = restart;
            }
            if (restart) {
                /* restart!, pdlp.jl:172-196 */
                pd sc = from_avg ? avg : sol;
                eta_sum = 0.0;
                double dx = dist2(n, sc.x, rst.x), dy = dist2(m, sc.y, rst.y);
                omega = orc_primal_weight_update(omega, dx, dy, p->primal_weight_damping, p->zero_tol);
                if (trace_this) tr->omega_after[ti] = omega;
                if (from_avg) {
                    memcpy(sol.x, avg.x, (size_t)n * sizeof(double));
                    memcpy(sol.y, avg.y, (size_t)m * sizeof(double));
                }
                memset(avg.x, 0, (size_t)n * sizeof(double));
                memset(avg.y, 0, (size_t)m * sizeof(double));
                memcpy(rst.x, sol.x, (size_t)n * sizeof(double));
                memcpy(rst.y, sol.y, (size_t)m * sizeof(double));
                inner = 0;                                 /* add_outer! */
                st->restarts += 1;
            }
        }
        if (!pdlp) break;  /* PDHG's loop ends at its single termination check */
        /* second termination_check! after the inner loop, pdlp.jl:88-90 */
        st->time_elapsed = now_s() - t_start;
        kkt_errors(&L, sol.x, sol.y, &st->err);
        record(st, hist, hist_cap, p->record_error_history);
        st->termination_status = term_status(st, p);
        if (terminated && st->termination_status != 3) break;
    }

finish:
    /* get_solution: unprecondition, preconditioning.jl:53-59 */
    for (int64_t j = 0; j < n; ++j) x_out[j] = D2[j] * sol.x[j];
    for (int64_t i = 0; i < m; ++i) y_out[i] = D1[i] * sol.y[i];
    if (tr && tr->x_avg) for (int64_t j = 0; j < n; ++j) tr->x_avg[j] = D2[j] * avg.x[j];
    if (tr && tr->y_avg) for (int64_t i = 0; i < m; ++i) tr->y_avg[i] = D1[i] * avg.y[i];
    st->eta = eta;
    st->omega = omega;
    free(buf); free(L.sx); free(L.sr); free(L.sy);
    free(a_rowptr); free(a_col); free(a_val);
    free(c); free(lv); free(uv); free(lc); free(uc);
    free(D1); free(D2); free(a_csc); free(at_csc);
    return 0;
}
