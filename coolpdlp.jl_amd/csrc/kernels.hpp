// kernels.hpp -- device kernels and epilogues of the PDLP/PDHG loop.
//
// Step epilogues follow the reference's broadcasts operation by operation
// (src/algorithms/pdlp.jl:110-117, pdhg.jl:286-293) and the averaging axpby!
// (pdlp.jl:126-136), fused into the row loop of the product that feeds them.  KKT
// epilogues follow kkt_errors! (src/components/errors.jl:65-109); they produce raw
// sums, from which the host forms the six KKTErrors fields.
#pragma once
#include "spmv.hpp"

namespace cpdlp {

// per-step scalars of one check_every chunk (device table, one entry per step)
struct StepCoef {
    double tau, sigma, isig;   // tau = eta/omega, sigma = eta*omega, isig = inv(sigma)
    double wa, wb;             // averaging weights eta/(eta+eta_sum), eta_sum/(eta+eta_sum)
    int32_t last;              // last step of the chunk: keep avg_last
    int32_t pad;
};

// ---------------------------------------------------------------- step epilogues
// primal half-step on the rows of A' (= columns j):
//   x+ = proj_box(x - tau*(c - A'y), lv, uv); xbar = 2x+ - x; avg update
template <class V>
struct EpiPrimal {
    static constexpr int NQ = 0;
    double* partials = nullptr;
    const V *x, *c, *lv, *uv;
    V *xn, *xbar, *xavg, *xavgl;
    const StepCoef* coef;
    int32_t step, pdlp;
    __device__ bool skip() const { return false; }
    template <class Acc>
    __device__ __forceinline__ void row(int j, const Acc* acc, double*) const {
        const StepCoef& cf = coef[step];
        const V tau = (V)cf.tau;
        const V xj = x[j];
        const V xp = proj_box<V>(xj - tau * (c[j] - (V)acc[0]), lv[j], uv[j]);
        xn[j] = xp;
        xbar[j] = V(2) * xp - xj;
        if (pdlp) {
            const V a = xavg[j];
            if (cf.last) xavgl[j] = a;
            xavg[j] = (V)cf.wa * xp + (V)cf.wb * a;
        }
    }
};

// dual half-step on the rows of A:
//   y+ = (y - sigma*A xbar) - sigma*proj_box(inv(sigma)*y - A xbar, -uc, -lc); avg update
template <class V>
struct EpiDual {
    static constexpr int NQ = 0;
    double* partials = nullptr;
    const V *y, *lc, *uc;
    V *yn, *yavg, *yavgl;
    const StepCoef* coef;
    int32_t step, pdlp;
    __device__ bool skip() const { return false; }
    template <class Acc>
    __device__ __forceinline__ void row(int i, const Acc* acc, double*) const {
        const StepCoef& cf = coef[step];
        const V sigma = (V)cf.sigma, isig = (V)cf.isig;
        const V yi = y[i], ax = (V)acc[0];
        const V yp = (yi - sigma * ax) - sigma * proj_box<V>(isig * yi - ax, -uc[i], -lc[i]);
        yn[i] = yp;
        if (pdlp) {
            const V a = yavg[i];
            if (cf.last) yavgl[i] = a;
            yavg[i] = (V)cf.wa * yp + (V)cf.wb * a;
        }
    }
};

// ---------------------------------------------------------------- KKT epilogues
// rows of A, K candidates: sum of (inv(D1)*(Ax - proj_box(Ax, lc, uc)))^2 and of
// pc = safeprod_left(uc, (-y)+) - safeprod_left(lc, (-y)-); plus |y_k - y_restart|^2
// for candidates 0 and 1 (the primal-weight distances of step_size.jl:58-75).
template <class V, int K>
struct EpiKktRows {
    static constexpr int NQ = 2 * K + 2;
    double* partials;
    const V* Y[K];
    const V* y_rst;
    const double *lc, *uc, *invD1;
    int32_t dist;
    __device__ bool skip() const { return false; }
    template <class Acc>
    __device__ __forceinline__ void row(int i, const Acc* acc, double* q) const {
        const double l = lc[i], u = uc[i], id = invD1[i];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const double ax = (double)acc[k];
            const double pd = id * (ax - proj_box(ax, l, u));
            q[2 * k] += pd * pd;
            const double yk = (double)Y[k][i];
            q[2 * k + 1] += safeprod_left(u, pos_part(-yk)) - safeprod_left(l, neg_part(-yk));
        }
        if (dist) {
            const double yr = (double)y_rst[i];
            const double d0 = (double)Y[0][i] - yr, d1 = (double)Y[K > 1 ? 1 : 0][i] - yr;
            q[2 * K] += d0 * d0;
            q[2 * K + 1] += d1 * d1;
        }
    }
};

// rows of A' (= columns j), K candidates: r = proj_multiplier(c - A'y, lv, uv);
// sums of (inv(D2)*((c - A'y) - r))^2, of pv = safeprod_left(uv, (-r)+) -
// safeprod_left(lv, (-r)-), of c_j*x_j; plus |x_k - x_restart|^2 for k = 0, 1.
template <class V, int K>
struct EpiKktCols {
    static constexpr int NQ = 3 * K + 2;
    double* partials;
    const V* X[K];
    const V* x_rst;
    const double *c, *lv, *uv, *invD2;
    int32_t dist;
    __device__ bool skip() const { return false; }
    template <class Acc>
    __device__ __forceinline__ void row(int j, const Acc* acc, double* q) const {
        const double cj = c[j], l = lv[j], u = uv[j], id = invD2[j];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const double g = cj - (double)acc[k];
            const double r = proj_multiplier(g, l, u);
            const double dd = id * (g - r);
            q[3 * k] += dd * dd;
            q[3 * k + 1] += safeprod_left(u, pos_part(-r)) - safeprod_left(l, neg_part(-r));
            q[3 * k + 2] += cj * (double)X[k][j];
        }
        if (dist) {
            const double xr = (double)x_rst[j];
            const double d0 = (double)X[0][j] - xr, d1 = (double)X[K > 1 ? 1 : 0][j] - xr;
            q[3 * K] += d0 * d0;
            q[3 * K + 1] += d1 * d1;
        }
    }
};

// ---------------------------------------------------------------- power method
struct PowState {
    double theta, nrm, residual, tol;
    int64_t iter, maxiter;
    int32_t done, pad;
};

template <class V>
struct EpiStore {   // t = A x (skipped once the power method has converged)
    static constexpr int NQ = 0;
    double* partials = nullptr;
    V* out;
    const PowState* st;
    __device__ bool skip() const { return st && st->done; }
    template <class Acc>
    __device__ __forceinline__ void row(int i, const Acc* acc, double*) const { out[i] = (V)acc[0]; }
};

template <class V>
struct EpiPowAt {   // Ax = A' t; sums of x.Ax and Ax.Ax
    static constexpr int NQ = 2;
    double* partials;
    V* Ax;
    const V* x;
    const PowState* st;
    __device__ bool skip() const { return st->done; }
    template <class Acc>
    __device__ __forceinline__ void row(int j, const Acc* acc, double* q) const {
        const V a = (V)acc[0];
        Ax[j] = a;
        q[0] += (double)(x[j] * a);
        q[1] += (double)(a * a);
    }
};

// plain mul!: out[idx[r]] = alpha*acc + beta*out[idx[r]] (beta == 0: out not read)
struct EpiAxpby {
    static constexpr int NQ = 0;
    double* partials = nullptr;
    double* out;
    const int32_t* out_index;
    double alpha, beta;
    __device__ bool skip() const { return false; }
    template <class Acc>
    __device__ __forceinline__ void row(int r, const Acc* acc, double*) const {
        const int i = out_index ? out_index[r] : r;
        const double s = (double)acc[0];
        out[i] = beta == 0.0 ? alpha * s : alpha * s + beta * out[i];
    }
};

}  // namespace cpdlp
