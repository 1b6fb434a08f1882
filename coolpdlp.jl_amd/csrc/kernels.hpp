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
    struct Pre {
        V x, c, lv, uv, avg;
    };
    __device__ bool skip() const { return false; }
    __device__ __forceinline__ Pre load(int j) const {
        Pre p;
        p.x = x[j];
        p.c = c[j];
        p.lv = lv[j];
        p.uv = uv[j];
        p.avg = pdlp ? xavg[j] : V(0);
        return p;
    }
    template <class Acc>
    __device__ __forceinline__ void row(int j, const Acc* acc, double*, const Pre& p) const {
        const StepCoef& cf = coef[step];
        const V tau = (V)cf.tau;
        const V xp = proj_box<V>(p.x - tau * (p.c - (V)acc[0]), p.lv, p.uv);
        xn[j] = xp;
        xbar[j] = V(2) * xp - p.x;
        if (pdlp) {
            if (cf.last) xavgl[j] = p.avg;
            xavg[j] = (V)cf.wa * xp + (V)cf.wb * p.avg;
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
    struct Pre {
        V y, lc, uc, avg;
    };
    __device__ bool skip() const { return false; }
    __device__ __forceinline__ Pre load(int i) const {
        Pre p;
        p.y = y[i];
        p.lc = lc[i];
        p.uc = uc[i];
        p.avg = pdlp ? yavg[i] : V(0);
        return p;
    }
    template <class Acc>
    __device__ __forceinline__ void row(int i, const Acc* acc, double*, const Pre& p) const {
        const StepCoef& cf = coef[step];
        const V sigma = (V)cf.sigma, isig = (V)cf.isig;
        const V ax = (V)acc[0];
        const V yp = (p.y - sigma * ax) - sigma * proj_box<V>(isig * p.y - ax, -p.uc, -p.lc);
        yn[i] = yp;
        if (pdlp) {
            if (cf.last) yavgl[i] = p.avg;
            yavg[i] = (V)cf.wa * yp + (V)cf.wb * p.avg;
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
    struct Pre {
        double l, u, id, yr;
        double y[K];
    };
    __device__ bool skip() const { return false; }
    __device__ __forceinline__ Pre load(int i) const {
        Pre p;
        p.l = lc[i];
        p.u = uc[i];
        p.id = invD1[i];
#pragma unroll
        for (int k = 0; k < K; ++k) p.y[k] = (double)Y[k][i];
        p.yr = dist ? (double)y_rst[i] : 0.0;
        return p;
    }
    template <class Acc>
    __device__ __forceinline__ void row(int i, const Acc* acc, double* q, const Pre& p) const {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const double ax = (double)acc[k];
            const double pd = p.id * (ax - proj_box(ax, p.l, p.u));
            q[2 * k] += pd * pd;
            const double yk = p.y[k];
            q[2 * k + 1] += safeprod_left(p.u, pos_part(-yk)) - safeprod_left(p.l, neg_part(-yk));
        }
        if (dist) {
            const double d0 = p.y[0] - p.yr, d1 = p.y[K > 1 ? 1 : 0] - p.yr;
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
    struct Pre {
        double c, l, u, id, xr;
        double x[K];
    };
    __device__ bool skip() const { return false; }
    __device__ __forceinline__ Pre load(int j) const {
        Pre p;
        p.c = c[j];
        p.l = lv[j];
        p.u = uv[j];
        p.id = invD2[j];
#pragma unroll
        for (int k = 0; k < K; ++k) p.x[k] = (double)X[k][j];
        p.xr = dist ? (double)x_rst[j] : 0.0;
        return p;
    }
    template <class Acc>
    __device__ __forceinline__ void row(int j, const Acc* acc, double* q, const Pre& p) const {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const double g = p.c - (double)acc[k];
            const double r = proj_multiplier(g, p.l, p.u);
            const double dd = p.id * (g - r);
            q[3 * k] += dd * dd;
            q[3 * k + 1] += safeprod_left(p.u, pos_part(-r)) - safeprod_left(p.l, neg_part(-r));
            q[3 * k + 2] += p.c * p.x[k];
        }
        if (dist) {
            const double d0 = p.x[0] - p.xr, d1 = p.x[K > 1 ? 1 : 0] - p.xr;
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

struct NoPre {};

template <class V>
struct EpiStore {   // t = A x (skipped once the power method has converged)
    static constexpr int NQ = 0;
    double* partials = nullptr;
    V* out;
    const PowState* st;
    using Pre = NoPre;
    __device__ bool skip() const { return st && st->done; }
    __device__ __forceinline__ Pre load(int) const { return {}; }
    template <class Acc>
    __device__ __forceinline__ void row(int i, const Acc* acc, double*, const Pre&) const {
        out[i] = (V)acc[0];
    }
};

template <class V>
struct EpiPowAt {   // Ax = A' t; sums of x.Ax and Ax.Ax
    static constexpr int NQ = 2;
    double* partials;
    V* Ax;
    const V* x;
    const PowState* st;
    struct Pre {
        V x;
    };
    __device__ bool skip() const { return st->done; }
    __device__ __forceinline__ Pre load(int j) const { return Pre{x[j]}; }
    template <class Acc>
    __device__ __forceinline__ void row(int j, const Acc* acc, double* q, const Pre& p) const {
        const V a = (V)acc[0];
        Ax[j] = a;
        q[0] += (double)(p.x * a);
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
    using Pre = NoPre;
    __device__ bool skip() const { return false; }
    __device__ __forceinline__ Pre load(int) const { return {}; }
    template <class Acc>
    __device__ __forceinline__ void row(int r, const Acc* acc, double*, const Pre&) const {
        const int i = out_index ? out_index[r] : r;
        const double s = (double)acc[0];
        out[i] = beta == 0.0 ? alpha * s : alpha * s + beta * out[i];
    }
};

}  // namespace cpdlp
