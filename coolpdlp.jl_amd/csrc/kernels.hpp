// kernels.hpp -- device kernels and epilogues of the PDLP/PDHG loop.
//
// Step epilogues follow the reference's broadcasts operation by operation
// (src/algorithms/pdlp.jl:110-117, pdhg.jl:91-97) and the averaging axpby!
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

// ---------------------------------------------------------------- infeasibility rays
// Opt-in detection (no counterpart in the reference).  Three candidate rays per side:
// k = 0, the current iterate minus the restart point; k = 1, the current iterate;
// k = 2, the current iterate minus the previous one.  The products run over
// {current, restart, previous} and give every candidate by linearity.  Sums are in
// original units -- x = D2 x_s, y = D1 y_s, A x = inv(D1) A_s x_s, A'y = inv(D2) A_s'y_s --
// and a bound times a ray component is the same number in either space.
__device__ __forceinline__ double ray_of(int k, double cur, double rst, double last) {
    return k == 0 ? cur - rst : k == 1 ? cur : cur - last;
}

// rows of A: for the primal ray d, |violation of A d in the recession cone of
// [lc, uc]|^2 (A d = 0 on rows with two finite bounds, >= 0 with lc only, <= 0 with uc
// only); for the dual ray e, the Farkas objective lc e+ - uc e- over finite bounds and
// |sign violation|^2 (e+ where lc = -inf, e- where uc = +inf).
template <class V>
struct EpiRayRows {
    static constexpr int NQ = 9;   // cone violation [3], objective [3], sign violation [3]
    double* partials;
    const V *y, *y_rst, *y_last;
    const double *lc, *uc, *invD1;
    struct Pre {
        double l, u, id, y, yr, yl;
    };
    __device__ bool skip() const { return false; }
    __device__ __forceinline__ Pre load(int i) const {
        return Pre{lc[i], uc[i], invD1[i], (double)y[i], (double)y_rst[i], (double)y_last[i]};
    }
    template <class Acc>
    __device__ __forceinline__ void row(int, const Acc* acc, double* q, const Pre& p) const {
        const bool lf = isfinite(p.l), uf = isfinite(p.u);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double v = p.id * ray_of(k, (double)acc[0], (double)acc[1], (double)acc[2]);
            const double cv = lf && uf ? v : lf ? fmin(v, 0.0) : uf ? fmax(v, 0.0) : 0.0;
            q[k] += cv * cv;
            const double e = ray_of(k, p.y, p.yr, p.yl);
            const double ep = pos_part(e), en = neg_part(e);
            q[3 + k] += (lf ? p.l * ep : 0.0) - (uf ? p.u * en : 0.0);
            const double sv = ((lf ? 0.0 : ep) + (uf ? 0.0 : en)) / p.id;
            q[6 + k] += sv * sv;
        }
    }
};

// rows of A' (= columns j): for the dual ray e with r = -A'e, the Farkas objective
// lv r+ - uv r- over finite bounds and |sign violation|^2 (r+ where lv = -inf, r- where
// uv = +inf); for the primal ray d, c'd and |violation of d in the recession cone of
// [lv, uv]|^2.
template <class V>
struct EpiRayCols {
    static constexpr int NQ = 12;   // objective [3], sign violation [3], c'd [3], cone violation [3]
    double* partials;
    const V *x, *x_rst, *x_last;
    const double *c, *lv, *uv, *invD2;
    struct Pre {
        double c, l, u, id, x, xr, xl;
    };
    __device__ bool skip() const { return false; }
    __device__ __forceinline__ Pre load(int j) const {
        return Pre{c[j], lv[j], uv[j], invD2[j], (double)x[j], (double)x_rst[j], (double)x_last[j]};
    }
    template <class Acc>
    __device__ __forceinline__ void row(int, const Acc* acc, double* q, const Pre& p) const {
        const bool lf = isfinite(p.l), uf = isfinite(p.u);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double r = -ray_of(k, (double)acc[0], (double)acc[1], (double)acc[2]);
            const double rp = pos_part(r), rn = neg_part(r);
            q[k] += (lf ? p.l * rp : 0.0) - (uf ? p.u * rn : 0.0);
            const double sv = p.id * ((lf ? 0.0 : rp) + (uf ? 0.0 : rn));
            q[3 + k] += sv * sv;
            const double d = ray_of(k, p.x, p.xr, p.xl);
            q[6 + k] += p.c * d;
            const double cv = (lf && uf ? d : lf ? fmin(d, 0.0) : uf ? fmax(d, 0.0) : 0.0) / p.id;
            q[9 + k] += cv * cv;
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
