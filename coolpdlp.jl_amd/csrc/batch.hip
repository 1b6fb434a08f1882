// batch.hip -- many independent LPs solved together (batched and served solves).
//
// One 256-thread workgroup per instance runs the reference's PDLP loop on it
// (src/algorithms/pdlp.jl:69-196), in three kernels: k_batch_setup (Ruiz +
// Chambolle-Pock scaling, src/components/preconditioning.jl:117-155, and the power
// method, src/utils/linalg.jl:86-96), k_batch_steps (check_every PDHG steps) and
// k_batch_check (the 5-candidate KKT check in FP64, src/components/errors.jl:65-109;
// termination, src/components/termination.jl:89-101; the restart rule,
// src/components/restart.jl:40-57; the primal-weight update,
// src/components/step_size.jl:58-75) -- all decisions on the device.
//
// The step kernel walks row tiles (<= 256 rows, <= 2048 nonzeros): the tile's entries
// are loaded coalesced, all gathers issued together, the products summed per row in
// LDS in the reference's order.  y (primal half-step) and xbar (dual half-step) are
// gathered from LDS when the instance's vectors fit the LDS budget, else from L2.
//
// Precision MIXED: FP32 matrix, iterates and products (FP32 accumulation); KKT sums,
// restart decisions, omega and eta in FP64; when the FP32 phase reaches polish_reltol
// or stops improving, the instance continues in FP64 from the same state (every
// iterate converted; a diverging FP32 phase restarts from its best iterate instead).
#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <type_traits>
#include <deque>
#include <map>
#include <memory>

#include "kernels.hpp"
#include "layout.hpp"

namespace cpdlp {

// ---------------------------------------------------------------- device structures
struct BInst {
    int32_t m, n, nnz, La, Lat, n_a_tiles, n_at_tiles;
    int32_t warm;                         // 1: xr64 / yr64 hold the caller's unscaled start
    int32_t *a_rp, *a_col, *at_rp, *at_col, *perm;
    int32_t* ecol;                        // column of each CSC entry (flat scaling passes)
    const TileMeta *a_tiles, *at_tiles;   // row tiles of A and At for the step kernel
    double *nz, *acsc, *atcsc, *a64, *at64;
    float *a32, *at32;
    double *c, *lv, *uv, *lc, *uc;
    double *cs, *lvs, *uvs, *lcs, *ucs, *D1, *D2, *iD1, *iD2;
    float *c32, *lv32, *uv32, *lc32, *uc32;
    float *x32[2], *y32[2], *xa32, *ya32, *xal32, *yal32, *xr32, *yr32, *xb32, *yb32;
    double *x64[2], *y64[2], *xa64, *ya64, *xal64, *yal64, *xr64, *yr64, *sn, *sm;
};

struct BState {
    double eta, omega, eta_sum, best_rel, best_omega, primal_scale, dual_scale;
    double err[6];
    int64_t kkt_passes, inner, total, restarts, powm_iters, fp32_passes;
    int32_t status, phase, cur, stall, setup, timeout;
};

struct BParams {
    int32_t check_every, ruiz_iter, pdlp, precision;
    double cp_alpha, invnorm, damping, zero_tol, suff, nec, art, reltol, polish_reltol;
    double powm_tol;   // NaN: eps(Float64) * n^3 rule
    int64_t max_kkt_passes, powm_maxiter;
    uint64_t seed;
    double inf_tol;    // opt-in infeasibility detection (k_batch_check<V, true>)
};

constexpr int BB = 256;   // threads per instance workgroup
constexpr int STALL = 10;

// block-wide sum of NQ doubles, result broadcast through sh[NQ] (deterministic order).
// For a block of NT > BB threads whose threads past BB hold zeros, the result is the
// 256-thread block's bit for bit: their waves add +0.0 to a sum that started at +0.0.
template <int NQ, int NT = BB>
__device__ __forceinline__ void bsum(double (&q)[NQ], double* sh) {
    block_sum<NQ, NT>(q, sh);
    __syncthreads();
}

// rows [0, rows) of a CSR with L lanes per row (L a power of two <= 64), used by the
// K-candidate KKT check (once per check_every steps): lane l of a
// group takes entries l, l+L, ... then a butterfly over the group; f(r, acc) runs on
// lane 0 of the group.  Every thread runs the same number of outer iterations, so all
// lanes of a group take part in the shuffles.
template <class Acc, int K, class T, class GX, class F>
__device__ __forceinline__ void vcsr(int rows, const int32_t* rp, const int32_t* col, const T* val,
                                     int L, GX gx, F f) {
    const int groups = BB / L, g = threadIdx.x / L, lane = threadIdx.x & (L - 1);
    for (int base = 0; base < rows; base += groups) {
        const int r = base + g;
        Acc acc[K];
#pragma unroll
        for (int kv = 0; kv < K; ++kv) acc[kv] = Acc(0);
        if (r < rows) {
            const int k1 = rp[r + 1];
            for (int k = rp[r] + lane; k < k1; k += L) {
                const Acc v = (Acc)val[k];
                const int c = col[k];
#pragma unroll
                for (int kv = 0; kv < K; ++kv) acc[kv] += v * gx(kv, c);
            }
        }
#pragma unroll
        for (int kv = 0; kv < K; ++kv)
            for (int o = L >> 1; o > 0; o >>= 1) acc[kv] += __shfl_xor(acc[kv], o, L);
        if (r < rows && lane == 0) f(r, acc);
    }
}

__device__ __forceinline__ uint64_t b_splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ double pnorm_acc(double acc, double v, double p) {
    const double a = fabs(v);
    if (isinf(p)) return a > acc ? a : acc;
    if (p == 1.0) return acc + a;
    if (p == 2.0) return acc + a * a;
    return acc + pow(a, p);
}
__device__ __forceinline__ double pnorm_fin(double acc, double p) {
    if (isinf(p) || p == 1.0) return acc;
    if (p == 2.0) return sqrt(acc);
    return pow(acc, 1.0 / p);
}

// ---------------------------------------------------------------- tiled products
// Tiles of a CSR's rows for the step kernel: <= BB rows and <= BT nonzeros each (a
// single row longer than BT is a tile of its own).
constexpr int BT = 2048;
constexpr int BPT = BT / BB;

// One half-step over the rows of a tiled CSR.  Per tile of <= BT nonzeros, thread tid
// owns row r0 + tid: its epilogue operands (pre) and row bounds, and the tile's
// (col, val) pairs (coalesced, clamped indices) are loaded while the previous tile is
// being summed; then all gathers are issued together, the products go to LDS, and each
// thread sums its row sequentially from zero (the reference's thread-per-row order,
// src/utils/mat_csr.jl:63-79); fin(r, sum, pre) finishes the row.  A tile of one row
// longer than BT is summed by the whole block (strided, then the four wave sums in wave
// order).  Ends synchronised.
template <class V, class GX, class PRE, class FIN>
__device__ __forceinline__ void tiled_pass(const TileMeta* tiles, int ntiles, const int32_t* rp,
                                           const int32_t* col, const V* val, V* prod, V* red,
                                           GX gx, PRE pre, FIN fin) {
    using PreT = decltype(pre(0));
    const int tid = threadIdx.x;
    if (ntiles <= 0) return;
    int cc[BPT];
    V vv[BPT];
    PreT pv{};
    int lo = 0, hi = 0;
    auto fetch = [&](const TileMeta& tm) {   // the loads of a tile of <= BT nonzeros
        const int r = tm.r0 + tid;
        if (r < tm.r1) {
            pv = pre(r);
            lo = rp[r] - tm.base;
            hi = rp[r + 1] - tm.base;
        }
        if (tm.cnt > 0) {
#pragma unroll
            for (int j = 0; j < BPT; ++j) {
                const int e = min(tid + j * BB, tm.cnt - 1);
                cc[j] = col[tm.base + e];
                vv[j] = val[tm.base + e];
            }
        }
    };
    TileMeta tm = tiles[0];
    if (tm.cnt <= BT) fetch(tm);
    for (int t = 0; t < ntiles; ++t) {
        const bool more = t + 1 < ntiles;
        const TileMeta tn = more ? tiles[t + 1] : TileMeta{0, 0, 0, 0};
        if (tm.cnt > BT) {   // uniform per block
            const int r = tm.r0;
            const auto pb = pre(r);
            V acc = V(0);
            const int k1 = tm.base + tm.cnt;
            constexpr int U = 4;
            for (int kb = tm.base + tid; kb < k1; kb += U * BB) {
                int c[U];
                V v[U], xv[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int k = min(kb + u * BB, k1 - 1);
                    c[u] = col[k];
                    v[u] = val[k];
                }
#pragma unroll
                for (int u = 0; u < U; ++u) xv[u] = gx(c[u]);
#pragma unroll
                for (int u = 0; u < U; ++u)
                    if (kb + u * BB < k1) acc += v[u] * xv[u];
            }
            acc = wave_sum(acc);
            if ((tid & (WAVE - 1)) == 0) red[tid / WAVE] = acc;
            __syncthreads();
            if (tid == 0) {
                V sum = red[0];
#pragma unroll
                for (int w = 1; w < BB / WAVE; ++w) sum += red[w];
                fin(r, sum, pb);
            }
            if (more && tn.cnt <= BT) fetch(tn);
            __syncthreads();
            tm = tn;
            continue;
        }
        const int r = tm.r0 + tid;
        const bool own = r < tm.r1;
        if (tm.cnt > 0) {   // uniform per block
            V xv[BPT];
#pragma unroll
            for (int j = 0; j < BPT; ++j) xv[j] = gx(cc[j]);
#pragma unroll
            for (int j = 0; j < BPT; ++j) prod[tid + j * BB] = vv[j] * xv[j];
        }
        __syncthreads();
        const PreT pc = pv;
        const int l0 = lo, h0 = hi;
        if (more && tn.cnt <= BT) fetch(tn);   // in flight while this tile is summed
        if (own) fin(r, seq_sum(prod, l0, h0), pc);
        __syncthreads();
        tm = tn;
    }
}

// tiled_pass for a block of NSUB * BB threads: quarter (sub-block) s sums the tiles
// s, s + NSUB, ... side by side with the others, each exactly as tiled_pass sums a tile
// (own LDS products prod[s], own wave sums red[s]), so the row sums are the same bits.
// Every trip has the two barriers of either tile kind, so sub-blocks with a long row, a
// short tile or no tile at all stay in step.  Ends synchronised.
template <int NSUB, class V, class GX, class PRE, class FIN>
__device__ __forceinline__ void tiled_pass_wide(const TileMeta* tiles, int ntiles, const int32_t* rp,
                                                const int32_t* col, const V* val, V* prod_all,
                                                V* red_all, GX gx, PRE pre, FIN fin) {
    using PreT = decltype(pre(0));
    const int sub = threadIdx.x / BB, tid = threadIdx.x % BB;
    V* const prod = prod_all + sub * BT;
    V* const red = red_all + sub * (BB / WAVE);
    if (ntiles <= 0) return;
    int cc[BPT];
    V vv[BPT];
    PreT pv{};
    int lo = 0, hi = 0;
    auto fetch = [&](const TileMeta& tm) {
        const int r = tm.r0 + tid;
        if (r < tm.r1) {
            pv = pre(r);
            lo = rp[r] - tm.base;
            hi = rp[r + 1] - tm.base;
        }
        if (tm.cnt > 0) {
#pragma unroll
            for (int j = 0; j < BPT; ++j) {
                const int e = min(tid + j * BB, tm.cnt - 1);
                cc[j] = col[tm.base + e];
                vv[j] = val[tm.base + e];
            }
        }
    };
    TileMeta tm = sub < ntiles ? tiles[sub] : TileMeta{0, 0, 0, 0};
    if (tm.cnt <= BT) fetch(tm);
    const int trips = (ntiles + NSUB - 1) / NSUB;
    for (int k = 0; k < trips; ++k) {
        const int tnext = (k + 1) * NSUB + sub;
        const bool more = tnext < ntiles;
        const TileMeta tn = more ? tiles[tnext] : TileMeta{0, 0, 0, 0};
        if (tm.cnt > BT) {   // uniform per sub-block
            const int r = tm.r0;
            const auto pb = pre(r);
            V acc = V(0);
            const int k1 = tm.base + tm.cnt;
            constexpr int U = 4;
            for (int kb = tm.base + tid; kb < k1; kb += U * BB) {
                int c[U];
                V v[U], xv[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int kk = min(kb + u * BB, k1 - 1);
                    c[u] = col[kk];
                    v[u] = val[kk];
                }
#pragma unroll
                for (int u = 0; u < U; ++u) xv[u] = gx(c[u]);
#pragma unroll
                for (int u = 0; u < U; ++u)
                    if (kb + u * BB < k1) acc += v[u] * xv[u];
            }
            acc = wave_sum(acc);
            if ((tid & (WAVE - 1)) == 0) red[tid / WAVE] = acc;
            __syncthreads();
            if (tid == 0) {
                V sum = red[0];
#pragma unroll
                for (int w = 1; w < BB / WAVE; ++w) sum += red[w];
                fin(r, sum, pb);
            }
            if (more && tn.cnt <= BT) fetch(tn);
            __syncthreads();
            tm = tn;
            continue;
        }
        const int r = tm.r0 + tid;
        const bool own = r < tm.r1;
        if (tm.cnt > 0) {   // uniform per sub-block
            V xv[BPT];
#pragma unroll
            for (int j = 0; j < BPT; ++j) xv[j] = gx(cc[j]);
#pragma unroll
            for (int j = 0; j < BPT; ++j) prod[tid + j * BB] = vv[j] * xv[j];
        }
        __syncthreads();
        const PreT pc = pv;
        const int l0 = lo, h0 = hi;
        if (more && tn.cnt <= BT) fetch(tn);
        if (own) fin(r, seq_sum(prod, l0, h0), pc);
        __syncthreads();
        tm = tn;
    }
}

// K products over the rows of a tiled CSR, accumulated in FP64 (the KKT check): per tile
// the (col, val) pairs are loaded coalesced once, then for each vector kv the gathers
// gx(kv, col) are issued together, the products go to LDS and each thread sums its row
// sequentially from zero (the reference's order); fin(r, acc) gets the K row sums.  A
// tile of one row longer than BT is summed by the whole block (strided, then the wave
// sums in wave order).  Ends synchronised.
template <int K, class V, class GX, class FIN>
__device__ __forceinline__ void tiled_pass_k(const TileMeta* tiles, int ntiles, const int32_t* rp,
                                             const int32_t* col, const V* val, double* prod,
                                             double (*red)[BB / WAVE], GX gx, FIN fin) {
    const int tid = threadIdx.x;
    for (int t = 0; t < ntiles; ++t) {
        const TileMeta tm = tiles[t];
        if (tm.cnt > BT) {   // uniform per block
            double acc[K];
#pragma unroll
            for (int kv = 0; kv < K; ++kv) acc[kv] = 0.0;
            const int k1 = tm.base + tm.cnt;
            for (int k = tm.base + tid; k < k1; k += BB) {
                const double v = (double)val[k];
                const int c = col[k];
#pragma unroll
                for (int kv = 0; kv < K; ++kv) acc[kv] += v * gx(kv, c);
            }
#pragma unroll
            for (int kv = 0; kv < K; ++kv) {
                const double w = wave_sum(acc[kv]);
                if ((tid & (WAVE - 1)) == 0) red[kv][tid / WAVE] = w;
            }
            __syncthreads();
            if (tid == 0) {
                double sum[K];
#pragma unroll
                for (int kv = 0; kv < K; ++kv) {
                    sum[kv] = red[kv][0];
                    for (int w = 1; w < BB / WAVE; ++w) sum[kv] += red[kv][w];
                }
                fin(tm.r0, sum);
            }
            __syncthreads();
            continue;
        }
        const int r = tm.r0 + tid;
        const bool own = r < tm.r1;
        int lo = 0, hi = 0;
        if (own) {
            lo = rp[r] - tm.base;
            hi = rp[r + 1] - tm.base;
        }
        int cc[BPT];
        double vv[BPT];
        if (tm.cnt > 0) {
#pragma unroll
            for (int j = 0; j < BPT; ++j) {
                const int e = min(tid + j * BB, tm.cnt - 1);
                cc[j] = col[tm.base + e];
                vv[j] = (double)val[tm.base + e];
            }
        }
        double acc[K];
#pragma unroll
        for (int kv = 0; kv < K; ++kv) {
            if (tm.cnt > 0) {
                double xv[BPT];
#pragma unroll
                for (int j = 0; j < BPT; ++j) xv[j] = gx(kv, cc[j]);
#pragma unroll
                for (int j = 0; j < BPT; ++j) prod[tid + j * BB] = vv[j] * xv[j];
            }
            __syncthreads();
            acc[kv] = own ? seq_sum(prod, lo, hi) : 0.0;
            __syncthreads();
        }
        if (own) fin(r, acc);
    }
}

// ---------------------------------------------------------------- setup kernel
// preprocess + initialize for each listed instance (one workgroup each).  NSUB > 1: a
// block of NSUB * BB threads (large instances).  Element-wise and per-row work is spread
// over all threads; the power method's products run NSUB tiles side by side (the same
// row sums); every sum over a vector is accumulated by the first BB threads in the
// 256-thread block's order -- so the result is the same bits for any NSUB.
template <int NSUB = 1>
__global__ __launch_bounds__(NSUB * BB) void k_batch_setup(const int32_t* list, const BInst* insts,
                                                           BState* states, BParams P) {
    constexpr int NT = NSUB * BB;
    __shared__ double sh[8];
    __shared__ double prod[NSUB * BT];
    __shared__ double red[NSUB * (BB / WAVE)];
    const int slot = list[blockIdx.x];
    const BInst I = insts[slot];
    BState* S = states + slot;
    const int m = I.m, n = I.n, nnz = I.nnz, tid = threadIdx.x;
    double* d1 = I.sm;
    double* d2 = I.sn;
    // identity + working copies (CSC order: A values and A' values)
    for (int i = tid; i < m; i += NT) I.D1[i] = 1.0;
    for (int j = tid; j < n; j += NT) I.D2[j] = 1.0;
    for (int k = tid; k < nnz; k += NT) {
        I.acsc[k] = I.nz[k];
        I.atcsc[k] = I.nz[k];
    }
    __syncthreads();
    auto norms = [&](double p_row, double p_col) {
        for (int j = tid; j < n; j += NT) {   // columns of A = rows of CSR(A')
            double acc = 0.0;
            for (int k = I.at_rp[j]; k < I.at_rp[j + 1]; ++k) acc = pnorm_acc(acc, I.acsc[k], p_col);
            const double v = pnorm_fin(acc, p_col);
            d2[j] = v == 0.0 ? 1.0 : 1.0 / sqrt(v);
        }
        for (int i = tid; i < m; i += NT) {   // rows of A, from the A' values
            double acc = 0.0;
            for (int k = I.a_rp[i]; k < I.a_rp[i + 1]; ++k) acc = pnorm_acc(acc, I.atcsc[I.perm[k]], p_row);
            const double v = pnorm_fin(acc, p_row);
            d1[i] = v == 0.0 ? 1.0 : 1.0 / sqrt(v);
        }
        __syncthreads();
    };
    // (d1_i * a) * d2_j into A, (d2_j * a') * d1_i into A'  (preconditioning.jl:37-43)
    auto apply = [&](const double* ain, const double* atin, const double* D1, const double* D2) {
        for (int j = tid; j < n; j += NT) {
            const double dj = D2[j];
            for (int k = I.at_rp[j]; k < I.at_rp[j + 1]; ++k) {
                const double di = D1[I.at_col[k]];
                const double a = ain[k], at = atin[k];
                I.acsc[k] = (di * a) * dj;
                I.atcsc[k] = (dj * at) * di;
            }
        }
        __syncthreads();
    };
    // Ruiz's infinity norms in one flat pass over the entries: exact maxima (order-free,
    // so the same values as a sequential max) via atomicMax on the bit patterns of |a|,
    // which order like the values for non-negative doubles; NaN entries are skipped as
    // the sequential max skips them
    auto inf_norms = [&]() {
        unsigned long long* b1 = (unsigned long long*)d1;
        unsigned long long* b2 = (unsigned long long*)d2;
        for (int i = tid; i < m; i += NT) b1[i] = 0ull;
        for (int j = tid; j < n; j += NT) b2[j] = 0ull;
        __syncthreads();
        for (int k = tid; k < nnz; k += NT) {
            const double a = fabs(I.acsc[k]), at = fabs(I.atcsc[k]);
            if (a == a) atomicMax(&b2[I.ecol[k]], (unsigned long long)__double_as_longlong(a));
            if (at == at) atomicMax(&b1[I.at_col[k]], (unsigned long long)__double_as_longlong(at));
        }
        __syncthreads();
        for (int i = tid; i < m; i += NT) {
            const double v = __longlong_as_double((long long)b1[i]);
            d1[i] = v == 0.0 ? 1.0 : 1.0 / sqrt(v);
        }
        for (int j = tid; j < n; j += NT) {
            const double v = __longlong_as_double((long long)b2[j]);
            d2[j] = v == 0.0 ? 1.0 : 1.0 / sqrt(v);
        }
        __syncthreads();
    };
    // (d1_i * a) * d2_j into A, (d2_j * a') * d1_i into A', entry by entry
    auto apply_flat = [&](const double* D1v, const double* D2v) {
        for (int k = tid; k < nnz; k += NT) {
            const double di = D1v[I.at_col[k]], dj = D2v[I.ecol[k]];
            I.acsc[k] = (di * I.acsc[k]) * dj;
            I.atcsc[k] = (dj * I.atcsc[k]) * di;
        }
        __syncthreads();
    };
    for (int it = 0; it < P.ruiz_iter; ++it) {   // ruiz_preconditioner, :117-125
        inf_norms();
        apply_flat(d1, d2);
        for (int i = tid; i < m; i += NT) I.D1[i] = d1[i] * I.D1[i];
        for (int j = tid; j < n; j += NT) I.D2[j] = I.D2[j] * d2[j];
        __syncthreads();
    }
    apply(I.nz, I.nz, I.D1, I.D2);   // cons_r from the originals
    norms(2.0 - P.cp_alpha, P.cp_alpha);
    for (int i = tid; i < m; i += NT) I.D1[i] = I.D1[i] * d1[i];
    for (int j = tid; j < n; j += NT) I.D2[j] = d2[j] * I.D2[j];
    __syncthreads();
    apply(I.nz, I.nz, I.D1, I.D2);   // A_p = D1*A*D2, At_p = D2*At*D1
    for (int k = tid; k < nnz; k += NT) {
        const double a = I.acsc[I.perm[k]];   // CSR(A) order
        I.a64[k] = a;
        I.a32[k] = (float)a;
        I.at64[k] = I.atcsc[k];
        I.at32[k] = (float)I.atcsc[k];
    }
    for (int j = tid; j < n; j += NT) {
        const double d = I.D2[j];
        I.cs[j] = d * I.c[j];
        I.lvs[j] = I.lv[j] / d;
        I.uvs[j] = I.uv[j] / d;
        I.iD2[j] = 1.0 / d;
        I.c32[j] = (float)I.cs[j];
        I.lv32[j] = (float)I.lvs[j];
        I.uv32[j] = (float)I.uvs[j];
    }
    for (int i = tid; i < m; i += NT) {
        const double d = I.D1[i];
        I.lcs[i] = d * I.lc[i];
        I.ucs[i] = d * I.uc[i];
        I.iD1[i] = 1.0 / d;
        I.lc32[i] = (float)I.lcs[i];
        I.uc32[i] = (float)I.ucs[i];
    }
    __syncthreads();
    // constant KKT scales and the primal_weight_init norms
    {
        double q[5] = {0, 0, 0, 0, 0};
        for (int i = tid; i < m && tid < BB; i += BB) {
            const double cb = combine(I.lcs[i], I.ucs[i]);
            const double v = I.iD1[i] * cb;
            q[0] += v * v;
            q[2] += cb * cb;
        }
        for (int j = tid; j < n && tid < BB; j += BB) {
            const double v = I.iD2[j] * I.cs[j];
            q[1] += v * v;
            q[3] += I.cs[j] * I.cs[j];
            q[4] += I.c[j] != 0.0 ? 1.0 : 0.0;
        }
        bsum<5, NT>(q, sh);
    }
    const double primal_scale = 1.0 + sqrt(sh[0]), dual_scale = 1.0 + sqrt(sh[1]);
    const double comb_norm = sqrt(sh[2]), c_norm = sqrt(sh[3]);
    // no constraints and c == 0: the early exit of common.jl:194-198, sol.x =
    // proj_box(0, lv, uv) of the preprocessed problem, OPTIMAL without a step
    const bool early = m == 0 && sh[4] == 0.0;
    __syncthreads();
    // power method on A'A (FP64): x = I.x64[0], Ax = I.x64[1], t = I.y64[0]
    double* x = I.x64[0];
    double* Ax = I.x64[1];
    double* t = I.y64[0];
    for (int j = tid; j < n; j += NT) {
        double acc = 0.0;
        for (int r = 0; r < 12; ++r) {
            const uint64_t h = b_splitmix64(P.seed * 0x100000001B3ull ^ b_splitmix64((uint64_t)j * 16u + (uint64_t)r));
            acc += (double)(h >> 11) * 0x1.0p-53;
        }
        x[j] = acc - 6.0;
    }
    __syncthreads();
    {
        double q[1] = {0.0};
        for (int j = tid; j < n && tid < BB; j += BB) q[0] += x[j] * x[j];
        bsum<1, NT>(q, sh);
    }
    {
        const double inv = 1.0 / sqrt(sh[0]);
        __syncthreads();
        for (int j = tid; j < n; j += NT) x[j] = x[j] * inv;
    }
    __syncthreads();
    const uint64_t u = (uint64_t)n;
    const double tol = isnan(P.powm_tol) ? DBL_EPSILON * (double)(int64_t)(u * u * u) : P.powm_tol;
    const int64_t maxiter = P.powm_maxiter < 0 ? n : P.powm_maxiter;
    double theta = 0.0, residual = 1.7976931348623157e308;
    int64_t it = 0;
    while (!(it > maxiter || residual <= tol * fabs(theta))) {
        auto pass = [&](auto... a) {
            if constexpr (NSUB == 1) tiled_pass<double>(a...);
            else tiled_pass_wide<NSUB, double>(a...);
        };
        pass(I.a_tiles, I.n_a_tiles, I.a_rp, I.a_col, I.a64, prod, red,
             [&](int c) { return x[c]; }, [&](int) { return 0; },
             [&](int r, double acc, int) { t[r] = acc; });
        // Ax = A't; then x.Ax and Ax.Ax
        pass(I.at_tiles, I.n_at_tiles, I.at_rp, I.at_col, I.at64, prod, red,
             [&](int c) { return t[c]; }, [&](int) { return 0; },
             [&](int r, double acc, int) { Ax[r] = acc; });
        double q[2] = {0.0, 0.0};
        for (int j = tid; j < n && tid < BB; j += BB) {
            const double a = Ax[j];
            q[0] += x[j] * a;
            q[1] += a * a;
        }
        bsum<2, NT>(q, sh);
        theta = sh[0];
        const double inv = 1.0 / sqrt(sh[1]);
        __syncthreads();
        double q2[1] = {0.0};
        for (int j = tid; j < n && tid < BB; j += BB) {
            const double r = Ax[j] + (-theta) * x[j];
            q2[0] += r * r;
            x[j] = Ax[j] * inv;
        }
        bsum<1, NT>(q2, sh);
        residual = sqrt(sh[0]);
        __syncthreads();
        ++it;
    }
    const double eta = P.invnorm * (1.0 / sqrt(theta));
    double omega = 1.0;
    if (P.pdlp && c_norm > P.zero_tol && comb_norm > P.zero_tol) omega = c_norm / comb_norm;
    // initial iterates: sol = precondition(sol_init) = (x0 / D2, y0 / D1), or zero
    // without a start (common.jl:200-207); sol_last = 0, avg = sol, avg_last = 0,
    // restart = sol (pdlp.jl:51-54).  The start was uploaded into the restart buffers,
    // which nothing above touches; each thread reads its own entries before writing them.
    for (int j = tid; j < n; j += NT) {
        const double xs = early ? proj_box(0.0, I.lvs[j], I.uvs[j]) : I.warm ? I.xr64[j] / I.D2[j] : 0.0;
        const float xf = (float)xs;
        I.x32[0][j] = I.xa32[j] = I.xr32[j] = xf;
        I.x32[1][j] = I.xal32[j] = I.xb32[j] = 0.f;
        I.x64[0][j] = I.xa64[j] = I.xr64[j] = xs;
        I.x64[1][j] = I.xal64[j] = 0.0;
    }
    for (int i = tid; i < m; i += NT) {
        const double ys = I.warm ? I.yr64[i] / I.D1[i] : 0.0;
        const float yf = (float)ys;
        I.y32[0][i] = I.ya32[i] = I.yr32[i] = yf;
        I.y32[1][i] = I.yal32[i] = I.yb32[i] = 0.f;
        I.y64[0][i] = I.ya64[i] = I.yr64[i] = ys;
        I.y64[1][i] = I.yal64[i] = 0.0;
    }
    // publish: every thread's writes are in this XCD's L2 after the barrier; thread 0 writes
    // the state, writes the L2 back (agent-scope fence) and only then sets the flag, so the
    // host may take the instance up while the rest of the launch still runs
    __syncthreads();
    if (tid == 0) {
        S->eta = eta;
        S->omega = omega;
        S->eta_sum = 0.0;
        S->best_rel = INFINITY;
        S->best_omega = omega;
        S->primal_scale = primal_scale;
        S->dual_scale = dual_scale;
        for (int f = 0; f < 6; ++f) S->err[f] = NAN;
        S->kkt_passes = S->inner = S->total = S->restarts = S->fp32_passes = 0;
        S->powm_iters = it;
        S->status = early ? CPDLP_OPTIMAL : CPDLP_STILL_RUNNING;
        S->phase = early || P.precision == CPDLP_PREC_FP64 ? 1 : 0;
        S->cur = 0;
        S->stall = 0;
        __threadfence();
        __hip_atomic_store(&S->setup, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// ---------------------------------------------------------------- iteration kernel
template <class V>
struct Iter {   // the precision-V view of an instance's iterates and problem data
    V *x[2], *y[2], *xa, *ya, *xal, *yal, *xr, *yr;
    const V *a, *at, *c, *lv, *uv, *lc, *uc;
};
__device__ __forceinline__ Iter<float> iter_of(const BInst& I, float) {
    return Iter<float>{{I.x32[0], I.x32[1]}, {I.y32[0], I.y32[1]}, I.xa32, I.ya32, I.xal32, I.yal32,
                       I.xr32, I.yr32, I.a32, I.at32, I.c32, I.lv32, I.uv32, I.lc32, I.uc32};
}
__device__ __forceinline__ Iter<double> iter_of(const BInst& I, double) {
    return Iter<double>{{I.x64[0], I.x64[1]}, {I.y64[0], I.y64[1]}, I.xa64, I.ya64, I.xal64,
                        I.yal64, I.xr64, I.yr64, I.a64, I.at64, I.cs, I.lvs, I.uvs, I.lcs, I.ucs};
}

__device__ __forceinline__ double rel_of(const double* e) {
    return fmax(fmax(e[0] / e[1], e[2] / e[3]), e[4] / e[5]);
}
__device__ __forceinline__ double abs_of(const double* e, double w) {
    return sqrt(w * w * e[0] * e[0] + (1.0 / (w * w)) * e[2] * e[2] + e[4] * e[4]);
}

template <class V>
struct PrePrimal {
    V x, c, lv, uv, a;
};
template <class V>
struct PreDual {
    V y, lc, uc, a;
};

// check_every PDHG steps (pdlp.jl:96-124 with update_average!, :126-136; pdhg.jl:76-102)
// on each listed instance.  LDSG: y and xbar live in LDS (dynamic shared memory of
// (n + m) * sizeof(V) bytes) and the gathers read them there; else from global memory.
// NSUB > 1: a block of NSUB * BB threads per instance, NSUB tiles at a time (large
// instances, whose single workgroup would otherwise set the batch's tail).
template <class V, bool LDSG, int NSUB = 1>
__global__ __launch_bounds__(NSUB * BB) void k_batch_steps(const int32_t* list, const BInst* insts,
                                                           BState* states, BParams P) {
    constexpr int NT = NSUB * BB;
    extern __shared__ __align__(16) unsigned char dyn[];
    __shared__ V prod[NSUB * BT];
    __shared__ V red[NSUB * (BB / WAVE)];
    const int slot = list[blockIdx.x];
    const BInst I = insts[slot];
    BState* S = states + slot;
    const int m = I.m, n = I.n, tid = threadIdx.x;
    const int want_phase = std::is_same<V, double>::value ? 1 : 0;
    if (S->setup != 1 || S->status != CPDLP_STILL_RUNNING || S->phase != want_phase) return;
    const Iter<V> it = iter_of(I, V(0));
    V* xbs = LDSG ? (V*)dyn : nullptr;          // xbar (n)
    V* ys = LDSG ? (V*)dyn + n : nullptr;       // y (m)
    V* xbar_g = (V*)I.sn;   // global xbar when the vectors do not fit in LDS
    const double eta = S->eta, omega = S->omega;
    double eta_sum = S->eta_sum;
    int cur = S->cur;
    if (LDSG) {   // y of the current solution into LDS
        const V* yc = cur ? it.y[1] : it.y[0];
        for (int i = tid; i < m; i += NT) ys[i] = yc[i];
    }
    __syncthreads();
    const double tau = eta / omega, sigma = eta * omega, isig = 1.0 / sigma;
    const V tv = (V)tau, sv = (V)sigma, iv = (V)isig;
    const bool pdlp = P.pdlp != 0;
    // plain pointers (no runtime-indexed arrays, which would live in scratch memory)
    V *const x0 = it.x[0], *const x1 = it.x[1], *const y0 = it.y[0], *const y1 = it.y[1];
    V *const xa = it.xa, *const xal = it.xal, *const ya = it.ya, *const yal = it.yal;
    const V *const cv = it.c, *const lvv = it.lv, *const uvv = it.uv, *const lcv = it.lc,
            *const ucv = it.uc, *const aval = it.a, *const atval = it.at;
    const TileMeta *const a_tiles = I.a_tiles, *const at_tiles = I.at_tiles;
    const int na_t = I.n_a_tiles, nat_t = I.n_at_tiles;
    const int32_t *const a_rp = I.a_rp, *const a_col = I.a_col, *const at_rp = I.at_rp,
                  *const at_col = I.at_col;
    for (int s = 0; s < P.check_every; ++s) {
        V* const xp_ = cur ? x1 : x0;
        V* const xq_ = cur ? x0 : x1;
        const V* const yp = cur ? y1 : y0;
        V* const yq_ = cur ? y0 : y1;
        const int q = 1 - cur;
        const double wa = eta / (eta + eta_sum), wb = eta_sum / (eta + eta_sum);
        const V wav = (V)wa, wbv = (V)wb;
        const bool last = s == P.check_every - 1;
        auto pass = [&](auto... a) {
            if constexpr (NSUB == 1) tiled_pass<V>(a...);
            else tiled_pass_wide<NSUB, V>(a...);
        };
        // primal half-step over the rows of A'
        pass(
            at_tiles, nat_t, at_rp, at_col, atval, prod, red,
            [&](int c) { return LDSG ? ys[c] : yp[c]; },
            [&](int j) {
                PrePrimal<V> r;
                r.x = xp_[j];
                r.c = cv[j];
                r.lv = lvv[j];
                r.uv = uvv[j];
                r.a = pdlp ? xa[j] : V(0);
                return r;
            },
            [&](int j, V acc, const PrePrimal<V>& r) {
                const V xp = proj_box<V>(r.x - tv * (r.c - acc), r.lv, r.uv);
                xq_[j] = xp;
                const V xb = V(2) * xp - r.x;
                if (LDSG) xbs[j] = xb;
                else xbar_g[j] = xb;
                if (pdlp) {
                    if (last) xal[j] = r.a;
                    xa[j] = wav * xp + wbv * r.a;
                }
            });
        // dual half-step over the rows of A
        pass(
            a_tiles, na_t, a_rp, a_col, aval, prod, red,
            [&](int c) { return LDSG ? xbs[c] : xbar_g[c]; },
            [&](int i) {
                PreDual<V> r;
                r.y = LDSG ? ys[i] : yp[i];
                r.lc = lcv[i];
                r.uc = ucv[i];
                r.a = pdlp ? ya[i] : V(0);
                return r;
            },
            [&](int i, V ax, const PreDual<V>& r) {
                const V yn = (r.y - sv * ax) - sv * proj_box<V>(iv * r.y - ax, -r.uc, -r.lc);
                yq_[i] = yn;
                if (LDSG) ys[i] = yn;
                if (pdlp) {
                    if (last) yal[i] = r.a;
                    ya[i] = wav * yn + wbv * r.a;
                }
            });
        cur = q;
        if (pdlp) eta_sum += eta;
    }
    if (tid == 0) {
        S->eta_sum = eta_sum;
        S->cur = cur;
        S->kkt_passes += P.check_every;
        if (pdlp) {
            S->inner += P.check_every;
            S->total += P.check_every;
        }
    }
}

// ---------------------------------------------------------------- infeasibility rays
// The plain engine's EpiRayRows / EpiRayCols on one instance: products over
// {current, restart, previous} give the three candidate rays by linearity; sums in
// original units.  Returns PRIMAL_INFEASIBLE, DUAL_INFEASIBLE or STILL_RUNNING (block-
// uniform).  Ends synchronised.
template <class V>
__device__ int32_t batch_rays(const BInst& I, const Iter<V>& it, int cur, double tol, double* kprod,
                              double (*kred)[BB / WAVE], double* sh) {
    const V* const xc = cur ? it.x[1] : it.x[0];
    const V* const yc = cur ? it.y[1] : it.y[0];
    const V* const xl = cur ? it.x[0] : it.x[1];
    const V* const yl = cur ? it.y[0] : it.y[1];
    auto ray = [](int k, double a, double r, double l) { return k == 0 ? a - r : k == 1 ? a : a - l; };
    double qa[9], qb[12];
    for (int t = 0; t < 9; ++t) qa[t] = 0.0;
    for (int t = 0; t < 12; ++t) qb[t] = 0.0;
    tiled_pass_k<3>(I.a_tiles, I.n_a_tiles, I.a_rp, I.a_col, it.a, kprod, kred,
                    [&](int kv, int col) {
                        return (double)(kv == 0 ? xc[col] : kv == 1 ? it.xr[col] : xl[col]);
                    },
                    [&](int i, const double* acc) {
                        const double lo = I.lcs[i], up = I.ucs[i], id = I.iD1[i];
                        const bool lf = isfinite(lo), uf = isfinite(up);
                        const double y = (double)yc[i], yr = (double)it.yr[i], yo = (double)yl[i];
#pragma unroll
                        for (int k = 0; k < 3; ++k) {
                            const double v = id * ray(k, acc[0], acc[1], acc[2]);
                            const double cv = lf && uf ? v : lf ? fmin(v, 0.0) : uf ? fmax(v, 0.0) : 0.0;
                            qa[k] += cv * cv;
                            const double e = ray(k, y, yr, yo);
                            const double ep = pos_part(e), en = neg_part(e);
                            qa[3 + k] += (lf ? lo * ep : 0.0) - (uf ? up * en : 0.0);
                            const double sv = ((lf ? 0.0 : ep) + (uf ? 0.0 : en)) / id;
                            qa[6 + k] += sv * sv;
                        }
                    });
    tiled_pass_k<3>(I.at_tiles, I.n_at_tiles, I.at_rp, I.at_col, it.at, kprod, kred,
                    [&](int kv, int col) {
                        return (double)(kv == 0 ? yc[col] : kv == 1 ? it.yr[col] : yl[col]);
                    },
                    [&](int j, const double* acc) {
                        const double cj = I.cs[j], lo = I.lvs[j], up = I.uvs[j], id = I.iD2[j];
                        const bool lf = isfinite(lo), uf = isfinite(up);
                        const double x = (double)xc[j], xr = (double)it.xr[j], xo = (double)xl[j];
#pragma unroll
                        for (int k = 0; k < 3; ++k) {
                            const double r = -ray(k, acc[0], acc[1], acc[2]);
                            const double rp = pos_part(r), rn = neg_part(r);
                            qb[k] += (lf ? lo * rp : 0.0) - (uf ? up * rn : 0.0);
                            const double sv = id * ((lf ? 0.0 : rp) + (uf ? 0.0 : rn));
                            qb[3 + k] += sv * sv;
                            const double d = ray(k, x, xr, xo);
                            qb[6 + k] += cj * d;
                            const double cv = (lf && uf ? d : lf ? fmin(d, 0.0) : uf ? fmax(d, 0.0) : 0.0) / id;
                            qb[9 + k] += cv * cv;
                        }
                    });
    bsum<9>(qa, sh);
    double ra[9];
    for (int t = 0; t < 9; ++t) ra[t] = sh[t];
    __syncthreads();
    bsum<12>(qb, sh);
    double rb[12];
    for (int t = 0; t < 12; ++t) rb[t] = sh[t];
    __syncthreads();
    for (int k = 0; k < 3; ++k) {
        const double obj = ra[3 + k] + rb[k];
        if (obj > 0.0 && sqrt(ra[6 + k] + rb[3 + k]) <= tol * obj) return CPDLP_PRIMAL_INFEASIBLE;
        const double cd = rb[6 + k];
        if (cd < 0.0 && sqrt(ra[k] + rb[9 + k]) <= tol * -cd) return CPDLP_DUAL_INFEASIBLE;
    }
    return CPDLP_STILL_RUNNING;
}

// ---------------------------------------------------------------- check kernel
// After each chunk of steps: the 5-candidate KKT check in FP64 (errors.jl:65-109),
// termination (termination.jl:89-101), restart_check! / restart! and the primal-weight
// update (pdlp.jl:138-196, restart.jl:40-57, step_size.jl:58-75), and the mixed
// precision's hand-over to FP64.  RAYS (opt-in detect_infeasibility): at a check that is
// not OPTIMAL, NaN / infinite errors stop the instance, and three candidate rays -- the
// iterate minus the restart point, the iterate, the iterate minus the previous one --
// are tested as a Farkas dual ray and as a primal ray, as the plain engine's ray_check.
template <class V, bool RAYS = false>
__global__ __launch_bounds__(BB) void k_batch_check(const int32_t* list, const BInst* insts,
                                                    BState* states, const int32_t* timeout,
                                                    BParams P) {
    __shared__ double sh[32];
    __shared__ double kprod[BT];
    __shared__ double kred[5][BB / WAVE];
    const int slot = list[blockIdx.x];
    const BInst I = insts[slot];
    BState* S = states + slot;
    const int m = I.m, n = I.n, tid = threadIdx.x;
    const int want_phase = std::is_same<V, double>::value ? 1 : 0;
    if (S->setup != 1 || S->status != CPDLP_STILL_RUNNING || S->phase != want_phase) return;
    const Iter<V> it = iter_of(I, V(0));
    double omega = S->omega, eta_sum = S->eta_sum;
    const int cur = S->cur;
    const int64_t passes = S->kkt_passes, total = S->total;
    int64_t inner = S->inner;
    const double reltol = want_phase == 0 && P.precision == CPDLP_PREC_MIXED
                              ? fmax(P.polish_reltol, P.reltol) : P.reltol;
    const bool polish_next = want_phase == 0 && P.precision == CPDLP_PREC_MIXED;
    int32_t status = CPDLP_STILL_RUNNING;
    {
        // ---- check: KKT errors (FP64) of sol, avg, sol_last, avg_last, restart
        const int c = cur, l = 1 - cur;
        constexpr int K = 5;
        // selects, not it.x[c]: a runtime index into the pointer pair sends it to scratch
        V* const xc = c ? it.x[1] : it.x[0];
        V* const yc = c ? it.y[1] : it.y[0];
        V* const xl = l ? it.x[1] : it.x[0];
        V* const yl = l ? it.y[1] : it.y[0];
        const V* X[K] = {xc, it.xa, xl, it.xal, it.xr};
        const V* Y[K] = {yc, it.ya, yl, it.yal, it.yr};
        const int KK = P.pdlp ? K : 1;
        double qa[2 * K + 2], qb[3 * K + 2];
        for (int t2 = 0; t2 < 2 * K + 2; ++t2) qa[t2] = 0.0;
        for (int t2 = 0; t2 < 3 * K + 2; ++t2) qb[t2] = 0.0;
        tiled_pass_k<K>(I.a_tiles, I.n_a_tiles, I.a_rp, I.a_col, it.a, kprod, kred,
                        [&](int kv, int col) { return kv < KK ? (double)X[kv][col] : 0.0; },
                        [&](int i, const double* acc) {
                            const double lo = I.lcs[i], up = I.ucs[i], id = I.iD1[i];
#pragma unroll
                            for (int k = 0; k < K; ++k) {
                                if (k >= KK) break;
                                const double pd = id * (acc[k] - proj_box(acc[k], lo, up));
                                qa[2 * k] += pd * pd;
                                const double yk = (double)Y[k][i];
                                qa[2 * k + 1] += safeprod_left(up, pos_part(-yk)) - safeprod_left(lo, neg_part(-yk));
                            }
                            if (P.pdlp) {
                                const double yr = (double)it.yr[i];
                                const double d0 = (double)Y[0][i] - yr, d1 = (double)Y[1][i] - yr;
                                qa[2 * K] += d0 * d0;
                                qa[2 * K + 1] += d1 * d1;
                            }
                        });
        tiled_pass_k<K>(I.at_tiles, I.n_at_tiles, I.at_rp, I.at_col, it.at, kprod, kred,
                        [&](int kv, int col) { return kv < KK ? (double)Y[kv][col] : 0.0; },
                        [&](int j, const double* acc) {
                            const double cj = I.cs[j], lo = I.lvs[j], up = I.uvs[j], id = I.iD2[j];
#pragma unroll
                            for (int k = 0; k < K; ++k) {
                                if (k >= KK) break;
                                const double g = cj - acc[k];
                                const double r = proj_multiplier(g, lo, up);
                                const double dd = id * (g - r);
                                qb[3 * k] += dd * dd;
                                qb[3 * k + 1] += safeprod_left(up, pos_part(-r)) - safeprod_left(lo, neg_part(-r));
                                qb[3 * k + 2] += cj * (double)X[k][j];
                            }
                            if (P.pdlp) {
                                const double xr = (double)it.xr[j];
                                const double d0 = (double)X[0][j] - xr, d1 = (double)X[1][j] - xr;
                                qb[3 * K] += d0 * d0;
                                qb[3 * K + 1] += d1 * d1;
                            }
                        });
        bsum<2 * K + 2>(qa, sh);
        double ra[2 * K + 2];
        for (int t2 = 0; t2 < 2 * K + 2; ++t2) ra[t2] = sh[t2];
        __syncthreads();
        bsum<3 * K + 2>(qb, sh);
        double rb[3 * K + 2];
        for (int t2 = 0; t2 < 3 * K + 2; ++t2) rb[t2] = sh[t2];
        __syncthreads();
        double e[K][6];
#pragma unroll
        for (int k = 0; k < K; ++k) {   // errors.jl:77-98
            if (k >= KK) break;
            e[k][0] = sqrt(ra[2 * k]);
            e[k][1] = S->primal_scale;
            e[k][2] = sqrt(rb[3 * k]);
            e[k][3] = S->dual_scale;
            const double pc = ra[2 * k + 1], pv = rb[3 * k + 1], cx = rb[3 * k + 2];
            e[k][4] = fabs(cx + pc + pv);
            e[k][5] = 1.0 + fabs(pc + pv) + fabs(cx);
        }
        // termination (termination.jl:89-101); time limit comes from the host flag
        const double rel = rel_of(e[0]);
        if (rel <= reltol) status = CPDLP_OPTIMAL;
        else if (timeout[slot]) status = CPDLP_TIME_LIMIT;
        else if (passes >= P.max_kkt_passes) status = CPDLP_ITERATION_LIMIT;
        if (tid == 0) {
            for (int f = 0; f < 6; ++f) S->err[f] = e[0][f];
        }
        int restart = 0, from_avg = 0, switch_phase = 0, improve = 0, blowup = 0, reached = 0;
        if constexpr (RAYS) {
            if (status != CPDLP_OPTIMAL) {   // uniform: rel and status come from block sums
                const int32_t rs = isfinite(rel) ? batch_rays(I, it, cur, P.inf_tol, kprod, kred, sh)
                                                 : CPDLP_NUMERICAL_ERROR;
                if (rs != CPDLP_STILL_RUNNING) status = rs;
            }
        }
        if (status == CPDLP_STILL_RUNNING && polish_next) {
            // FP32 phase: keep the best iterate; hand over on blow-up (at once, restart
            // from the best) or on stall (after this check's restart decision, continuing)
            if (rel < S->best_rel) improve = 1;
            else if (rel > 100.0 * S->best_rel) switch_phase = blowup = 1;
            else if (S->stall + 1 >= STALL) switch_phase = 1;
        }
        if (status == CPDLP_OPTIMAL && polish_next && P.reltol < reltol) {
            switch_phase = reached = 1;   // continue in FP64 from this state
            status = CPDLP_STILL_RUNNING;
        }
        if (status == CPDLP_STILL_RUNNING && !blowup && !reached && P.pdlp) {
            // restart_check! + should_restart (pdlp.jl:138-170, restart.jl:40-57)
            // candidates compared by value (no pointer into e: keeps e in registers)
            const double a0 = abs_of(e[0], omega), a1 = abs_of(e[1], omega);
            const double a2 = abs_of(e[2], omega), a3 = abs_of(e[3], omega);
            from_avg = !(a0 < a1);
            const double ac = from_avg ? a1 : a0, ar = abs_of(e[4], omega), acl = a2 < a3 ? a2 : a3;
            const bool sufficient = ac <= P.suff * ar;
            const bool necessary = ac <= P.nec * ar;
            const bool no_local = ac > acl;
            const bool long_inner = (double)inner >= P.art * (double)total;
            if (sufficient || (necessary && no_local) || long_inner) {
                restart = 1;
                const double dx = sqrt(from_avg ? rb[3 * K + 1] : rb[3 * K]);
                const double dy = sqrt(from_avg ? ra[2 * K + 1] : ra[2 * K]);
                if (dx > P.zero_tol && dy > P.zero_tol)
                    omega = exp(P.damping * log(dy / dx) + (1.0 - P.damping) * log(omega));
                eta_sum = 0.0;
                inner = 0;
            }
        }
        __syncthreads();
        if (improve) {
            for (int j = tid; j < n; j += BB) I.xb32[j] = (float)X[0][j];
            for (int i = tid; i < m; i += BB) I.yb32[i] = (float)Y[0][i];
            if (tid == 0) {
                S->best_rel = rel;
                S->best_omega = omega;
                S->stall = 0;
            }
        } else if (polish_next && status == CPDLP_STILL_RUNNING && !switch_phase && tid == 0) {
            S->stall += 1;
        }
        if (restart) {   // restart!: sol <- candidate; avg <- 0; restart <- sol
            for (int j = tid; j < n; j += BB) {
                V s0 = xc[j];
                if (from_avg) s0 = it.xa[j];
                xc[j] = s0;
                it.xa[j] = V(0);
                it.xr[j] = s0;
            }
            for (int i = tid; i < m; i += BB) {
                V s0 = yc[i];
                if (from_avg) s0 = it.ya[i];
                yc[i] = s0;
                it.ya[i] = V(0);
                it.yr[i] = s0;
            }
            if (tid == 0) S->restarts += 1;
        }
        __syncthreads();
        if (switch_phase) {
            if (blowup) {
                // FP64 phase from the best FP32 iterate: sol = avg = restart = best
                for (int j = tid; j < n; j += BB) {
                    const double v = (double)I.xb32[j];
                    I.x64[0][j] = I.xa64[j] = I.xr64[j] = v;
                    I.x64[1][j] = I.xal64[j] = 0.0;
                }
                for (int i = tid; i < m; i += BB) {
                    const double v = (double)I.yb32[i];
                    I.y64[0][i] = I.ya64[i] = I.yr64[i] = v;
                    I.y64[1][i] = I.yal64[i] = 0.0;
                }
            } else {
                // continuation: every iterate of the state in FP64, same parity
                for (int j = tid; j < n; j += BB) {
                    I.x64[0][j] = (double)I.x32[0][j];
                    I.x64[1][j] = (double)I.x32[1][j];
                    I.xa64[j] = (double)I.xa32[j];
                    I.xal64[j] = (double)I.xal32[j];
                    I.xr64[j] = (double)I.xr32[j];
                }
                for (int i = tid; i < m; i += BB) {
                    I.y64[0][i] = (double)I.y32[0][i];
                    I.y64[1][i] = (double)I.y32[1][i];
                    I.ya64[i] = (double)I.ya32[i];
                    I.yal64[i] = (double)I.yal32[i];
                    I.yr64[i] = (double)I.yr32[i];
                }
            }
            __syncthreads();
            if (tid == 0) {
                S->phase = 1;
                S->fp32_passes = passes;
                if (blowup) {
                    S->omega = S->best_omega;
                    S->eta_sum = 0.0;
                    S->inner = 0;
                    S->total = 0;
                    S->cur = 0;
                } else {
                    S->omega = omega;
                    S->eta_sum = eta_sum;
                    S->inner = inner;
                }
                S->kkt_passes = passes;
                S->status = CPDLP_STILL_RUNNING;
            }
            return;
        }
    }
    if (tid == 0) {
        S->omega = omega;
        S->eta_sum = eta_sum;
        S->inner = inner;
        S->status = status;
    }
}

// unscaled solution of an instance into a staging buffer: x = D2 .* x~, y = D1 .* y~
__global__ void k_batch_unscale(const BInst* insts, const BState* states, int32_t slot, double* x,
                                double* y) {
    const BInst I = insts[slot];
    const BState* S = states + slot;
    const int c = S->cur;
    const bool f64 = S->phase == 1;
    const double* x64 = c ? I.x64[1] : I.x64[0];
    const double* y64 = c ? I.y64[1] : I.y64[0];
    const float* x32 = c ? I.x32[1] : I.x32[0];
    const float* y32 = c ? I.y32[1] : I.y32[0];
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < I.n; j += gridDim.x * blockDim.x)
        x[j] = I.D2[j] * (f64 ? x64[j] : (double)x32[j]);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < I.m; i += gridDim.x * blockDim.x)
        y[i] = I.D1[i] * (f64 ? y64[i] : (double)y32[i]);
}

// an instance's loop state for cpdlp_batch_take_state: the five iterates of its current
// precision (sol, sol_last, avg, avg_last, restart), scaled, widened to double, and its
// scaling D2 / D1; one workgroup grid-strides over n and m
__global__ void k_batch_state(const BInst* insts, const BState* states, int32_t slot, double* x, double* y,
                              double* D2, double* D1) {
    const BInst I = insts[slot];
    const BState* S = states + slot;
    const int c = S->cur;
    const bool f64 = S->phase == 1;
    const double* xd[5] = {c ? I.x64[1] : I.x64[0], c ? I.x64[0] : I.x64[1], I.xa64, I.xal64, I.xr64};
    const double* yd[5] = {c ? I.y64[1] : I.y64[0], c ? I.y64[0] : I.y64[1], I.ya64, I.yal64, I.yr64};
    const float* xf[5] = {c ? I.x32[1] : I.x32[0], c ? I.x32[0] : I.x32[1], I.xa32, I.xal32, I.xr32};
    const float* yf[5] = {c ? I.y32[1] : I.y32[0], c ? I.y32[0] : I.y32[1], I.ya32, I.yal32, I.yr32};
    const int64_t n = I.n, m = I.m;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
#pragma unroll
        for (int k = 0; k < 5; ++k) x[k * n + j] = f64 ? xd[k][j] : (double)xf[k][j];
        D2[j] = I.D2[j];
    }
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
#pragma unroll
        for (int k = 0; k < 5; ++k) y[k * m + i] = f64 ? yd[k][i] : (double)yf[k][i];
        D1[i] = I.D1[i];
    }
}

// ================================================================ host side

static double bnow() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

static size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

static_assert(BT == TILE_NNZ && BB == BLOCK, "batch tiles are built by build_tiles");

// most tiles build_tiles can make: a tile closes at BB rows, at the end, or because the
// next row would pass BT nonzeros -- then it and the next tile hold > BT together, so
// there are fewer than nnz / (BT / 2) of those
static int64_t max_tiles(int64_t rows, int64_t nnz) {
    return (rows + BB - 1) / BB + (nnz + BT / 2 - 1) / (BT / 2) + 1;
}

// bytes of one instance in the arena (same order as carve())
static size_t footprint(int64_t m, int64_t n, int64_t nnz) {
    size_t b = 0;
    auto add = [&](int64_t count, size_t sz) { b += align256((size_t)std::max<int64_t>(count, 1) * sz); };
    add(m + 1, 4); add(nnz, 4); add(n + 1, 4); add(nnz, 4); add(nnz, 4);       // ints
    add(nnz, 4);                                                                // ecol
    add(max_tiles(m, nnz), sizeof(TileMeta)); add(max_tiles(n, nnz), sizeof(TileMeta));
    for (int k = 0; k < 5; ++k) add(nnz, 8);                                    // nz, acsc, atcsc, a64, at64
    add(nnz, 4); add(nnz, 4);                                                   // a32, at32
    for (int k = 0; k < 3; ++k) add(n, 8);                                      // c, lv, uv
    for (int k = 0; k < 2; ++k) add(m, 8);                                      // lc, uc
    for (int k = 0; k < 3; ++k) add(n, 8);                                      // cs, lvs, uvs
    for (int k = 0; k < 2; ++k) add(m, 8);                                      // lcs, ucs
    add(m, 8); add(n, 8); add(m, 8); add(n, 8);                                 // D1, D2, iD1, iD2
    for (int k = 0; k < 3; ++k) add(n, 4);                                      // c32, lv32, uv32
    for (int k = 0; k < 2; ++k) add(m, 4);                                      // lc32, uc32
    for (int k = 0; k < 6; ++k) add(n, 4);                                      // x32 x2, xa, xal, xr, xb
    for (int k = 0; k < 6; ++k) add(m, 4);                                      // y32 x2, ya, yal, yr, yb
    for (int k = 0; k < 6; ++k) add(n, 8);                                      // x64 x2, xa, xal, xr, sn
    for (int k = 0; k < 6; ++k) add(m, 8);                                      // y64 x2, ya, yal, yr, sm
    return b;
}

static BInst carve(char* base, int64_t m, int64_t n, int64_t nnz) {
    BInst I{};
    size_t off = 0;
    auto take = [&](int64_t count, size_t sz) {
        char* p = base + off;
        off += align256((size_t)std::max<int64_t>(count, 1) * sz);
        return (void*)p;
    };
    I.m = (int32_t)m;
    I.n = (int32_t)n;
    I.nnz = (int32_t)nnz;
    I.a_rp = (int32_t*)take(m + 1, 4);
    I.a_col = (int32_t*)take(nnz, 4);
    I.at_rp = (int32_t*)take(n + 1, 4);
    I.at_col = (int32_t*)take(nnz, 4);
    I.perm = (int32_t*)take(nnz, 4);
    I.ecol = (int32_t*)take(nnz, 4);
    I.a_tiles = (const TileMeta*)take(max_tiles(m, nnz), sizeof(TileMeta));
    I.at_tiles = (const TileMeta*)take(max_tiles(n, nnz), sizeof(TileMeta));
    I.nz = (double*)take(nnz, 8);
    I.acsc = (double*)take(nnz, 8);
    I.atcsc = (double*)take(nnz, 8);
    I.a64 = (double*)take(nnz, 8);
    I.at64 = (double*)take(nnz, 8);
    I.a32 = (float*)take(nnz, 4);
    I.at32 = (float*)take(nnz, 4);
    I.c = (double*)take(n, 8);
    I.lv = (double*)take(n, 8);
    I.uv = (double*)take(n, 8);
    I.lc = (double*)take(m, 8);
    I.uc = (double*)take(m, 8);
    I.cs = (double*)take(n, 8);
    I.lvs = (double*)take(n, 8);
    I.uvs = (double*)take(n, 8);
    I.lcs = (double*)take(m, 8);
    I.ucs = (double*)take(m, 8);
    I.D1 = (double*)take(m, 8);
    I.D2 = (double*)take(n, 8);
    I.iD1 = (double*)take(m, 8);
    I.iD2 = (double*)take(n, 8);
    I.c32 = (float*)take(n, 4);
    I.lv32 = (float*)take(n, 4);
    I.uv32 = (float*)take(n, 4);
    I.lc32 = (float*)take(m, 4);
    I.uc32 = (float*)take(m, 4);
    I.x32[0] = (float*)take(n, 4);
    I.x32[1] = (float*)take(n, 4);
    I.xa32 = (float*)take(n, 4);
    I.xal32 = (float*)take(n, 4);
    I.xr32 = (float*)take(n, 4);
    I.xb32 = (float*)take(n, 4);
    I.y32[0] = (float*)take(m, 4);
    I.y32[1] = (float*)take(m, 4);
    I.ya32 = (float*)take(m, 4);
    I.yal32 = (float*)take(m, 4);
    I.yr32 = (float*)take(m, 4);
    I.yb32 = (float*)take(m, 4);
    I.x64[0] = (double*)take(n, 8);
    I.x64[1] = (double*)take(n, 8);
    I.xa64 = (double*)take(n, 8);
    I.xal64 = (double*)take(n, 8);
    I.xr64 = (double*)take(n, 8);
    I.sn = (double*)take(n, 8);
    I.y64[0] = (double*)take(m, 8);
    I.y64[1] = (double*)take(m, 8);
    I.ya64 = (double*)take(m, 8);
    I.yal64 = (double*)take(m, 8);
    I.yr64 = (double*)take(m, 8);
    I.sm = (double*)take(m, 8);
    return I;
}

// lanes per row: a power of two near half the mean row length, in [2, 64]
static int lanes_for(int64_t nnz, int64_t rows) {
    const double mean = rows > 0 ? (double)nnz / rows : 1.0;
    int L = 2;
    while (L < 64 && L * 2 <= mean) L *= 2;
    return L;
}

// first-fit arena allocator with coalescing free list (offsets in bytes)
struct Arena {
    size_t size = 0;
    std::map<size_t, size_t> free_;   // offset -> length
    size_t used = 0;
    void init(size_t s) {
        size = s;
        free_.clear();
        free_[0] = s;
        used = 0;
    }
    bool alloc(size_t len, size_t& off) {
        for (auto it = free_.begin(); it != free_.end(); ++it)
            if (it->second >= len) {
                off = it->first;
                const size_t rest = it->second - len;
                free_.erase(it);
                if (rest) free_[off + len] = rest;
                used += len;
                return true;
            }
        return false;
    }
    void release(size_t off, size_t len) {
        auto it = free_.emplace(off, len).first;
        used -= len;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

struct BHost {
    int64_t id;
    int32_t slot, m, n, nnz;
    size_t off, bytes;
    bool finished = false, polled = false;
    bool ready = false;   // its setup kernel has completed and its state is on the host
    double t_start = 0.0, t_done = 0.0;
    int32_t status = CPDLP_STILL_RUNNING;
};

struct Batch {
    cpdlp_params P;
    int device = 0;
    std::vector<hipStream_t> streams;
    DevBuf<char> arena_mem;
    Arena arena;
    int32_t max_slots = 0;
    DevBuf<BInst> d_inst;
    DevBuf<BState> d_state;
    DevBuf<int32_t> d_timeout, d_list;
    // setup runs on its own stream, asynchronously: rounds go on for the instances that
    // are ready while new ones are set up.  An admission's setup is split by size into
    // launches run smallest first, so the small instances become ready long before the
    // largest one's power method ends.
    struct SetupChunk {
        hipEvent_t ev = nullptr;
        std::vector<int32_t> slots;   // host copy stays alive until the launch completes
        std::vector<int64_t> ids;
        DevBuf<int32_t> list;
    };
    hipStream_t setup_stream = nullptr;
    std::deque<SetupChunk> setups;   // in flight, in launch order
    std::vector<BInst> h_inst;
    std::vector<BState> h_state;
    std::vector<int32_t> h_timeout, free_slots;
    std::map<int64_t, BHost> inst;   // by id
    std::deque<int64_t> done;
    std::vector<int32_t> pending_setup;
    std::vector<int64_t> pending_ids;
    int64_t next_id = 0;
    int32_t hi_slot = 0;   // slots [0, hi_slot) may be in use
    size_t lds_max = 65536;  // dynamic LDS a workgroup may take (device limit minus static use)
    // wide workgroups (WIDE * BB threads) for instances of at least wide_nnz nonzeros;
    // their kernel's static LDS is WIDE times larger, so their budget is per phase
    static constexpr int WIDE = 4;
    int32_t wide_nnz = 16384;
    size_t wide_lds_max[2] = {0, 0};
    int cus = 256;   // compute units of the device

    Batch(const cpdlp_params& p, int64_t bytes) : P(p) {
        int ndev = 0;
        HIP_CHECK(hipGetDeviceCount(&ndev));
        if (P.device < 0 || P.device >= ndev) throw Error(CPDLP_ERR_ARGUMENT, "invalid device ordinal");
        if (P.check_every <= 0) throw Error(CPDLP_ERR_ARGUMENT, "check_every must be positive");
        if (P.detect_infeasibility && !(P.infeasibility_tol > 0.0))
            throw Error(CPDLP_ERR_ARGUMENT, "infeasibility_tol must be positive");
        device = P.device;
        HIP_CHECK(hipSetDevice(device));
        streams.assign((size_t)std::min(8, std::max(1, P.batch_streams)), nullptr);
        for (auto& s : streams) HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        HIP_CHECK(hipStreamCreateWithFlags(&setup_stream, hipStreamNonBlocking));
        if (bytes <= 0) throw Error(CPDLP_ERR_ARGUMENT, "arena size must be positive");
        arena_mem.alloc((size_t)bytes);
        arena.init((size_t)bytes);
        int lds = 0;
        HIP_CHECK(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
        // room for the step kernel's static LDS (BT products in FP64 + the reductions)
        constexpr size_t static_lds = BT * sizeof(double) + 4096;
        lds_max = (size_t)lds > static_lds ? (size_t)lds - static_lds : 0;
        // instances whose y and xbar need more LDS than this gather from global memory
        // (L2) instead: more workgroups fit on a CU
        if (const char* e = std::getenv("CPDLP_BATCH_LDS_BYTES")) lds_max = std::min<size_t>(lds_max, std::strtoull(e, nullptr, 10));
        for (int ph = 0; ph < 2; ++ph) {
            const size_t st = (size_t)WIDE * (BT + BB / WAVE) * (ph ? 8 : 4) + 4096;
            wide_lds_max[ph] = std::min(lds_max, (size_t)lds > st ? (size_t)lds - st : 0);
        }
        if (const char* e = std::getenv("CPDLP_BATCH_WIDE_NNZ")) wide_nnz = (int32_t)std::strtol(e, nullptr, 10);
        HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
        max_slots = 1 << 16;
        d_inst.alloc(max_slots);
        d_state.alloc(max_slots);
        d_timeout.alloc(max_slots);
        d_list.alloc(max_slots);
        h_inst.resize(max_slots);
        h_state.resize(max_slots);
        h_timeout.assign(max_slots, 0);
        for (int32_t s = max_slots - 1; s >= 0; --s) free_slots.push_back(s);
    }
    ~Batch() {
        (void)hipSetDevice(device);
        for (auto& s : streams)
            if (s) (void)hipStreamSynchronize(s);   // buffers go back to the block cache
        if (setup_stream) (void)hipStreamSynchronize(setup_stream);
        for (auto& s : streams)
            if (s) (void)hipStreamDestroy(s);
        if (setup_stream) (void)hipStreamDestroy(setup_stream);
        for (auto& c : setups)
            if (c.ev) (void)hipEventDestroy(c.ev);
    }

    BParams params() const {
        BParams b{};
        b.check_every = P.check_every;
        b.ruiz_iter = P.ruiz_iter;
        b.pdlp = P.algorithm == CPDLP_ALGO_PDLP;
        b.precision = P.precision == CPDLP_PREC_FP64 ? CPDLP_PREC_FP64 : CPDLP_PREC_MIXED;
        b.cp_alpha = P.chambolle_pock_alpha;
        b.invnorm = P.invnorm_scaling;
        b.damping = P.primal_weight_damping;
        b.zero_tol = P.zero_tol;
        b.suff = P.sufficient_decay;
        b.nec = P.necessary_decay;
        b.art = P.artificial_decay;
        b.reltol = P.termination_reltol;
        b.polish_reltol = P.polish_reltol;
        b.powm_tol = P.powm_tol;
        b.max_kkt_passes = P.max_kkt_passes;
        b.powm_maxiter = P.powm_maxiter;
        b.seed = P.powm_seed;
        b.inf_tol = P.infeasibility_tol;
        return b;
    }

    int64_t add(const cpdlp_lp& lp, const double* x0 = nullptr, const double* y0 = nullptr) {
        HIP_CHECK(hipSetDevice(device));
        std::vector<int32_t> cp, rv, rp, col, perm;
        read_csc(lp.m, lp.n, lp.index_base, lp.index_bits, lp.colptr, lp.rowval, cp, rv);
        const int64_t m = lp.m, n = lp.n, nnz = cp[n];
        if (m + n > 65536) throw Error(CPDLP_ERR_ARGUMENT, "batched instances need m + n <= 65536");
        if (nnz > 0 && !lp.nzval) throw Error(CPDLP_ERR_ARGUMENT, "nzval is NULL");
        for (const void* v : {(const void*)lp.c, (const void*)lp.lv, (const void*)lp.uv})
            if (n > 0 && !v) throw Error(CPDLP_ERR_ARGUMENT, "NULL variable vector");
        for (const void* v : {(const void*)lp.lc, (const void*)lp.uc})
            if (m > 0 && !v) throw Error(CPDLP_ERR_ARGUMENT, "NULL constraint vector");
        const size_t bytes = footprint(m, n, nnz);
        if (bytes > arena.size)   // can never fit: not a transient shortage
            throw Error(CPDLP_ERR_ARGUMENT, "LP needs " + std::to_string(bytes) + " bytes, more than the whole batch arena (" +
                                                std::to_string(arena.size) + ")");
        if (free_slots.empty()) throw Error(CPDLP_ERR_MEMORY, "no free batch slot");
        // host-side preparation before any arena bytes are taken
        csc_to_csr(m, n, cp, rv, rp, col, perm);
        std::vector<TileMeta> ta, tat;
        build_tiles(rp.data(), 0, (int32_t)m, ta);
        build_tiles(cp.data(), 0, (int32_t)n, tat);
        if ((int64_t)ta.size() > max_tiles(m, nnz) || (int64_t)tat.size() > max_tiles(n, nnz))
            throw Error(CPDLP_ERR_STATE, "internal: tile count above its bound");
        std::vector<int32_t> ecol(nnz);
        for (int64_t j = 0; j < n; ++j)
            for (int32_t k = cp[j]; k < cp[j + 1]; ++k) ecol[k] = (int32_t)j;
        size_t off = 0;
        if (!arena.alloc(bytes, off)) throw Error(CPDLP_ERR_MEMORY, "batch arena full");
        struct RegionGuard {   // returns the region to the arena unless the add completes
            Arena& a;
            size_t off, len;
            bool armed = true;
            ~RegionGuard() {
                if (armed) a.release(off, len);
            }
        } region{arena, off, bytes};
        BInst I = carve(arena_mem.p + off, m, n, nnz);
        I.La = lanes_for(nnz, m);
        I.Lat = lanes_for(nnz, n);
        I.n_a_tiles = (int32_t)ta.size();
        I.n_at_tiles = (int32_t)tat.size();
        const hipStream_t s = streams[0];
        auto up = [&](void* dst, const void* src, size_t b) {
            if (b) HIP_CHECK(hipMemcpyAsync(dst, src, b, hipMemcpyHostToDevice, s));
        };
        up(I.a_rp, rp.data(), (m + 1) * 4);
        up(I.a_col, col.data(), nnz * 4);
        up(I.at_rp, cp.data(), (n + 1) * 4);
        up(I.at_col, rv.data(), nnz * 4);
        up(I.perm, perm.data(), nnz * 4);
        up(I.ecol, ecol.data(), nnz * 4);
        up((void*)I.a_tiles, ta.data(), ta.size() * sizeof(TileMeta));
        up((void*)I.at_tiles, tat.data(), tat.size() * sizeof(TileMeta));
        up(I.nz, lp.nzval, nnz * 8);
        up(I.c, lp.c, n * 8);
        up(I.lv, lp.lv, n * 8);
        up(I.uv, lp.uv, n * 8);
        up(I.lc, lp.lc, m * 8);
        up(I.uc, lp.uc, m * 8);
        // a start goes into the restart buffers, where the setup kernel scales it
        std::vector<double> zeros;
        if (x0 || y0) {
            I.warm = 1;
            if (!x0 || !y0) zeros.assign((size_t)std::max(m, n), 0.0);
            up(I.xr64, x0 ? x0 : zeros.data(), n * 8);
            up(I.yr64, y0 ? y0 : zeros.data(), m * 8);
        }
        const int32_t slot = free_slots.back();   // taken below, once the copies have completed
        h_inst[slot] = I;
        BState st{};
        st.status = CPDLP_STILL_RUNNING;
        st.setup = 0;
        h_state[slot] = st;
        h_timeout[slot] = 0;
        HIP_CHECK(hipMemcpyAsync(d_inst.p + slot, &h_inst[slot], sizeof(BInst), hipMemcpyHostToDevice, s));
        HIP_CHECK(hipMemcpyAsync(d_state.p + slot, &h_state[slot], sizeof(BState), hipMemcpyHostToDevice, s));
        HIP_CHECK(hipMemcpyAsync(d_timeout.p + slot, &h_timeout[slot], 4, hipMemcpyHostToDevice, s));
        HIP_CHECK(hipStreamSynchronize(s));   // host vectors above go out of scope
        free_slots.pop_back();
        hi_slot = std::max(hi_slot, slot + 1);
        region.armed = false;
        BHost h;
        h.id = next_id++;
        h.slot = slot;
        h.m = (int32_t)m;
        h.n = (int32_t)n;
        h.nnz = (int32_t)nnz;
        h.off = off;
        h.bytes = bytes;
        h.t_start = bnow();
        inst[h.id] = h;
        pending_setup.push_back(slot);
        pending_ids.push_back(h.id);
        return h.id;
    }

    void sync_all() {
        for (auto& s : streams) HIP_CHECK(hipStreamSynchronize(s));
    }

    int64_t run(int32_t checks) {
        HIP_CHECK(hipSetDevice(device));
        if (checks <= 0) throw Error(CPDLP_ERR_ARGUMENT, "max_checks must be positive");
        const BParams bp = params();
        std::vector<int32_t> lists;
        // 1. setup of newly admitted instances, on the setup stream; a round never waits
        //    for it unless nothing else is ready to run
        bool newly_ready = false;
        auto retire_done = [&](bool wait_first) {   // completed launches, in order
            while (!setups.empty()) {
                SetupChunk& c = setups.front();
                if (wait_first) {
                    HIP_CHECK(hipEventSynchronize(c.ev));
                    wait_first = false;
                } else if (hipEventQuery(c.ev) != hipSuccess) {
                    break;
                }
                for (int64_t id : c.ids) {
                    auto it = inst.find(id);
                    if (it != inst.end()) it->second.ready = true;
                }
                (void)hipEventDestroy(c.ev);
                setups.pop_front();
                newly_ready = true;
            }
        };
        retire_done(false);
        if (setups.empty() && !pending_setup.empty()) {
            // smallest first, in launches whose largest instance has at most 1/64, 1/8 and
            // all of the largest one's work (nonzeros + rows + columns)
            std::vector<size_t> order(pending_setup.size());
            for (size_t i = 0; i < order.size(); ++i) order[i] = i;
            auto w = [&](size_t i) {
                const BInst& I = h_inst[pending_setup[i]];
                return (double)I.nnz + I.m + I.n;
            };
            std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return w(a) < w(b); });
            const double wmax = w(order.back());
            size_t k = 0;
            for (const double frac : {1.0 / 64, 1.0 / 8, 1.0}) {
                SetupChunk c;
                while (k < order.size() && (frac == 1.0 || w(order[k]) <= frac * wmax)) {
                    c.slots.push_back(pending_setup[order[k]]);
                    c.ids.push_back(pending_ids[order[k]]);
                    ++k;
                }
                if (c.slots.empty()) continue;
                c.list.alloc(c.slots.size());
                HIP_CHECK(hipMemcpyAsync(c.list.p, c.slots.data(), c.slots.size() * 4,
                                         hipMemcpyHostToDevice, setup_stream));
                // wide workgroups for the large instances, which are at the end of the list
                int n_narrow = 0;
                while (n_narrow < (int)c.slots.size() && h_inst[c.slots[n_narrow]].nnz < wide_nnz) ++n_narrow;
                const int n_wide = (int)c.slots.size() - n_narrow;
                if (n_narrow > 0)
                    hipLaunchKernelGGL(k_batch_setup<1>, dim3(n_narrow), dim3(BB), 0, setup_stream, c.list.p,
                                       d_inst.p, d_state.p, bp);
                if (n_wide > 0)
                    hipLaunchKernelGGL(k_batch_setup<WIDE>, dim3(n_wide), dim3(WIDE * BB), 0, setup_stream,
                                       c.list.p + n_narrow, d_inst.p, d_state.p, bp);
                HIP_CHECK(hipGetLastError());
                HIP_CHECK(hipEventCreateWithFlags(&c.ev, hipEventDisableTiming));
                HIP_CHECK(hipEventRecord(c.ev, setup_stream));
                setups.push_back(std::move(c));
            }
            pending_setup.clear();
            pending_ids.clear();
        }
        // instances of setup launches still running become ready one by one: each
        // publishes its state and then its setup flag (k_batch_setup), so a slow power
        // method holds back only its own instance
        if (!setups.empty() && hi_slot > 0) {
            d_state.download(h_state.data(), hi_slot, streams[0]);
            HIP_CHECK(hipStreamSynchronize(streams[0]));
            for (const SetupChunk& c : setups)
                for (size_t q = 0; q < c.ids.size(); ++q) {
                    auto it = inst.find(c.ids[q]);
                    if (it != inst.end() && !it->second.ready && h_state[c.slots[q]].setup == 1) {
                        it->second.ready = true;
                        newly_ready = true;
                    }
                }
        }
        bool any_ready = false;
        for (auto& kv : inst)
            if (kv.second.ready && !kv.second.finished) {
                any_ready = true;
                break;
            }
        if (!any_ready && !setups.empty()) retire_done(true);
        if (newly_ready && hi_slot > 0) {   // their set-up states (phase, eta, ...) to the host,
                                            // read after their flags: whole
            d_state.download(h_state.data(), hi_slot, streams[0]);
            HIP_CHECK(hipStreamSynchronize(streams[0]));
        }
        // 2. time limits (host clock per instance)
        const double t = bnow();
        bool any_timeout = false;
        for (auto& kv : inst) {
            BHost& h = kv.second;
            if (h.finished) continue;
            if (t - h.t_start >= P.time_limit && !h_timeout[h.slot]) {
                h_timeout[h.slot] = 1;
                any_timeout = true;
            }
        }
        if (any_timeout) {
            d_timeout.upload(h_timeout.data(), hi_slot, streams[0]);
            HIP_CHECK(hipStreamSynchronize(streams[0]));
        }
        // 3. one launch per (phase, class); an instance's y and xbar take
        //    (n + m) * sizeof(V) bytes of LDS, in the smallest class that holds them.
        //    Classes [0, NLDS + 1): one workgroup of BB threads per instance, LDS budgets
        //    lds_bytes, then gathers from L2; classes NCLS - 2 and NCLS - 1: instances with
        //    at least wide_nnz nonzeros, WIDE * BB threads each, vectors in LDS when they
        //    fit the wide kernel's budget, else gathers from L2.
        constexpr int NLDS = 4, NCLS = NLDS + 3, CW_LDS = NCLS - 2, CW_GLB = NCLS - 1;
        const size_t lds_bytes[NLDS + 1] = {std::min<size_t>(16384, lds_max), std::min<size_t>(40960, lds_max),
                                            std::min<size_t>(98304, lds_max), lds_max, 0};
        std::vector<int32_t> cls[2][NCLS];
        size_t wide_lds[2] = {0, 0};   // dynamic LDS of each phase's wide LDS launch
        // once fewer than two instances per CU are left, the GPU has room to spare: half
        // the threshold then (the results are the same bits either way)
        int64_t n_active = 0;
        for (auto& kv : inst) n_active += !kv.second.finished && kv.second.ready;
        const int32_t wide_min = n_active <= 2 * (int64_t)cus ? wide_nnz / 2 : wide_nnz;
        for (auto& kv : inst) {
            const BHost& h = kv.second;
            if (h.finished || !h.ready) continue;
            const BState& st = h_state[h.slot];
            const int ph = st.phase;
            const size_t lds = (size_t)(h.n + h.m) * (ph ? 8 : 4);
            int c = 0;
            if (h.nnz >= wide_min) {
                c = lds <= wide_lds_max[ph] ? CW_LDS : CW_GLB;
                if (c == CW_LDS) wide_lds[ph] = std::max(wide_lds[ph], lds);
            } else {
                while (c < NLDS && lds > lds_bytes[c]) ++c;   // class NLDS: gathers from L2
            }
            cls[ph][c].push_back(h.slot);
        }
        // largest instances first: workgroups are dispatched in list order, so the long
        // ones start early and the short ones fill the CUs at the end of each launch
        for (auto& ph_cls : cls)
            for (auto& v : ph_cls)
                std::stable_sort(v.begin(), v.end(), [&](int32_t a, int32_t b) {
                    return h_inst[a].nnz + h_inst[a].m + h_inst[a].n > h_inst[b].nnz + h_inst[b].m + h_inst[b].n;
                });
        lists.clear();
        std::vector<std::pair<int, int>> spans;   // (offset, count) into lists per launch
        for (int ph = 0; ph < 2; ++ph)
            for (int c = 0; c < NCLS; ++c) {
                spans.emplace_back((int)lists.size(), (int)cls[ph][c].size());
                lists.insert(lists.end(), cls[ph][c].begin(), cls[ph][c].end());
            }
        int64_t active = (int64_t)lists.size();
        if (active > 0) {
            HIP_CHECK(hipMemcpyAsync(d_list.p, lists.data(), lists.size() * 4, hipMemcpyHostToDevice, streams[0]));
            hipEvent_t ready;
            HIP_CHECK(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
            HIP_CHECK(hipEventRecord(ready, streams[0]));
            hipStream_t gs[2 * NCLS];
            int si = 0;
            for (int g = 0; g < 2 * NCLS; ++g) {
                gs[g] = nullptr;
                if (spans[g].second == 0) continue;
                gs[g] = streams[si++ % streams.size()];
                HIP_CHECK(hipStreamWaitEvent(gs[g], ready, 0));
            }
            // Work-budgeted round: an instance of work w (nonzeros + rows + columns, x2 in
            // FP64) runs floor(budget / w) checks, at least 1 and at most `checks`; budget
            // = max(largest w * max(1, checks / 4), checks * smallest w): the largest
            // instance runs a quarter of `checks` while the smaller ones run up to
            // `checks` in the same time, and instances of similar size all run `checks`.
            // Lists are sorted by decreasing w, so the instances of check k are a suffix
            // of their group's list.
            auto work = [&](int32_t slot, int ph) {
                const BInst& I = h_inst[slot];
                return (double)(I.nnz + I.m + I.n) * (ph ? 2.0 : 1.0);
            };
            double wmax = 0.0, wmin = INFINITY;
            for (int ph = 0; ph < 2; ++ph)
                for (int c = 0; c < NCLS; ++c)
                    if (!cls[ph][c].empty()) {
                        wmax = std::max(wmax, work(cls[ph][c].front(), ph));
                        wmin = std::min(wmin, work(cls[ph][c].back(), ph));
                    }
            const double budget = std::max(wmax * std::max(1, checks / 4), (double)checks * wmin);
            int start[2 * NCLS] = {};
            for (int chk = 0; chk < checks; ++chk)
                for (int ph = 0; ph < 2; ++ph)
                    for (int c = 0; c < NCLS; ++c) {
                        const int g = ph * NCLS + c;
                        const auto sp = spans[g];
                        if (sp.second == 0) continue;
                        const std::vector<int32_t>& v = cls[ph][c];
                        while (chk > 0 && start[g] < sp.second && work(v[start[g]], ph) * (chk + 1) > budget)
                            ++start[g];
                        const int cnt = sp.second - start[g];
                        if (cnt <= 0) continue;
                        const hipStream_t s = gs[g];
                        const int32_t* L = d_list.p + sp.first + start[g];
                        const dim3 grid(cnt), blk(BB), wblk(WIDE * BB);
                        if (ph == 0) {
                            if (c < NLDS)
                                hipLaunchKernelGGL((k_batch_steps<float, true>), grid, blk, lds_bytes[c], s, L,
                                                   d_inst.p, d_state.p, bp);
                            else if (c == NLDS)
                                hipLaunchKernelGGL((k_batch_steps<float, false>), grid, blk, 0, s, L, d_inst.p,
                                                   d_state.p, bp);
                            else if (c == CW_LDS)
                                hipLaunchKernelGGL((k_batch_steps<float, true, WIDE>), grid, wblk, wide_lds[0], s,
                                                   L, d_inst.p, d_state.p, bp);
                            else
                                hipLaunchKernelGGL((k_batch_steps<float, false, WIDE>), grid, wblk, 0, s, L,
                                                   d_inst.p, d_state.p, bp);
                            if (P.detect_infeasibility)
                                hipLaunchKernelGGL((k_batch_check<float, true>), grid, blk, 0, s, L, d_inst.p,
                                                   d_state.p, d_timeout.p, bp);
                            else
                                hipLaunchKernelGGL(k_batch_check<float>, grid, blk, 0, s, L, d_inst.p, d_state.p,
                                                   d_timeout.p, bp);
                        } else {
                            if (c < NLDS)
                                hipLaunchKernelGGL((k_batch_steps<double, true>), grid, blk, lds_bytes[c], s, L,
                                                   d_inst.p, d_state.p, bp);
                            else if (c == NLDS)
                                hipLaunchKernelGGL((k_batch_steps<double, false>), grid, blk, 0, s, L, d_inst.p,
                                                   d_state.p, bp);
                            else if (c == CW_LDS)
                                hipLaunchKernelGGL((k_batch_steps<double, true, WIDE>), grid, wblk, wide_lds[1], s,
                                                   L, d_inst.p, d_state.p, bp);
                            else
                                hipLaunchKernelGGL((k_batch_steps<double, false, WIDE>), grid, wblk, 0, s, L,
                                                   d_inst.p, d_state.p, bp);
                            if (P.detect_infeasibility)
                                hipLaunchKernelGGL((k_batch_check<double, true>), grid, blk, 0, s, L, d_inst.p,
                                                   d_state.p, d_timeout.p, bp);
                            else
                                hipLaunchKernelGGL(k_batch_check<double>, grid, blk, 0, s, L, d_inst.p, d_state.p,
                                                   d_timeout.p, bp);
                        }
                        HIP_CHECK(hipGetLastError());
                    }
            sync_all();
            (void)hipEventDestroy(ready);
        }
        // 4. states back; harvest finished instances
        if (hi_slot > 0) {
            d_state.download(h_state.data(), hi_slot, streams[0]);
            HIP_CHECK(hipStreamSynchronize(streams[0]));
        }
        int64_t running = 0;
        const double t2 = bnow();
        for (auto& kv : inst) {
            BHost& h = kv.second;
            if (h.finished) continue;
            if (!h.ready) {   // being set up (its state on the host is not valid yet)
                ++running;
                continue;
            }
            const BState& st = h_state[h.slot];
            if (st.status != CPDLP_STILL_RUNNING) {
                h.finished = true;
                h.status = st.status;
                h.t_done = t2;
                done.push_back(h.id);
            } else {
                ++running;
            }
        }
        return running;
    }

    BHost& get_host(int64_t id) {
        auto it = inst.find(id);
        if (it == inst.end()) throw Error(CPDLP_ERR_ARGUMENT, "unknown batch instance id");
        return it->second;
    }

    void get(int64_t id, double* x, double* y, cpdlp_stats* st) {
        HIP_CHECK(hipSetDevice(device));
        BHost& h = get_host(id);
        if (x || y) {
            if (!h.ready) throw Error(CPDLP_ERR_STATE, "instance not set up yet: run the batch first");
            DevBuf<double> xo(std::max<int32_t>(h.n, 1)), yo(std::max<int32_t>(h.m, 1));
            hipLaunchKernelGGL(k_batch_unscale, dim3(64), dim3(256), 0, streams[0], d_inst.p, d_state.p, h.slot,
                               xo.p, yo.p);
            HIP_CHECK(hipGetLastError());
            if (x) xo.download(x, h.n, streams[0]);
            if (y) yo.download(y, h.m, streams[0]);
            HIP_CHECK(hipStreamSynchronize(streams[0]));
        }
        if (st) {
            const BState& s = h_state[h.slot];
            *st = cpdlp_stats{};
            st->err = cpdlp_kkt{s.err[0], s.err[1], s.err[2], s.err[3], s.err[4], s.err[5]};
            st->time_elapsed = (h.finished ? h.t_done : bnow()) - h.t_start;
            st->kkt_passes = s.kkt_passes;
            st->termination_status = s.status;
            st->format_A = st->format_At = CPDLP_FMT_CSR;
            st->eta = s.eta;
            st->omega = s.omega;
            st->powm_iters = s.powm_iters;
            st->restarts = s.restarts;
            st->fp32_passes = s.phase == 1 ? s.fp32_passes : s.kkt_passes;
        }
    }

    void release(int64_t id) {
        BHost& h = get_host(id);
        if (!h.finished) throw Error(CPDLP_ERR_STATE, "instance still running");
        arena.release(h.off, h.bytes);
        free_slots.push_back(h.slot);
        inst.erase(id);
    }

    // a running instance's loop state, then the instance leaves the batch (run() has
    // completed its rounds, so no kernel reads its slot or arena region any more)
    void take_state(int64_t id, cpdlp_state* st, double* x, double* y, double* D2, double* D1) {
        HIP_CHECK(hipSetDevice(device));
        BHost& h = get_host(id);
        if (!h.ready) throw Error(CPDLP_ERR_STATE, "instance not set up yet: run the batch first");
        if (h.finished) throw Error(CPDLP_ERR_STATE, "instance finished: take its result with cpdlp_batch_get");
        if ((h.n && (!x || !D2)) || (h.m && (!y || !D1))) throw Error(CPDLP_ERR_ARGUMENT, "NULL state buffer");
        const int64_t n = h.n, m = h.m;
        DevBuf<double> xs(std::max<int64_t>(5 * n, 1)), ys(std::max<int64_t>(5 * m, 1));
        DevBuf<double> d2(std::max<int64_t>(n, 1)), d1(std::max<int64_t>(m, 1));
        hipLaunchKernelGGL(k_batch_state, dim3(64), dim3(256), 0, streams[0], d_inst.p, d_state.p, h.slot, xs.p,
                           ys.p, d2.p, d1.p);
        HIP_CHECK(hipGetLastError());
        BState s{};
        HIP_CHECK(hipMemcpyAsync(&s, d_state.p + h.slot, sizeof(BState), hipMemcpyDeviceToHost, streams[0]));
        if (n) {
            xs.download(x, 5 * n, streams[0]);
            d2.download(D2, n, streams[0]);
        }
        if (m) {
            ys.download(y, 5 * m, streams[0]);
            d1.download(D1, m, streams[0]);
        }
        HIP_CHECK(hipStreamSynchronize(streams[0]));
        if (s.status != CPDLP_STILL_RUNNING) throw Error(CPDLP_ERR_STATE, "instance finished: run the batch first");
        if (st) {
            st->eta = s.eta;
            st->omega = s.omega;
            st->eta_sum = s.eta_sum;
            st->kkt_passes = s.kkt_passes;
            st->inner = s.inner;
            st->total = s.total;
            st->restarts = s.restarts;
            st->fp32_passes = s.phase == 1 ? s.fp32_passes : s.kkt_passes;
        }
        arena.release(h.off, h.bytes);
        free_slots.push_back(h.slot);
        inst.erase(id);
    }
};

}  // namespace cpdlp

using namespace cpdlp;

struct cpdlp_batch {
    std::unique_ptr<Batch> b;
};

extern "C" {

int cpdlp_batch_create(const cpdlp_params* params, int64_t hbm_bytes, cpdlp_batch** out) {
    return guarded([&] {
        if (!out) throw Error(CPDLP_ERR_ARGUMENT, "NULL output");
        cpdlp_params p;
        if (params) p = *params;
        else cpdlp_default_params(&p, CPDLP_ALGO_PDLP);
        auto* b = new cpdlp_batch;
        try {
            b->b = std::make_unique<Batch>(p, hbm_bytes);
        } catch (...) {
            delete b;
            throw;
        }
        *out = b;
    });
}

#define BATCH_OR_THROW(bp) \
    if (!(bp) || !(bp)->b) throw Error(CPDLP_ERR_ARGUMENT, "NULL batch");

int cpdlp_batch_add(cpdlp_batch* b, const cpdlp_lp* lp, int64_t* id) {
    return cpdlp_batch_add_start(b, lp, nullptr, nullptr, id);
}

int cpdlp_batch_add_start(cpdlp_batch* b, const cpdlp_lp* lp, const double* x0, const double* y0,
                          int64_t* id) {
    return guarded([&] {
        BATCH_OR_THROW(b);
        if (!lp || !id) throw Error(CPDLP_ERR_ARGUMENT, "NULL argument");
        *id = b->b->add(*lp, x0, y0);
    });
}

int64_t cpdlp_batch_footprint(int64_t m, int64_t n, int64_t nnz) {
    return (int64_t)footprint(m, n, nnz);
}

int cpdlp_batch_run(cpdlp_batch* b, int32_t max_checks, int64_t* running) {
    return guarded([&] {
        BATCH_OR_THROW(b);
        const int64_t r = b->b->run(max_checks);
        if (running) *running = r;
    });
}

int cpdlp_batch_poll(cpdlp_batch* b, int64_t* ids, int64_t cap, int64_t* count) {
    return guarded([&] {
        BATCH_OR_THROW(b);
        int64_t k = 0;
        auto& d = b->b->done;
        while (!d.empty() && k < cap) {
            if (ids) ids[k] = d.front();
            d.pop_front();
            ++k;
        }
        if (count) *count = k;
    });
}

int cpdlp_batch_take_state(cpdlp_batch* b, int64_t id, cpdlp_state* st, double* x, double* y, double* D2,
                           double* D1) {
    return guarded([&] {
        BATCH_OR_THROW(b);
        b->b->take_state(id, st, x, y, D2, D1);
    });
}

int cpdlp_batch_get(cpdlp_batch* b, int64_t id, double* x, double* y, cpdlp_stats* st) {
    return guarded([&] {
        BATCH_OR_THROW(b);
        b->b->get(id, x, y, st);
    });
}

int cpdlp_batch_release(cpdlp_batch* b, int64_t id) {
    return guarded([&] {
        BATCH_OR_THROW(b);
        b->b->release(id);
    });
}

int cpdlp_batch_memory(cpdlp_batch* b, int64_t* used, int64_t* size) {
    return guarded([&] {
        BATCH_OR_THROW(b);
        if (used) *used = (int64_t)b->b->arena.used;
        if (size) *size = (int64_t)b->b->arena.size;
    });
}

void cpdlp_batch_destroy(cpdlp_batch* b) { delete b; }

}  // extern "C"
