// solver.hip -- libcoolpdlp_hip.so: handle, setup (scaling, step size), the PDLP/PDHG
// loop and the C ABI declared in include/coolpdlp_hip.h.
//
// Data layout in HBM (internal order, see layout.hpp): problem vectors c, lv, uv, D2
// over the n internal columns and lc, uc, D1 over the m internal rows, FP64 (plus
// FP32 copies for the FP32/mixed loop); CSR of A and of A' with int32 indices.  The
// iterate set of the reference's PDLPState (pdlp.jl:20-43) -- sol, sol_last,
// sol_avg, sol_avg_last, sol_restart -- lives on the device for the whole solve; only
// the check scalars (about 30 doubles) come back to the host once per check.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <type_traits>

#include "kernels.hpp"
#include "layout.hpp"
#include "solver.hpp"

namespace cpdlp {

static thread_local std::string g_last_error;
void set_last_error(const char* msg) { g_last_error = msg ? msg : ""; }

static double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// ================================================================ setup kernels

// norm over CSC column j (rows ascending): p = inf -> max|v|; p = 1 -> sequential
// sum of |v|; p = 2 -> sqrt of sequential sum of squares; else (sum |v|^p)^(1/p).
// d = iszero(norm) ? 1 : inv(sqrt(norm))   (preconditioning.jl:102-111)
__device__ __forceinline__ double pn_acc(double acc, double v, double p) {
    const double a = fabs(v);
    if (isinf(p)) return a > acc ? a : acc;
    if (p == 1.0) return acc + a;
    if (p == 2.0) return acc + a * a;
    return acc + pow(a, p);
}
__device__ __forceinline__ double pn_fin(double acc, double p) {
    if (isinf(p) || p == 1.0) return acc;
    if (p == 2.0) return sqrt(acc);
    return pow(acc, 1.0 / p);
}

__global__ void k_col_norm(int32_t n, const int32_t* colptr, const double* vals, double p,
                           double* d) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    double acc = 0.0;
    for (int k = colptr[j]; k < colptr[j + 1]; ++k) acc = pn_acc(acc, vals[k], p);
    const double nv = pn_fin(acc, p);
    d[j] = nv == 0.0 ? 1.0 : 1.0 / sqrt(nv);
}

// row norms of A from the A' values: row i's entries in CSR(A) order through perm
__global__ void k_row_norm(int32_t m, const int32_t* rowptr, const int32_t* perm,
                           const double* vals, double p, double* d) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    double acc = 0.0;
    for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) acc = pn_acc(acc, vals[perm[k]], p);
    const double nv = pn_fin(acc, p);
    d[i] = nv == 0.0 ? 1.0 : 1.0 / sqrt(nv);
}

// precondition(cons, prec), preconditioning.jl:37-43:
// A <- (d1_i * a_ij) * d2_j ; A' <- (d2_j * a'_ji) * d1_i   (CSC(A) layout)
__global__ void k_apply_prec(int32_t n, const int32_t* colptr, const int32_t* rowval,
                             const double* a_in, const double* at_in, const double* d1,
                             const double* d2, double* a_out, double* at_out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const double dj = d2[j];
    for (int k = colptr[j]; k < colptr[j + 1]; ++k) {
        const double di = d1[rowval[k]];
        const double a = a_in[k], at = at_in[k];
        a_out[k] = (di * a) * dj;
        at_out[k] = (dj * at) * di;
    }
}

__global__ void k_mul_into(int32_t n, double* D, const double* d) {   // D .*= d
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) D[i] = d[i] * D[i];
}
__global__ void k_fill(int64_t n, double* v, double a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = a;
}
template <class T>
__global__ void k_fill_t(int64_t n, T* v, T a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = a;
}

// values of an internal side: out[k] = src[vmap[k]] (vmap < 0: padding -> 0)
template <class T>
__global__ void k_gather_vals(int64_t nnz, const int32_t* vmap, const double* src, T* out) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nnz) return;
    const int32_t s = vmap[k];
    out[k] = s >= 0 ? (T)src[s] : T(0);
}

// internal scaled column vectors: j = col_of_int[s]; c~ = D2*c, lv~ = lv/D2, uv~ = uv/D2
// (precondition(milp, prec), preconditioning.jl:61-91)
__global__ void k_scale_cols(int32_t n, const int32_t* col_of_int, const double* D2,
                             const double* c, const double* lv, const double* uv, double* cs,
                             double* lvs, double* uvs, double* D2s, double* invD2s) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int j = col_of_int[s];
    const double d = D2[j];
    cs[s] = d * c[j];
    lvs[s] = lv[j] / d;
    uvs[s] = uv[j] / d;
    D2s[s] = d;
    invD2s[s] = 1.0 / d;
}
__global__ void k_scale_rows(int32_t m, const int32_t* row_of_int, const double* D1,
                             const double* lc, const double* uc, double* lcs, double* ucs,
                             double* D1s, double* invD1s) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= m) return;
    const int i = row_of_int[r];
    const double d = D1[i];
    lcs[r] = d * lc[i];
    ucs[r] = d * uc[i];
    D1s[r] = d;
    invD1s[r] = 1.0 / d;
}

// original -> internal and scaled: out[s] = v[of_int[s]] / D[of_int[s]]
template <class T>
__global__ void k_to_internal(int32_t n, const int32_t* of_int, const double* v, const double* D,
                              T* out) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int j = of_int[s];
    out[s] = (T)(v[j] / D[j]);
}
// internal -> original and unscaled: out[j] = D[j] * v[int_of[j]]  (unprecondition)
template <class T>
__global__ void k_from_internal(int32_t n, const int32_t* int_of, const double* D, const T* v,
                                double* out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    out[j] = D[j] * (double)v[int_of[j]];
}
// four vectors interleaved, out[4 i + k] = v_k[i]: one 4-wide store per entry
template <class V>
__global__ void k_interleave4(int64_t n, const V* a, const V* b, const V* c, const V* d, V* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    VecK<V, 4> t;
    t.v[0] = a[i];
    t.v[1] = b[i];
    t.v[2] = c[i];
    t.v[3] = d[i];
    *(VecK<V, 4>*)(out + 4 * i) = t;
}
template <class S, class T>
__global__ void k_convert(int64_t n, const S* in, T* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (T)in[i];
}

// constant scales of kkt_errors!: sum ((1/D1) combine(lc, uc))^2 and sum ((1/D2) c)^2
__global__ void k_scale_sums(int32_t m, int32_t n, const double* lc, const double* uc,
                             const double* invD1, const double* c, const double* invD2,
                             double* partials) {
    double q[4] = {0, 0, 0, 0};
    for (int r = blockIdx.x * BLOCK + threadIdx.x; r < m; r += gridDim.x * BLOCK) {
        const double v = invD1[r] * combine(lc[r], uc[r]);
        q[0] += v * v;
        const double w = combine(lc[r], uc[r]);
        q[2] += w * w;
    }
    for (int s = blockIdx.x * BLOCK + threadIdx.x; s < n; s += gridDim.x * BLOCK) {
        const double v = invD2[s] * c[s];
        q[1] += v * v;
        q[3] += c[s] * c[s];
    }
    block_sum<4>(q, partials + (size_t)blockIdx.x * 4);
}

// power-method step finished by the reduction that produces its sums (POW_FIN1: theta =
// x.Ax and |Ax| from the two sums; POW_FIN2: residual, iteration count and convergence)
enum : int32_t { POW_NONE = 0, POW_FIN1 = 1, POW_FIN2 = 2 };
__device__ __forceinline__ void pow_fin(PowState* st, int32_t fin, const double* sums) {
    if (fin == POW_FIN1) {
        st->theta = sums[0];
        st->nrm = sqrt(sums[1]);
    } else if (fin == POW_FIN2) {
        st->residual = sqrt(sums[0]);
        st->iter += 1;
        st->done = (st->iter > st->maxiter) || (st->residual <= st->tol * fabs(st->theta));
    }
}

// deterministic final reduction of partials[nb][nq] -> out[nq] (one block); each thread
// adds its partials b, b + BLOCK, ... in that order, eight loads in flight at a time.
// With st: a no-op once the power method has converged, and pow_fin(fin) at the end.
__global__ __launch_bounds__(BLOCK) void k_reduce_partials(int32_t nb, int32_t nq,
                                                           const double* partials, double* out,
                                                           PowState* st, int32_t fin) {
    if (st && st->done) return;
    __shared__ double red[BLOCK / WAVE];
    double first[2] = {0.0, 0.0};
    for (int q = 0; q < nq; ++q) {
        double s = 0.0;
        int b = threadIdx.x;
        for (; b + 7 * BLOCK < nb; b += 8 * BLOCK) {
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = partials[(size_t)(b + u * BLOCK) * nq + q];
#pragma unroll
            for (int u = 0; u < 8; ++u) s += v[u];
        }
        for (; b < nb; b += BLOCK) s += partials[(size_t)b * nq + q];
        s = wave_sum(s);
        if ((threadIdx.x & (WAVE - 1)) == 0) red[threadIdx.x / WAVE] = s;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0.0;
            for (int w = 0; w < BLOCK / WAVE; ++w) t += red[w];
            out[q] = t;
            if (q < 2) first[q] = t;
        }
        __syncthreads();
    }
    if (st && threadIdx.x == 0) pow_fin(st, fin, first);
}

// ---------------------------------------------------------------- power method kernels
// start vector: sum of 12 counter-hashed uniforms - 6 (exact adds; same bits as the
// oracle's orc_powm_start on any IEEE device)
__host__ __device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
template <class V>
__global__ void k_powm_start(int32_t n, const int32_t* col_of_int, uint64_t seed, V* x) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const uint64_t i = (uint64_t)col_of_int[s];
    double acc = 0.0;
    for (int r = 0; r < 12; ++r) {
        const uint64_t h = splitmix64(seed * 0x100000001B3ull ^ splitmix64(i * 16u + (uint64_t)r));
        acc += (double)(h >> 11) * 0x1.0p-53;
    }
    x[s] = (V)(acc - 6.0);
}
template <class V>
__global__ void k_sumsq(int32_t n, const V* x, double* partials) {
    double q[1] = {0.0};
    for (int s = blockIdx.x * BLOCK + threadIdx.x; s < n; s += gridDim.x * BLOCK)
        q[0] += (double)(x[s] * x[s]);
    block_sum<1>(q, partials + blockIdx.x);
}
template <class V>
__global__ void k_scale_by_inv_sqrt(int32_t n, V* x, const double* sumsq) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const V inv = (V)(1.0 / sqrt(sumsq[0]));
    x[s] = x[s] * inv;
}
// theta and |Ax| from the sums (one thread)
// residual sum of (Ax - theta x)^2 and x <- Ax * inv(|Ax|)
template <class V>
__global__ void k_pow_update(int32_t n, V* x, const V* Ax, const PowState* st, double* partials) {
    if (st->done) return;
    double q[1] = {0.0};
    const V th = (V)st->theta;
    const V inv = (V)(1.0 / st->nrm);
    for (int s = blockIdx.x * BLOCK + threadIdx.x; s < n; s += gridDim.x * BLOCK) {
        const V a = Ax[s];
        const V r = a + (-th) * x[s];
        q[0] += (double)(r * r);
        x[s] = a * inv;
    }
    block_sum<1>(q, partials + blockIdx.x);
}

// ---------------------------------------------------------------- restart!
// pdlp.jl:189-192: sol <- candidate; avg <- 0; restart <- sol
template <class V>
__global__ void k_restart(int64_t len, int32_t from_avg, V* sol, V* avg, V* rst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= len) return;
    V s = sol[i];
    if (from_avg) s = avg[i];
    sol[i] = s;
    avg[i] = V(0);
    rst[i] = s;
}

// ================================================================ device side of a matrix

template <class V>
MatView<V> DevSide::view(bool allow_slices) const {
    MatView<V> v;
    v.rows = rows;
    v.rowptr = rowptr.p;
    v.col = col.p;
    v.val = vals<V>();
    v.n_long = n_long;
    v.n_long_blocks = ceil_div(n_long, LONG_ROWS_PER_BLOCK);
    v.tiles = tiles.p;
    v.n_tiles = n_tiles;
    if (lg_built && lg_on) {
        v.lg_tiles = lg_tiles.p;
        v.lg_n_tiles = lg_n_tiles;
        v.lg_col = lg_col.p;
        v.lg_val = lg_vals<V>();
    }
    v.format = format;
    v.ell_width = ell_w;
    v.ell_col = ell_col.p;
    v.ell_val = ell_vals<V>();
    v.ns = sl_ns;
    v.sl_all = format == CPDLP_FMT_XSLICE_ALL;
    v.sl_rows = v.sl_all ? rows : n_slice;
    v.sl_n_long = sl_n_group;
    v.sl_stride = sl_built + 1;
    v.sl_rowptr = sl_rowptr.p;
    v.sl_col = sl_col.p;
    v.sl_val = sl_vals<V>();
    v.sl_tiles = sl_tiles.p;
    v.sl_tile_off = sl_tile_off.p;
    v.sl_tile_n = v.sl_all ? sl_tile_n.p : sl_tile_n_pre.p;
    v.sl_max_tiles = v.sl_all ? sl_max_tiles : sl_max_tiles_pre;
    v.sl_part = sl_part.p;
    v.xs_tiles = xs_tiles.p;
    v.n_xs_tiles = n_xs_tiles;
    // both slice variants run the sliced kernels
    if (format == CPDLP_FMT_XSLICE_ALL) v.format = CPDLP_FMT_XSLICE;
    if (v.format == CPDLP_FMT_XSLICE && !allow_slices) v.format = CPDLP_FMT_CSR;
    return v;
}
template <>
const double* DevSide::vals<double>() const { return val64.p; }
template <>
const float* DevSide::vals<float>() const { return val32.p; }
template <>
const double* DevSide::ell_vals<double>() const { return ell_val64.p; }
template <>
const float* DevSide::ell_vals<float>() const { return ell_val32.p; }
template <>
const double* DevSide::sl_vals<double>() const { return sl_val64.p; }
template <>
const float* DevSide::sl_vals<float>() const { return sl_val32.p; }
template <>
const double* DevSide::lg_vals<double>() const { return lg_val64.p; }
template <>
const float* DevSide::lg_vals<float>() const { return lg_val32.p; }

void DevSide::build_line_tiles(const HostLineTiles& T, hipStream_t s) {
    lg_n_tiles = (int32_t)T.tiles.size();
    lg_tiles.upload(T.tiles, s);
    lg_col.upload(T.col, s);
    lg_vmap.upload(T.vmap, s);
    lg_built = true;
    lg_on = true;
    HIP_CHECK(hipStreamSynchronize(s));
}

// NS column slices of rows [0, nrows); NS divides the 8 XCDs of the MI355X, and slice s
// is summed on the XCDs x with x % NS == s (workgroups go to the XCDs round-robin)
void DevSide::build_slices(const HostSide& S, int32_t nrows, int32_t NS, hipStream_t s, int32_t n_group) {
    if (NS != 2 && NS != 4 && NS != 8) throw Error(CPDLP_ERR_ARGUMENT, "slice count must be 2, 4 or 8");
    HostSlices H;
    cpdlp::build_slices(S, NS, nrows, H, n_group);
    sl_ns = NS;
    sl_n_group = H.n_group;
    sl_built = nrows;
    sl_max_tiles = H.max_tiles;
    sl_max_tiles_pre = H.max_tiles_pre;
    sl_nnz = (int64_t)H.col.size();
    sl_rowptr.upload(H.rowptr, s);
    sl_col.alloc(std::max<int64_t>(sl_nnz, 1));
    sl_col.upload(H.col.data(), H.col.size(), s);
    sl_vmap.alloc(std::max<int64_t>(sl_nnz, 1));
    sl_vmap.upload(H.vmap.data(), H.vmap.size(), s);
    sl_tiles.upload(H.tiles, s);
    sl_tile_off.upload(H.tile_off, s);
    std::vector<int32_t> tn(NS);
    for (int32_t k = 0; k < NS; ++k) tn[k] = H.tile_off[k + 1] - H.tile_off[k];
    sl_tile_n.upload(tn, s);
    sl_tile_n_pre.upload(H.tile_n_pre, s);
    sl_part.alloc(std::max<int64_t>((int64_t)SL_KMAX * NS * nrows, 1));
    HIP_CHECK(hipStreamSynchronize(s));
}

void DevSide::upload_structure(const HostSide& S, hipStream_t s) {
    rows = S.rows;
    cols = S.cols;
    nnz = S.nnz;
    n_long = S.n_long;
    n_slice = std::max(S.n_slice, S.n_long);
    n_tiles = (int32_t)S.tiles.size();
    n_xs_tiles = (int32_t)S.xs_tiles.size();
    rowptr.upload(S.rowptr, s);
    col.alloc(std::max<int64_t>(nnz, 1));
    col.upload(S.col.data(), S.col.size(), s);
    vmap.alloc(std::max<int64_t>(nnz, 1));
    vmap.upload(S.vmap.data(), S.vmap.size(), s);
    tiles.upload(S.tiles, s);
    xs_tiles.upload(S.xs_tiles, s);
}

void DevSide::fill_values(const double* src_csc, bool want32, hipStream_t s) {
    if (val64.n < (size_t)std::max<int64_t>(nnz, 1)) val64.alloc(std::max<int64_t>(nnz, 1));
    if (nnz)
        hipLaunchKernelGGL(k_gather_vals<double>, dim3(ceil_div(nnz, BLOCK)), dim3(BLOCK), 0, s,
                           nnz, vmap.p, src_csc, val64.p);
    if (want32) {
        if (val32.n < (size_t)std::max<int64_t>(nnz, 1)) val32.alloc(std::max<int64_t>(nnz, 1));
        if (nnz)
            hipLaunchKernelGGL(k_gather_vals<float>, dim3(ceil_div(nnz, BLOCK)), dim3(BLOCK), 0, s,
                               nnz, vmap.p, src_csc, val32.p);
    }
    if (sl_ns > 0 && sl_nnz > 0) {
        if (sl_val64.n < (size_t)sl_nnz) sl_val64.alloc(sl_nnz);
        hipLaunchKernelGGL(k_gather_vals<double>, dim3(ceil_div(sl_nnz, BLOCK)), dim3(BLOCK), 0, s,
                           sl_nnz, sl_vmap.p, src_csc, sl_val64.p);
        if (want32) {
            if (sl_val32.n < (size_t)sl_nnz) sl_val32.alloc(sl_nnz);
            hipLaunchKernelGGL(k_gather_vals<float>, dim3(ceil_div(sl_nnz, BLOCK)), dim3(BLOCK), 0,
                               s, sl_nnz, sl_vmap.p, src_csc, sl_val32.p);
        }
    }
    if (lg_built && nnz) {
        if (lg_val64.n < (size_t)nnz) lg_val64.alloc(nnz);
        hipLaunchKernelGGL(k_gather_vals<double>, dim3(ceil_div(nnz, BLOCK)), dim3(BLOCK), 0, s,
                           nnz, lg_vmap.p, src_csc, lg_val64.p);
        if (want32) {
            if (lg_val32.n < (size_t)nnz) lg_val32.alloc(nnz);
            hipLaunchKernelGGL(k_gather_vals<float>, dim3(ceil_div(nnz, BLOCK)), dim3(BLOCK), 0, s,
                               nnz, lg_vmap.p, src_csc, lg_val32.p);
        }
    }
    if (ell_w > 0) {
        const int64_t len = (int64_t)ell_w * rows;
        if (ell_val64.n < (size_t)len) ell_val64.alloc(len);
        hipLaunchKernelGGL(k_gather_vals<double>, dim3(ceil_div(len, BLOCK)), dim3(BLOCK), 0, s,
                           len, ell_vmap.p, src_csc, ell_val64.p);
        if (want32) {
            if (ell_val32.n < (size_t)len) ell_val32.alloc(len);
            hipLaunchKernelGGL(k_gather_vals<float>, dim3(ceil_div(len, BLOCK)), dim3(BLOCK), 0, s,
                               len, ell_vmap.p, src_csc, ell_val32.p);
        }
    }
    HIP_CHECK(hipGetLastError());
}

void DevSide::build_ell(const HostSide& S, hipStream_t s, double max_fill) {
    std::vector<int32_t> ecol, evmap;
    const int32_t w = cpdlp::build_ell(S, max_fill, ecol, evmap);
    if (w <= 0) {
        ell_w = 0;
        return;
    }
    ell_w = w;
    ell_col.upload(ecol, s);
    ell_vmap.upload(evmap, s);
    HIP_CHECK(hipStreamSynchronize(s));
}

template MatView<double> DevSide::view<double>(bool) const;
template MatView<float> DevSide::view<float>(bool) const;

// ================================================================ iterate state

template <class V>
void IterState<V>::alloc(int64_t n, int64_t m) {
    const int64_t nn = std::max<int64_t>(n, 1), mm = std::max<int64_t>(m, 1);
    for (int b = 0; b < 2; ++b) {
        x[b].alloc(nn);
        y[b].alloc(mm);
    }
    xbar.alloc(nn);
    xavg.alloc(nn);
    yavg.alloc(mm);
    xavgl.alloc(nn);
    yavgl.alloc(mm);
    xrst.alloc(nn);
    yrst.alloc(mm);
    tmp_n.alloc(nn);
    tmp_m.alloc(mm);
    cur = 0;
}
template struct IterState<double>;
template struct IterState<float>;

// ================================================================ handle

static int64_t wrap_cube(int64_t n) {
    const uint64_t u = (uint64_t)n;
    return (int64_t)(u * u * u);
}

Handle::Handle(const cpdlp_lp& lp, const cpdlp_params& prm) : P(prm) {
    no_rst_cache = std::getenv("CPDLP_KKT_FULL_CHECK") != nullptr;
    no_aos = std::getenv("CPDLP_KKT_SEPARATE") != nullptr;
    if (P.check_every <= 0) throw Error(CPDLP_ERR_ARGUMENT, "check_every must be positive");
    if (P.precision < 0 || P.precision > 2) throw Error(CPDLP_ERR_ARGUMENT, "unknown precision");
    if (P.format < CPDLP_FMT_AUTO || P.format > CPDLP_FMT_XSLICE_ALL)
        throw Error(CPDLP_ERR_ARGUMENT, "unknown format");
    if (P.algorithm != CPDLP_ALGO_PDLP && P.algorithm != CPDLP_ALGO_PDHG)
        throw Error(CPDLP_ERR_ARGUMENT, "unknown algorithm");
    if (P.detect_infeasibility && !(P.infeasibility_tol > 0.0))
        throw Error(CPDLP_ERR_ARGUMENT, "infeasibility_tol must be positive");
    if (P.long_row <= 0) P.long_row = 64;
    if (P.long_row > TILE_NNZ) P.long_row = TILE_NNZ;
    m = lp.m;
    n = lp.n;
    for (const void* v : {(const void*)lp.c, (const void*)lp.lv, (const void*)lp.uv})
        if (n > 0 && !v) throw Error(CPDLP_ERR_ARGUMENT, "NULL variable vector");
    for (const void* v : {(const void*)lp.lc, (const void*)lp.uc})
        if (m > 0 && !v) throw Error(CPDLP_ERR_ARGUMENT, "NULL constraint vector");
    std::vector<int32_t> cp, rv;
    read_csc(m, n, lp.index_base, lp.index_bits, lp.colptr, lp.rowval, cp, rv);
    nnz = cp[n];
    if (nnz > 0 && !lp.nzval) throw Error(CPDLP_ERR_ARGUMENT, "nzval is NULL");
    int ndev = 0;
    HIP_CHECK(hipGetDeviceCount(&ndev));
    if (P.device < 0 || P.device >= ndev) throw Error(CPDLP_ERR_ARGUMENT, "invalid device ordinal");
    device = P.device;
    HIP_CHECK(hipSetDevice(device));
    HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    HIP_CHECK(hipEventCreate(&ev0));
    HIP_CHECK(hipEventCreate(&ev1));
    // row and column order: first touch unless CPDLP_ORDER names "density" or "given",
    // for measurements
    RowOrder order = RowOrder::first_touch;
    if (const char* e = std::getenv("CPDLP_ORDER")) {
        const std::string v(e);
        if (v == "density") order = RowOrder::density;
        else if (v == "given") order = RowOrder::given;
        else if (v != "first_touch") throw Error(CPDLP_ERR_ARGUMENT, "CPDLP_ORDER must be first_touch, density or given");
    }
    // rows longer than slice_row (16; CPDLP_SLICE_ROW for measurements) are cut by the XCD
    // slices in the sliced format; in CSR only those longer than long_row leave the
    // row-ordered tiles.  Measured on the large LP (primal / dual step, us): rows longer
    // than 64 sliced 105.8 / 142.8, than 32 104.8 / 132.7, 20 104.1 / 128.8, 16 102.2 /
    // 126.7, 12 102.6 / 127.2, 8 109.1 / 128.5 (the split also orders the other side's
    // rows longer than it by length, not by first touch: at 8 most of A''s).
    int32_t slice_row = 16;
    if (const char* e = std::getenv("CPDLP_SLICE_ROW")) slice_row = (int32_t)std::strtol(e, nullptr, 10);
    if (slice_row <= 0) throw Error(CPDLP_ERR_ARGUMENT, "CPDLP_SLICE_ROW must be positive");
    L = build_layout(m, n, std::move(cp), std::move(rv), P.long_row, order, slice_row);
    const hipStream_t s = stream;
    auto up64 = [&](DevBuf<double>& d, const double* h, int64_t len) {
        d.alloc(std::max<int64_t>(len, 1));
        d.upload(h, len, s);
    };
    up64(d_nzval, lp.nzval, nnz);
    up64(d_c, lp.c, n);
    up64(d_lv, lp.lv, n);
    up64(d_uv, lp.uv, n);
    up64(d_lc, lp.lc, m);
    up64(d_uc, lp.uc, m);
    d_colptr.upload(L.colptr, s);
    d_rowval.alloc(std::max<int64_t>(nnz, 1));
    d_rowval.upload(L.rowval.data(), nnz, s);
    d_csr_rowptr.upload(L.csr_rowptr, s);
    d_perm.alloc(std::max<int64_t>(nnz, 1));
    d_perm.upload(L.perm.data(), nnz, s);
    auto upi = [&](DevBuf<int32_t>& d, const std::vector<int32_t>& h) {
        d.alloc(std::max<size_t>(h.size(), 1));
        d.upload(h.data(), h.size(), s);
    };
    upi(d_row_of_int, L.row_of_int);
    upi(d_col_of_int, L.col_of_int);
    upi(d_int_of_row, L.int_of_row);
    upi(d_int_of_col, L.int_of_col);
    A.upload_structure(L.A, s);
    At.upload_structure(L.At, s);
    if (P.format == CPDLP_FMT_ELL || P.format == CPDLP_FMT_AUTO) {
        // AUTO considers ELL only when padding stays within 1.5x nnz; an explicit ELL
        // request accepts up to 8x (beyond that the side stays CSR, see stats.format_*)
        const double fill = P.format == CPDLP_FMT_ELL ? 8.0 : 1.5;
        A.build_ell(L.A, s, fill);
        At.build_ell(L.At, s, fill);
    }
    // XCD slices: XSLICE cuts the rows longer than slice_row, XSLICE_ALL every row.  AUTO
    // builds slices of every row for a large side whose gathered vector outgrows one XCD's
    // 4 MiB L2 (the tuner then times both variants), else of the rows longer than
    // slice_row when it has any and its gathered vector exceeds 2 MiB (a smaller one stays
    // in L2 without them: small LPs skip building and timing them).
    const size_t elem = P.precision == CPDLP_PREC_FP64 ? 8 : 4;
    // slices per side: 8 (one per XCD) unless CPDLP_SLICES="<A>,<A'>" (2, 4 or 8) says
    // otherwise, for measurements
    int32_t ns_of[2] = {8, 8};
    if (const char* e = std::getenv("CPDLP_SLICES")) {
        const std::string spec(e);
        const size_t comma = spec.find(',');
        ns_of[0] = (int32_t)std::strtol(spec.substr(0, comma).c_str(), nullptr, 10);
        ns_of[1] = comma == std::string::npos ? ns_of[0] : (int32_t)std::strtol(spec.substr(comma + 1).c_str(), nullptr, 10);
    }
    // In the slices, the parts of rows longer than CPDLP_SLICE_GROUP_ROW entries (default
    // 1,024, at least long_row) are summed by 16-lane groups, the shorter ones a thread per
    // part (a part holds about 1/8 of its row).  Measured on the large power-law LP
    // (ms per iteration, bench.py --long-row, which set both thresholds then): 0.2308 at
    // 32, 0.2295 at 64, 0.2282 at 128, 0.2268 at 256, 0.2261 at 512, 0.2259 at 1,024,
    // 0.2301 at 2,048.
    int32_t group_row = 1024;
    if (const char* e = std::getenv("CPDLP_SLICE_GROUP_ROW")) group_row = (int32_t)std::strtol(e, nullptr, 10);
    group_row = std::max(group_row, P.long_row);
    int side_k = 0;
    for (auto side : {std::make_pair(&A, &L.A), std::make_pair(&At, &L.At)}) {
        const HostSide& H = *side.second;
        const int32_t ns = ns_of[side_k++];
        int32_t n_group = 0;   // the long rows are longest first: a prefix
        while (n_group < H.n_long && H.rowptr[n_group + 1] - H.rowptr[n_group] > group_row) ++n_group;
        const bool big = (size_t)H.cols * elem > ((size_t)4 << 20) && H.nnz >= ((int64_t)4 << 20);
        if (P.format == CPDLP_FMT_XSLICE_ALL || (P.format == CPDLP_FMT_AUTO && big))
            side.first->build_slices(H, H.rows, ns, s, n_group);
        else if (P.format == CPDLP_FMT_XSLICE ||
                 (P.format == CPDLP_FMT_AUTO && H.n_slice > 0 && (size_t)H.cols * elem > ((size_t)2 << 20)))
            side.first->build_slices(H, H.n_slice, ns, s, n_group);
    }
    // Line-grouped tiles of the short rows (the steps of a CSR side): built for a side
    // whose short rows (>= 1M entries) would need >= 10% fewer gather requests in them,
    // counted over 128-byte lines of the loop's element size; the format tuner then keeps
    // them only if they time faster.  CPDLP_LINE_TILES=0 never builds them, =1 always
    // does (and keeps them).
    lg_mode = -1;
    if (const char* e = std::getenv("CPDLP_LINE_TILES")) lg_mode = std::strtol(e, nullptr, 10) != 0 ? 1 : 0;
    if (lg_mode != 0 && (P.format == CPDLP_FMT_AUTO || P.format == CPDLP_FMT_CSR)) {
        for (auto side : {std::make_pair(&A, &L.A), std::make_pair(&At, &L.At)}) {
            const HostSide& H = *side.second;
            const int64_t short_nnz = H.nnz - H.rowptr[H.n_long];
            if (H.n_long == H.rows || !line_tiles_packable(H)) continue;
            if (lg_mode < 0 && short_nnz < ((int64_t)1 << 20)) continue;
            HostLineTiles T;
            if (lg_mode < 0) {   // estimate from every 8th tile before building them all
                line_group_tiles(H, (int32_t)(128 / elem), false, T, 8);
                if ((double)T.req_lines > 0.9 * (double)T.req_rows) continue;
            }
            line_group_tiles(H, (int32_t)(128 / elem), true, T);
            side.first->build_line_tiles(T, s);
        }
    }
    HIP_CHECK(hipStreamSynchronize(s));
    // free the big host-side copies that are no longer needed
    std::vector<int32_t>().swap(L.A.col);
    std::vector<int32_t>().swap(L.A.vmap);
    std::vector<int32_t>().swap(L.At.col);
    std::vector<int32_t>().swap(L.At.vmap);
    std::vector<int32_t>().swap(L.perm);
}

Handle::~Handle() {
    (void)hipSetDevice(device);
    if (stream) (void)hipStreamSynchronize(stream);   // buffers go back to the block cache
    for (auto& g : graphs)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
}

int Handle::grid_n() const { return std::max(1, ceil_div(n, BLOCK)); }
int Handle::grid_m() const { return std::max(1, ceil_div(m, BLOCK)); }

void Handle::sync() { HIP_CHECK(hipStreamSynchronize(stream)); }

// reduce nb blocks of nq partials on the device and copy the nq sums to the host
void Handle::reduce_to_host(int nb, int nq, const double* partials, double* host_out) {
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(BLOCK), 0, stream, nb, nq, partials,
                       d_sums.p, (PowState*)nullptr, POW_NONE);
    d_sums.download(host_out, nq, stream);
    sync();
}

void Handle::ensure_partials(size_t count) {
    if (d_partials.n >= count && d_partials2.n >= count) return;
    sync();   // no kernel may still use the old buffers
    if (d_partials.n < count) d_partials.alloc(count);
    if (d_partials2.n < count) d_partials2.alloc(count);
}

// ---------------------------------------------------------------- preprocess (device)
// pdlp_preconditioner (preconditioning.jl:146-155) then precondition(milp, prec)
void Handle::scale() {
    const hipStream_t s = stream;
    const int32_t mm = (int32_t)m, nn = (int32_t)n;
    const int64_t nz1 = std::max<int64_t>(nnz, 1);
    d_D1o.alloc(std::max<int64_t>(m, 1));
    d_D2o.alloc(std::max<int64_t>(n, 1));
    d_a_csc.alloc(nz1);
    d_at_csc.alloc(nz1);
    DevBuf<double> d1(std::max<int64_t>(m, 1)), d2(std::max<int64_t>(n, 1));
    const int gm = grid_m(), gn = grid_n();
    hipLaunchKernelGGL(k_fill, dim3(gm), dim3(BLOCK), 0, s, (int64_t)m, d_D1o.p, 1.0);
    hipLaunchKernelGGL(k_fill, dim3(gn), dim3(BLOCK), 0, s, (int64_t)n, d_D2o.p, 1.0);
    if (nnz) {
        HIP_CHECK(hipMemcpyAsync(d_a_csc.p, d_nzval.p, nnz * 8, hipMemcpyDeviceToDevice, s));
        HIP_CHECK(hipMemcpyAsync(d_at_csc.p, d_nzval.p, nnz * 8, hipMemcpyDeviceToDevice, s));
    }
    for (int it = 0; it < P.ruiz_iter; ++it) {   // ruiz_preconditioner, :117-125
        hipLaunchKernelGGL(k_col_norm, dim3(gn), dim3(BLOCK), 0, s, nn, d_colptr.p, d_a_csc.p,
                           (double)INFINITY, d2.p);
        hipLaunchKernelGGL(k_row_norm, dim3(gm), dim3(BLOCK), 0, s, mm, d_csr_rowptr.p, d_perm.p,
                           d_at_csc.p, (double)INFINITY, d1.p);
        hipLaunchKernelGGL(k_apply_prec, dim3(gn), dim3(BLOCK), 0, s, nn, d_colptr.p, d_rowval.p,
                           d_a_csc.p, d_at_csc.p, d1.p, d2.p, d_a_csc.p, d_at_csc.p);
        hipLaunchKernelGGL(k_mul_into, dim3(gm), dim3(BLOCK), 0, s, mm, d_D1o.p, d1.p);
        hipLaunchKernelGGL(k_mul_into, dim3(gn), dim3(BLOCK), 0, s, nn, d_D2o.p, d2.p);
    }
    // cons_r = precondition(cons, prec_r) from the original values; CP(alpha)
    const double alpha = P.chambolle_pock_alpha;
    hipLaunchKernelGGL(k_apply_prec, dim3(gn), dim3(BLOCK), 0, s, nn, d_colptr.p, d_rowval.p,
                       d_nzval.p, d_nzval.p, d_D1o.p, d_D2o.p, d_a_csc.p, d_at_csc.p);
    hipLaunchKernelGGL(k_col_norm, dim3(gn), dim3(BLOCK), 0, s, nn, d_colptr.p, d_a_csc.p, alpha,
                       d2.p);
    hipLaunchKernelGGL(k_row_norm, dim3(gm), dim3(BLOCK), 0, s, mm, d_csr_rowptr.p, d_perm.p,
                       d_at_csc.p, 2.0 - alpha, d1.p);
    hipLaunchKernelGGL(k_mul_into, dim3(gm), dim3(BLOCK), 0, s, mm, d_D1o.p, d1.p);
    hipLaunchKernelGGL(k_mul_into, dim3(gn), dim3(BLOCK), 0, s, nn, d_D2o.p, d2.p);
    // A_p = D1*A*D2, At_p = D2*At*D1 from the originals
    hipLaunchKernelGGL(k_apply_prec, dim3(gn), dim3(BLOCK), 0, s, nn, d_colptr.p, d_rowval.p,
                       d_nzval.p, d_nzval.p, d_D1o.p, d_D2o.p, d_a_csc.p, d_at_csc.p);
    HIP_CHECK(hipGetLastError());
    const bool want32 = P.precision != CPDLP_PREC_FP64;
    A.fill_values(d_a_csc.p, want32, s);
    At.fill_values(d_at_csc.p, want32, s);
    // scaled vectors in internal order
    const int64_t n1 = std::max<int64_t>(n, 1), m1 = std::max<int64_t>(m, 1);
    c_s.alloc(n1); lv_s.alloc(n1); uv_s.alloc(n1); D2_s.alloc(n1); invD2_s.alloc(n1);
    lc_s.alloc(m1); uc_s.alloc(m1); D1_s.alloc(m1); invD1_s.alloc(m1);
    hipLaunchKernelGGL(k_scale_cols, dim3(gn), dim3(BLOCK), 0, s, nn, d_col_of_int.p, d_D2o.p,
                       d_c.p, d_lv.p, d_uv.p, c_s.p, lv_s.p, uv_s.p, D2_s.p, invD2_s.p);
    hipLaunchKernelGGL(k_scale_rows, dim3(gm), dim3(BLOCK), 0, s, mm, d_row_of_int.p, d_D1o.p,
                       d_lc.p, d_uc.p, lc_s.p, uc_s.p, D1_s.p, invD1_s.p);
    if (want32) {
        c_32.alloc(n1); lv_32.alloc(n1); uv_32.alloc(n1); lc_32.alloc(m1); uc_32.alloc(m1);
        hipLaunchKernelGGL((k_convert<double, float>), dim3(gn), dim3(BLOCK), 0, s, (int64_t)n, c_s.p, c_32.p);
        hipLaunchKernelGGL((k_convert<double, float>), dim3(gn), dim3(BLOCK), 0, s, (int64_t)n, lv_s.p, lv_32.p);
        hipLaunchKernelGGL((k_convert<double, float>), dim3(gn), dim3(BLOCK), 0, s, (int64_t)n, uv_s.p, uv_32.p);
        hipLaunchKernelGGL((k_convert<double, float>), dim3(gm), dim3(BLOCK), 0, s, (int64_t)m, lc_s.p, lc_32.p);
        hipLaunchKernelGGL((k_convert<double, float>), dim3(gm), dim3(BLOCK), 0, s, (int64_t)m, uc_s.p, uc_32.p);
    }
    HIP_CHECK(hipGetLastError());
    // constant scales of the KKT errors + primal_weight_init norms
    const int nb = 64;
    ensure_partials((size_t)nb * 4);
    hipLaunchKernelGGL(k_scale_sums, dim3(nb), dim3(BLOCK), 0, s, mm, nn, lc_s.p, uc_s.p,
                       invD1_s.p, c_s.p, invD2_s.p, d_partials.p);
    double sums[4];
    reduce_to_host(nb, 4, d_partials.p, sums);
    primal_scale = 1.0 + std::sqrt(sums[0]);
    dual_scale = 1.0 + std::sqrt(sums[1]);
    comb_norm = std::sqrt(sums[2]);
    c_norm = std::sqrt(sums[3]);
}

// ---------------------------------------------------------------- step size (device)
// fixed_stepsize / spectral_norm via the powm! rule (step_size.jl:22-27, linalg.jl:86-96)
template <class V>
void Handle::power_method(double tol, int64_t maxiter) {
    const hipStream_t s = stream;
    const int32_t nn = (int32_t)n;
    DevBuf<V> x(std::max<int64_t>(n, 1)), Ax(std::max<int64_t>(n, 1)), t(std::max<int64_t>(m, 1));
    DevBuf<PowState> st(1);
    PowState h0{};
    h0.theta = 0.0;
    h0.residual = 1.7976931348623157e308;
    h0.tol = tol;
    h0.iter = 0;
    h0.maxiter = maxiter;
    h0.done = (0 > maxiter) || (h0.residual <= tol * 0.0);
    st.upload(&h0, 1, s);
    const int gn = grid_n();
    const int nbr = std::min(gn, 1024);
    const int gA = A.view<V>().partial_blocks(), gAt = At.view<V>().partial_blocks();
    ensure_partials((size_t)std::max({gA, gAt, nbr}) * 2 + 16);
    hipLaunchKernelGGL(k_powm_start<V>, dim3(gn), dim3(BLOCK), 0, s, nn, d_col_of_int.p,
                       P.powm_seed, x.p);
    hipLaunchKernelGGL(k_sumsq<V>, dim3(nbr), dim3(BLOCK), 0, s, nn, (const V*)x.p, d_partials.p);
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(BLOCK), 0, s, nbr, 1, d_partials.p,
                       d_sums.p, (PowState*)nullptr, POW_NONE);
    hipLaunchKernelGGL(k_scale_by_inv_sqrt<V>, dim3(gn), dim3(BLOCK), 0, s, nn, x.p,
                       (const double*)d_sums.p);
    const MatView<V> vA = A.view<V>(), vAt = At.view<V>();
    PowState hs{};
    // 32 iterations per host check; every kernel reads the device PowState (a converged
    // state makes the rest of the chunk a no-op), so the chunk is captured once in a
    // hipGraph and replayed (no memset nodes)
    hipGraphExec_t chunk = nullptr;
    auto launch_chunk = [&]() {
        for (int c = 0; c < 32; ++c) {
            EpiStore<V> e1;
            e1.out = t.p;
            e1.st = st.p;
            launch_spmv<V, V, 1>(vA, XPack<V, 1>{{x.p}}, e1, s);
            EpiPowAt<V> e2;
            e2.partials = d_partials.p;
            e2.Ax = Ax.p;
            e2.x = x.p;
            e2.st = st.p;
            launch_spmv<V, V, 1>(vAt, XPack<V, 1>{{t.p}}, e2, s);
            hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(BLOCK), 0, s, gAt, 2, d_partials.p,
                               d_sums.p, st.p, POW_FIN1);
            hipLaunchKernelGGL(k_pow_update<V>, dim3(nbr), dim3(BLOCK), 0, s, nn, x.p,
                               (const V*)Ax.p, (const PowState*)st.p, d_partials2.p);
            hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(BLOCK), 0, s, nbr, 1,
                               d_partials2.p, d_sums.p + 8, st.p, POW_FIN2);
        }
    };
    for (;;) {
        if (P.use_graph) {
            if (!chunk) {
                hipGraph_t graph;
                HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                launch_chunk();
                HIP_CHECK(hipStreamEndCapture(s, &graph));
                HIP_CHECK(hipGraphInstantiate(&chunk, graph, nullptr, nullptr, 0));
                HIP_CHECK(hipGraphDestroy(graph));
            }
            HIP_CHECK(hipGraphLaunch(chunk, s));
        } else {
            launch_chunk();
        }
        HIP_CHECK(hipGetLastError());
        st.download(&hs, 1, s);
        sync();
        if (hs.done) break;
    }
    if (chunk) (void)hipGraphExecDestroy(chunk);
    powm_iters = hs.iter;
    eta = P.invnorm_scaling * (1.0 / std::sqrt(hs.theta));
}

// Format autotuner: for each side with a feasible ELL (padding <= 1.5x nnz), time the
// plain product in both formats on the actual matrix and keep the faster.
void Handle::choose_formats() {
    for (DevSide* S : {&A, &At}) {
        S->format = CPDLP_FMT_CSR;
        if (P.format == CPDLP_FMT_CSR) continue;
        if (P.format == CPDLP_FMT_ELL) {
            if (S->ell_w > 0) S->format = CPDLP_FMT_ELL;
            continue;
        }
        if (P.format == CPDLP_FMT_XSLICE || P.format == CPDLP_FMT_XSLICE_ALL) {
            if (S->sl_ns > 0) S->format = P.format;
            continue;
        }
        // AUTO with CPDLP_FORMATS="<A>,<A'>" (csr / ell / xslice / xslice_all): those
        // formats where available, CSR otherwise -- for counter runs, where every
        // dispatch is serialised and timing would not show what a normal run picks
        if (const char* e = std::getenv("CPDLP_FORMATS")) {
            const std::string spec(e);
            const size_t comma = spec.find(',');
            const std::string want = S == &A ? spec.substr(0, comma)
                                             : (comma == std::string::npos ? "" : spec.substr(comma + 1));
            if (want == "ell" && S->ell_w > 0) S->format = CPDLP_FMT_ELL;
            if (want == "xslice" && S->sl_ns > 0 && S->n_slice > 0) S->format = CPDLP_FMT_XSLICE;
            if (want == "xslice_all" && S->slices_all()) S->format = CPDLP_FMT_XSLICE_ALL;
            continue;
        }
        // AUTO: time the plain product (K = 1, store epilogue) of every available
        // format; another format replaces the best so far only when it is >= 10% faster
        // (25% for the slices of every row)
        std::vector<int32_t> cand{CPDLP_FMT_CSR};
        if (S->ell_w > 0) cand.push_back(CPDLP_FMT_ELL);
        if (S->sl_ns > 0 && S->n_slice > 0) cand.push_back(CPDLP_FMT_XSLICE);
        if (S->slices_all()) cand.push_back(CPDLP_FMT_XSLICE_ALL);
        if (cand.size() == 1) continue;
        DevBuf<double> xin(std::max<int64_t>(S->cols, 1)), out(std::max<int64_t>(S->rows, 1));
        hipLaunchKernelGGL(k_fill, dim3(std::max(1, ceil_div(S->cols, BLOCK))), dim3(BLOCK), 0,
                           stream, (int64_t)S->cols, xin.p, 1.0);
        float best_ms = 0.f;
        int32_t best = CPDLP_FMT_CSR;
        for (int32_t f : cand) {
            S->format = f;
            const MatView<double> v = S->view<double>();
            EpiStore<double> e;
            e.out = out.p;
            e.st = nullptr;
            for (int rep = 0; rep < 3; ++rep)   // warm-up
                launch_spmv<double, double, 1>(v, XPack<double, 1>{{xin.p}}, e, stream);
            float ms = 0.f;
            HIP_CHECK(hipEventRecord(ev0, stream));
            for (int rep = 0; rep < 10; ++rep)
                launch_spmv<double, double, 1>(v, XPack<double, 1>{{xin.p}}, e, stream);
            HIP_CHECK(hipEventRecord(ev1, stream));
            HIP_CHECK(hipEventSynchronize(ev1));
            HIP_CHECK(hipEventElapsedTime(&ms, ev0, ev1));
            // slicing every row moves the whole epilogue into a second pass over all
            // rows, which the store-only timing here leaves out: it must win by 25%
            const float bar = f == CPDLP_FMT_XSLICE_ALL ? 0.75f : 0.9f;
            if (f == CPDLP_FMT_CSR || ms < bar * best_ms) {
                best_ms = ms;
                best = f;
            }
        }
        S->format = best;
    }
    // line-grouped short tiles (AUTO, no override): kept where they time >= 3% faster
    // with the chosen format
    for (DevSide* S : {&A, &At}) {
        if (!S->lg_built || lg_mode == 1 || P.format != CPDLP_FMT_AUTO || std::getenv("CPDLP_FORMATS")) continue;
        if (S->format != CPDLP_FMT_CSR) {
            S->lg_on = false;
            continue;
        }
        DevBuf<double> xin(std::max<int64_t>(S->cols, 1)), out(std::max<int64_t>(S->rows, 1));
        hipLaunchKernelGGL(k_fill, dim3(std::max(1, ceil_div(S->cols, BLOCK))), dim3(BLOCK), 0,
                           stream, (int64_t)S->cols, xin.p, 1.0);
        float ms_of[2] = {0.f, 0.f};
        for (int on = 1; on >= 0; --on) {
            S->lg_on = on != 0;
            const MatView<double> v = S->view<double>();
            EpiStore<double> e;
            e.out = out.p;
            e.st = nullptr;
            for (int rep = 0; rep < 3; ++rep)
                launch_spmv<double, double, 1>(v, XPack<double, 1>{{xin.p}}, e, stream);
            HIP_CHECK(hipEventRecord(ev0, stream));
            for (int rep = 0; rep < 10; ++rep)
                launch_spmv<double, double, 1>(v, XPack<double, 1>{{xin.p}}, e, stream);
            HIP_CHECK(hipEventRecord(ev1, stream));
            HIP_CHECK(hipEventSynchronize(ev1));
            HIP_CHECK(hipEventElapsedTime(&ms_of[on], ev0, ev1));
        }
        S->lg_on = ms_of[1] < 0.97f * ms_of[0];
    }
}

void Handle::prepare() {
    HIP_CHECK(hipSetDevice(device));
    const double t0 = now_s();
    if (d_sums.n == 0) d_sums.alloc(64);
    scale();
    choose_formats();
    // power method in T: Float32 semantics for the all-FP32 precision.  Mixed precision
    // runs it on the FP32 matrix too when the tolerance is coarse enough (>= 1e-4
    // relative) for FP32 products to reach it, else in FP64.
    const bool t32 = P.precision == CPDLP_PREC_FP32;
    const double eps = t32 ? (double)FLT_EPSILON : DBL_EPSILON;
    const double tol = std::isnan(P.powm_tol) ? eps * (double)wrap_cube(n) : P.powm_tol;
    const int64_t maxiter = P.powm_maxiter < 0 ? n : P.powm_maxiter;
    if (t32 || (P.precision == CPDLP_PREC_MIXED && tol >= 1e-4))
        power_method<float>(tol, maxiter);
    else
        power_method<double>(tol, maxiter);
    // primal_weight_init (step_size.jl:29-40); PDHG: omega = 1 (pdhg.jl:48)
    omega0 = 1.0;
    if (P.algorithm == CPDLP_ALGO_PDLP && c_norm > P.zero_tol && comb_norm > P.zero_tol)
        omega0 = c_norm / comb_norm;
    prepared = true;
    setup_seconds = now_s() - t0;
}

// ---------------------------------------------------------------- KKT evaluation
template <class V, int K>
void Handle::kkt_launch(const V* const* X, const V* const* Y, const V* xr, const V* yr, bool dist,
                        double* host_sums) {
    const int gA = A.view<V>().partial_blocks(), gAt = At.view<V>().partial_blocks();
    constexpr int NQR = EpiKktRows<V, K>::NQ, NQC = EpiKktCols<V, K>::NQ;
    // (the 4-candidate rows side may run as two passes of 2: 2 * 6 partials per block)
    ensure_partials((size_t)std::max(std::max(gA * NQR, 2 * gA * EpiKktRows<V, 2>::NQ), gAt * NQC) + 16);
    kkt_launch_into<V, K>(X, Y, xr, yr, dist, host_sums, d_partials.p, d_partials2.p, d_sums.p);
}

template <class V, int K>
void Handle::kkt_launch_into(const V* const* X, const V* const* Y, const V* xr, const V* yr, bool dist,
                             double* host_sums, double* partials, double* partials2, double* sums_dev) {
    const hipStream_t s = stream;
    const MatView<V> vA = A.view<V>(), vAt = At.view<V>();
    const int gA = vA.partial_blocks(), gAt = vAt.partial_blocks();
    constexpr int NQR = EpiKktRows<V, K>::NQ, NQC = EpiKktCols<V, K>::NQ;
    // 4 candidates: both sides gather them interleaved (one request per entry instead of
    // 4), the same products summed in the same order, so the same bits
    bool aos = false;
    if constexpr (K == 4)
        aos = !no_aos && vA.format != CPDLP_FMT_ELL && vAt.format != CPDLP_FMT_ELL;
    // Else the rows side gathers the K candidates' x through the column slices: K * n / ns
    // elements per XCD.  With 4 candidates beyond an XCD's 4 MiB L2 it runs as two passes
    // of 2 (each vector's sums are the same numbers in either form).
    bool split = false;
    if constexpr (K == 4)
        split = !aos && vA.format == CPDLP_FMT_XSLICE && vA.ns > 0 &&
                (size_t)n * sizeof(V) * K / (size_t)vA.ns > ((size_t)4 << 20);
    V* ax = nullptr;
    V* ay = nullptr;
    if constexpr (K == 4) {
        if (aos) {
            const size_t need_x = ((size_t)std::max<int64_t>(n, 1) * 4 * sizeof(V) + 7) / 8;
            const size_t need_y = ((size_t)std::max<int64_t>(m, 1) * 4 * sizeof(V) + 7) / 8;
            if (aos_x.n < need_x || aos_y.n < need_y) {
                sync();   // a buffer in use by queued work is never freed under it
                if (aos_x.n < need_x) aos_x.alloc(need_x);
                if (aos_y.n < need_y) aos_y.alloc(need_y);
            }
            ax = (V*)aos_x.p;
            ay = (V*)aos_y.p;
            if (n)
                hipLaunchKernelGGL(k_interleave4<V>, dim3(ceil_div(n, BLOCK)), dim3(BLOCK), 0, s, (int64_t)n, X[0],
                                   X[1], X[2], X[3], ax);
            if (m)
                hipLaunchKernelGGL(k_interleave4<V>, dim3(ceil_div(m, BLOCK)), dim3(BLOCK), 0, s, (int64_t)m, Y[0],
                                   Y[1], Y[2], Y[3], ay);
        }
    }
    if (aos) {
        if constexpr (K == 4) {
            EpiKktRows<V, K> er;
            er.partials = partials;
            for (int k = 0; k < K; ++k) er.Y[k] = Y[k];
            er.y_rst = yr ? yr : Y[0];
            er.lc = lc_s.p;
            er.uc = uc_s.p;
            er.invD1 = invD1_s.p;
            er.dist = dist;
            launch_spmv<V, double, K>(vA, XPackAoS<V, K>{ax}, er, s);
        }
    } else if (split) {
        if constexpr (K == 4) {
            constexpr int NQ2 = EpiKktRows<V, 2>::NQ;
            for (int half = 0; half < 2; ++half) {
                EpiKktRows<V, 2> er;
                er.partials = partials + (size_t)half * gA * NQ2;
                er.Y[0] = Y[2 * half];
                er.Y[1] = Y[2 * half + 1];
                er.y_rst = yr ? yr : Y[0];
                er.lc = lc_s.p;
                er.uc = uc_s.p;
                er.invD1 = invD1_s.p;
                er.dist = half == 0 ? dist : false;
                launch_spmv<V, double, 2>(vA, XPack<V, 2>{{X[2 * half], X[2 * half + 1]}}, er, s);
            }
        }
    } else {
        EpiKktRows<V, K> er;
        er.partials = partials;
        for (int k = 0; k < K; ++k) er.Y[k] = Y[k];
        er.y_rst = yr ? yr : Y[0];
        er.lc = lc_s.p;
        er.uc = uc_s.p;
        er.invD1 = invD1_s.p;
        er.dist = dist;
        XPack<V, K> xp;
        for (int k = 0; k < K; ++k) xp.p[k] = X[k];
        launch_spmv<V, double, K>(vA, xp, er, s);
    }
    EpiKktCols<V, K> ec;
    ec.partials = partials2;
    for (int k = 0; k < K; ++k) ec.X[k] = X[k];
    ec.x_rst = xr ? xr : X[0];
    ec.c = c_s.p;
    ec.lv = lv_s.p;
    ec.uv = uv_s.p;
    ec.invD2 = invD2_s.p;
    ec.dist = dist;
    if (aos) {
        if constexpr (K == 4) launch_spmv<V, double, K>(vAt, XPackAoS<V, K>{ay}, ec, s);
    } else {
        XPack<V, K> yp;
        for (int k = 0; k < K; ++k) yp.p[k] = Y[k];
        launch_spmv<V, double, K>(vAt, yp, ec, s);
    }
    if (split) {   // two passes' sums, reassembled below into the K = 4 layout
        constexpr int NQ2 = EpiKktRows<V, 2>::NQ;
        for (int half = 0; half < 2; ++half)
            hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(BLOCK), 0, s, gA, NQ2,
                               partials + (size_t)half * gA * NQ2, sums_dev + NQR + NQC + half * NQ2,
                               (PowState*)nullptr, POW_NONE);
    } else if (gA > 0)
        hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(BLOCK), 0, s, gA, NQR, partials,
                           sums_dev, (PowState*)nullptr, POW_NONE);
    else
        hipLaunchKernelGGL(k_fill, dim3(1), dim3(BLOCK), 0, s, (int64_t)NQR, sums_dev, 0.0);
    if (gAt > 0)
        hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(BLOCK), 0, s, gAt, NQC, partials2,
                           sums_dev + NQR, (PowState*)nullptr, POW_NONE);
    else
        hipLaunchKernelGGL(k_fill, dim3(1), dim3(BLOCK), 0, s, (int64_t)NQC, sums_dev + NQR, 0.0);
    HIP_CHECK(hipGetLastError());
    if (!split) {
        if (NQR + NQC) HIP_CHECK(hipMemcpyAsync(host_sums, sums_dev, (NQR + NQC) * sizeof(double), hipMemcpyDeviceToHost, s));
        sync();
        return;
    }
    // rows layout of K = 4: the 2 sums of each candidate (first pass: 0, 1; second: 2, 3),
    // then the 2 distances (first pass)
    constexpr int NQ2 = EpiKktRows<V, 2>::NQ;
    double tmp[64];
    HIP_CHECK(hipMemcpyAsync(tmp, sums_dev, (NQR + NQC + 2 * NQ2) * sizeof(double), hipMemcpyDeviceToHost, s));
    sync();
    const double* a = tmp + NQR + NQC;
    const double* b = a + NQ2;
    for (int q = 0; q < 4; ++q) {
        host_sums[q] = a[q];
        host_sums[4 + q] = b[q];
    }
    host_sums[8] = a[4];
    host_sums[9] = a[5];
    for (int q = 0; q < NQC; ++q) host_sums[NQR + q] = tmp[NQR + q];
}

// errors.jl:77-98 from the raw sums of candidate k
cpdlp_kkt Handle::kkt_from_sums(const double* sums, int K, int k) const {
    const double* r = sums;
    const double* c = sums + 2 * K + 2;
    cpdlp_kkt e;
    e.primal = std::sqrt(r[2 * k]);
    e.primal_scale = primal_scale;
    e.dual = std::sqrt(c[3 * k]);
    e.dual_scale = dual_scale;
    const double pc = r[2 * k + 1], pv = c[3 * k + 1], cx = c[3 * k + 2];
    e.gap = std::fabs(cx + pc + pv);
    e.gap_scale = 1.0 + std::fabs(pc + pv) + std::fabs(cx);
    return e;
}

// Opt-in infeasibility test at a check (no counterpart in the reference): the three
// candidate rays of EpiRayRows / EpiRayCols, in order.  A dual ray e with a positive Farkas
// objective and sign violation at most infeasibility_tol times that objective certifies
// primal infeasibility; a primal ray d with c'd < 0 and cone violation at most
// infeasibility_tol * |c'd| certifies dual infeasibility.  Returns a status, or
// CPDLP_STILL_RUNNING when neither qualifies.
template <class V>
int32_t Handle::ray_check(IterState<V>& S) {
    const hipStream_t s = stream;
    const MatView<V> vA = A.view<V>(), vAt = At.view<V>();
    const int gA = vA.partial_blocks(), gAt = vAt.partial_blocks();
    constexpr int NQR = EpiRayRows<V>::NQ, NQC = EpiRayCols<V>::NQ;
    ensure_partials((size_t)std::max(gA * NQR, gAt * NQC) + 16);
    const int c = S.cur, l = 1 - S.cur;
    EpiRayRows<V> er;
    er.partials = d_partials.p;
    er.y = S.y[c].p;
    er.y_rst = S.yrst.p;
    er.y_last = S.y[l].p;
    er.lc = lc_s.p;
    er.uc = uc_s.p;
    er.invD1 = invD1_s.p;
    launch_spmv<V, double, 3>(vA, XPack<V, 3>{{S.x[c].p, S.xrst.p, S.x[l].p}}, er, s);
    EpiRayCols<V> ec;
    ec.partials = d_partials2.p;
    ec.x = S.x[c].p;
    ec.x_rst = S.xrst.p;
    ec.x_last = S.x[l].p;
    ec.c = c_s.p;
    ec.lv = lv_s.p;
    ec.uv = uv_s.p;
    ec.invD2 = invD2_s.p;
    launch_spmv<V, double, 3>(vAt, XPack<V, 3>{{S.y[c].p, S.yrst.p, S.y[l].p}}, ec, s);
    if (gA > 0)
        hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(BLOCK), 0, s, gA, NQR, d_partials.p,
                           d_sums.p, (PowState*)nullptr, POW_NONE);
    else
        hipLaunchKernelGGL(k_fill, dim3(1), dim3(BLOCK), 0, s, (int64_t)NQR, d_sums.p, 0.0);
    if (gAt > 0)
        hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(BLOCK), 0, s, gAt, NQC, d_partials2.p,
                           d_sums.p + NQR, (PowState*)nullptr, POW_NONE);
    else
        hipLaunchKernelGGL(k_fill, dim3(1), dim3(BLOCK), 0, s, (int64_t)NQC, d_sums.p + NQR, 0.0);
    HIP_CHECK(hipGetLastError());
    double q[NQR + NQC];
    d_sums.download(q, NQR + NQC, s);
    sync();
    const double* r = q;
    const double* cl = q + NQR;
    const double tol = P.infeasibility_tol;
    for (int k = 0; k < 3; ++k) {
        const double obj = r[3 + k] + cl[k];
        if (obj > 0.0 && std::sqrt(r[6 + k] + cl[3 + k]) <= tol * obj) return CPDLP_PRIMAL_INFEASIBLE;
        const double cd = cl[6 + k];
        if (cd < 0.0 && std::sqrt(r[k] + cl[9 + k]) <= tol * -cd) return CPDLP_DUAL_INFEASIBLE;
    }
    return CPDLP_STILL_RUNNING;
}

static double relative(const cpdlp_kkt& e) {   // errors.jl:55-58
    return std::max(std::max(e.primal / e.primal_scale, e.dual / e.dual_scale), e.gap / e.gap_scale);
}
static double absolute(const cpdlp_kkt& e, double w) {   // errors.jl:60-63
    return std::sqrt(w * w * e.primal * e.primal + (1.0 / (w * w)) * e.dual * e.dual + e.gap * e.gap);
}

// ---------------------------------------------------------------- the loop
template <class V>
void Handle::launch_steps(IterState<V>& S, int nsteps, bool pdlp) {
    const hipStream_t s = stream;
    const MatView<V> vA = A.view<V>(), vAt = At.view<V>();
    const bool f64 = std::is_same<V, double>::value;
    const V* c = f64 ? (const V*)c_s.p : (const V*)c_32.p;
    const V* lv = f64 ? (const V*)lv_s.p : (const V*)lv_32.p;
    const V* uv = f64 ? (const V*)uv_s.p : (const V*)uv_32.p;
    const V* lc = f64 ? (const V*)lc_s.p : (const V*)lc_32.p;
    const V* uc = f64 ? (const V*)uc_s.p : (const V*)uc_32.p;
    for (int k = 0; k < nsteps; ++k) {
        const int p = S.cur, q = 1 - S.cur;   // read sol_last = buffers p, write sol = q
        EpiPrimal<V> ep;
        ep.x = S.x[p].p;
        ep.c = c;
        ep.lv = lv;
        ep.uv = uv;
        ep.xn = S.x[q].p;
        ep.xbar = S.xbar.p;
        ep.xavg = S.xavg.p;
        ep.xavgl = S.xavgl.p;
        ep.coef = d_coef.p;
        ep.step = k;
        ep.pdlp = pdlp;
        launch_spmv<V, V, 1>(vAt, XPack<V, 1>{{S.y[p].p}}, ep, s);
        EpiDual<V> ed;
        ed.y = S.y[p].p;
        ed.lc = lc;
        ed.uc = uc;
        ed.yn = S.y[q].p;
        ed.yavg = S.yavg.p;
        ed.yavgl = S.yavgl.p;
        ed.coef = d_coef.p;
        ed.step = k;
        ed.pdlp = pdlp;
        launch_spmv<V, V, 1>(vA, XPack<V, 1>{{S.xbar.p}}, ed, s);
        S.cur = q;
    }
}

template <class V>
void Handle::run_chunk(IterState<V>& S, bool pdlp) {
    const int ce = P.check_every;
    const bool use_graph = P.use_graph != 0;
    if (!use_graph) {
        launch_steps(S, ce, pdlp);
        return;
    }
    const int parity = S.cur;
    const int gid = (std::is_same<V, double>::value ? 0 : 2) + parity;
    GraphCache& g = graphs[gid];
    if (!g.exec) {   // capture once per (precision, start parity); kernels only
        hipGraph_t graph;
        HIP_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        launch_steps(S, ce, pdlp);
        HIP_CHECK(hipStreamEndCapture(stream, &graph));
        HIP_CHECK(hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0));
        HIP_CHECK(hipGraphDestroy(graph));
        g.end_parity = S.cur;
        S.cur = parity;
    }
    HIP_CHECK(hipGraphLaunch(g.exec, stream));
    S.cur = g.end_parity;
}

// one PDLP/PDHG solve phase in precision V (pdlp.jl:69-94, pdhg.jl:56-74).
// Returns when the phase terminates; phase_reltol replaces termination_reltol (mixed
// precision's FP32 phase), final = false marks a non-final phase (no duplicate entry).
template <class V>
void Handle::loop(IterState<V>& S, double phase_reltol, bool final_phase, int64_t max_checks) {
    const bool pdlp = P.algorithm == CPDLP_ALGO_PDLP;
    const int ce = P.check_every;
    std::vector<StepCoef> coefs(ce);
    if (d_coef.n < (size_t)ce) {
        sync();
        d_coef.alloc(ce);
    }
    double sums[64];
    auto term_status = [&](const cpdlp_kkt& e) -> int32_t {   // termination.jl:89-101
        if (relative(e) <= phase_reltol) return CPDLP_OPTIMAL;
        if (stats.time_elapsed >= P.time_limit) return CPDLP_TIME_LIMIT;
        if (stats.kkt_passes >= P.max_kkt_passes) return CPDLP_ITERATION_LIMIT;
        return CPDLP_STILL_RUNNING;
    };
    auto push_history = [&]() {
        if (P.record_error_history) history.emplace_back(stats.kkt_passes, stats.err);
    };
    for (int64_t chk = 0; chk < max_checks; ++chk) {
        // step coefficients of this chunk: tau, sigma fixed; eta_sum advances per step
        const double tau = eta / omega, sigma = eta * omega, isig = 1.0 / sigma;
        for (int k = 0; k < ce; ++k) {
            StepCoef& cf = coefs[k];
            cf.tau = tau;
            cf.sigma = sigma;
            cf.isig = isig;
            cf.wa = eta / (eta + eta_sum);
            cf.wb = eta_sum / (eta + eta_sum);
            cf.last = (k == ce - 1);
            cf.pad = 0;
            if (pdlp) eta_sum += eta;
        }
        d_coef.upload(coefs.data(), ce, stream);
        HIP_CHECK(hipEventRecord(ev0, stream));
        run_chunk(S, pdlp);
        HIP_CHECK(hipEventRecord(ev1, stream));
        HIP_CHECK(hipGetLastError());
        stats.kkt_passes += ce;
        if (pdlp) {
            inner += ce;
            total += ce;
        }
        step_launches += 2 * ce;
        // check: KKT of sol (and, for PDLP, avg, sol_last, avg_last, restart)
        const int c = S.cur, l = 1 - S.cur;
        int K = 1;
        if (pdlp && rst_err_valid && !no_rst_cache) {   // the restart point's errors are known
            const V* X[4] = {S.x[c].p, S.xavg.p, S.x[l].p, S.xavgl.p};
            const V* Y[4] = {S.y[c].p, S.yavg.p, S.y[l].p, S.yavgl.p};
            kkt_launch<V, 4>(X, Y, S.xrst.p, S.yrst.p, true, sums);
            K = 4;
        } else if (pdlp) {
            const V* X[5] = {S.x[c].p, S.xavg.p, S.x[l].p, S.xavgl.p, S.xrst.p};
            const V* Y[5] = {S.y[c].p, S.yavg.p, S.y[l].p, S.yavgl.p, S.yrst.p};
            kkt_launch<V, 5>(X, Y, S.xrst.p, S.yrst.p, true, sums);
            K = 5;
        } else {
            const V* X[1] = {S.x[c].p};
            const V* Y[1] = {S.y[c].p};
            kkt_launch<V, 1>(X, Y, nullptr, nullptr, false, sums);
        }
        float ms = 0.f;
        HIP_CHECK(hipEventElapsedTime(&ms, ev0, ev1));
        step_ms += ms;
        // termination_check!, common.jl:218-231
        stats.time_elapsed = now_s() - t_start;
        stats.err = kkt_from_sums(sums, K, 0);
        push_history();
        stats.termination_status = term_status(stats.err);
        if (P.detect_infeasibility && stats.termination_status != CPDLP_OPTIMAL) {
            // opt-in: NaN / infinite errors, then the ray certificates
            const int32_t st = std::isfinite(relative(stats.err)) ? ray_check<V>(S) : CPDLP_NUMERICAL_ERROR;
            if (st != CPDLP_STILL_RUNNING) {
                stats.termination_status = st;
                if (P.show_progress) progress(true);
                return;
            }
        }
        if (P.show_progress) progress(false);
        bool switch_after = false;
        if (stats.termination_status == CPDLP_STILL_RUNNING && !final_phase) {
            // mixed precision, FP32 phase: hand over to FP64 when the relative KKT error
            // has not improved for STALL_CHECKS checks (after this check's restart
            // decision; the FP64 phase continues the trajectory), or at once when it has
            // grown 100x past its best (the FP64 phase then restarts from the best
            // iterate, kept here)
            const double rel = relative(stats.err);
            if (rel < best_rel) {
                best_rel = rel;
                best_omega = omega;
                stall = 0;
                if (n) HIP_CHECK(hipMemcpyAsync(S.tmp_n.p, S.x[c].p, n * sizeof(V), hipMemcpyDeviceToDevice, stream));
                if (m) HIP_CHECK(hipMemcpyAsync(S.tmp_m.p, S.y[c].p, m * sizeof(V), hipMemcpyDeviceToDevice, stream));
            } else if (rel > 100.0 * best_rel) {
                phase_switch = true;
                blowup = true;
                return;
            } else if (++stall >= STALL_CHECKS) {
                phase_switch = true;
                switch_after = true;
            }
        }
        if (stats.termination_status != CPDLP_STILL_RUNNING) {
            if (P.show_progress) progress(true);
            if (final_phase && pdlp) {   // second check after the inner loop, pdlp.jl:88-90
                stats.time_elapsed = now_s() - t_start;
                push_history();
                stats.termination_status = term_status(stats.err);
            }
            return;
        }
        if (!pdlp) {
            if (switch_after) return;
            continue;
        }
        // restart_check!, pdlp.jl:138-170 and should_restart, restart.jl:40-57
        const cpdlp_kkt e_s = stats.err, e_a = kkt_from_sums(sums, K, 1);
        const cpdlp_kkt e_l = kkt_from_sums(sums, K, 2), e_al = kkt_from_sums(sums, K, 3);
        // the restart point's errors: computed when K == 5, else those kept (the same
        // vector, evaluated by the same per-vector sums, so the same bits)
        const cpdlp_kkt e_r = K == 5 ? kkt_from_sums(sums, K, 4) : rst_err;
        rst_err = e_r;
        rst_err_valid = true;
        const bool from_avg = !(absolute(e_s, omega) < absolute(e_a, omega));
        const cpdlp_kkt cand = from_avg ? e_a : e_s;
        const cpdlp_kkt cand_last = absolute(e_l, omega) < absolute(e_al, omega) ? e_l : e_al;
        const double ac = absolute(cand, omega), ar = absolute(e_r, omega);
        const double acl = absolute(cand_last, omega);
        const bool sufficient = ac <= P.sufficient_decay * ar;
        const bool necessary = ac <= P.necessary_decay * ar;
        const bool no_local = ac > acl;
        const bool long_inner = (double)inner >= P.artificial_decay * (double)total;
        if (sufficient || (necessary && no_local) || long_inner) {
            // restart!, pdlp.jl:172-196; primal_weight_update!, step_size.jl:58-75
            eta_sum = 0.0;
            const double* rs = sums;
            const double* cs = sums + 2 * K + 2;
            const int di = from_avg ? 1 : 0;
            const double dx = std::sqrt(cs[3 * K + di]), dy = std::sqrt(rs[2 * K + di]);
            const double th = P.primal_weight_damping;
            if (dx > P.zero_tol && dy > P.zero_tol)
                omega = std::exp(th * std::log(dy / dx) + (1.0 - th) * std::log(omega));
            rst_err = cand;   // the restart point is now the candidate
            hipLaunchKernelGGL(k_restart<V>, dim3(grid_n()), dim3(BLOCK), 0, stream, (int64_t)n,
                               (int32_t)from_avg, S.x[c].p, S.xavg.p, S.xrst.p);
            hipLaunchKernelGGL(k_restart<V>, dim3(grid_m()), dim3(BLOCK), 0, stream, (int64_t)m,
                               (int32_t)from_avg, S.y[c].p, S.yavg.p, S.yrst.p);
            HIP_CHECK(hipGetLastError());
            inner = 0;
            restarts += 1;
        }
        if (switch_after) return;
    }
}

// show_progress: one line on stderr at most every 2 s and at termination (the
// reference's ProgressMeter line, src/algorithms/common.jl:134-145)
void Handle::progress(bool final_line) {
    const double t = now_s();
    if (!final_line && t - last_progress < 2.0) return;
    last_progress = t;
    const cpdlp_kkt& e = stats.err;
    std::fprintf(stderr, "%s iterations: %lld  primal %.3e  dual %.3e  gap %.3e  (%.1f s)%s\n",
                 P.algorithm == CPDLP_ALGO_PDLP ? "PDLP" : "PDHG", (long long)stats.kkt_passes,
                 e.primal / e.primal_scale, e.dual / e.dual_scale, e.gap / e.gap_scale,
                 stats.time_elapsed, final_line ? "  done" : "");
    std::fflush(stderr);
}

template <class V>
void Handle::init_state(IterState<V>& S, const double* x0_dev, const double* y0_dev) {
    const hipStream_t s = stream;
    rst_err_valid = false;
    S.alloc(n, m);
    const int gn = grid_n(), gm = grid_m();
    const int32_t nn = (int32_t)n, mm = (int32_t)m;
    // sol = precondition(sol_init): x/D2, y/D1 (internal order)
    hipLaunchKernelGGL(k_to_internal<V>, dim3(gn), dim3(BLOCK), 0, s, nn, d_col_of_int.p, x0_dev,
                       d_D2o.p, S.x[0].p);
    hipLaunchKernelGGL(k_to_internal<V>, dim3(gm), dim3(BLOCK), 0, s, mm, d_row_of_int.p, y0_dev,
                       d_D1o.p, S.y[0].p);
    // sol_last = 0, avg = sol, avg_last = 0, restart = sol  (pdlp.jl:51-54)
    hipLaunchKernelGGL(k_fill_t<V>, dim3(gn), dim3(BLOCK), 0, s, (int64_t)n, S.x[1].p, V(0));
    hipLaunchKernelGGL(k_fill_t<V>, dim3(gm), dim3(BLOCK), 0, s, (int64_t)m, S.y[1].p, V(0));
    hipLaunchKernelGGL(k_fill_t<V>, dim3(gn), dim3(BLOCK), 0, s, (int64_t)n, S.xavgl.p, V(0));
    hipLaunchKernelGGL(k_fill_t<V>, dim3(gm), dim3(BLOCK), 0, s, (int64_t)m, S.yavgl.p, V(0));
    if (n) {
        HIP_CHECK(hipMemcpyAsync(S.xavg.p, S.x[0].p, n * sizeof(V), hipMemcpyDeviceToDevice, s));
        HIP_CHECK(hipMemcpyAsync(S.xrst.p, S.x[0].p, n * sizeof(V), hipMemcpyDeviceToDevice, s));
    }
    if (m) {
        HIP_CHECK(hipMemcpyAsync(S.yavg.p, S.y[0].p, m * sizeof(V), hipMemcpyDeviceToDevice, s));
        HIP_CHECK(hipMemcpyAsync(S.yrst.p, S.y[0].p, m * sizeof(V), hipMemcpyDeviceToDevice, s));
    }
    S.cur = 0;
    HIP_CHECK(hipGetLastError());
}

// FP32 phase -> FP64 polish: the FP64 state restarts (sol = avg = restart) from the
// FP32 phase's best iterate (kept in tmp_n / tmp_m)
static void upcast(IterState<float>& a, IterState<double>& b, int64_t n, int64_t m, hipStream_t s) {
    const int gn = std::max(1, ceil_div(n, BLOCK)), gm = std::max(1, ceil_div(m, BLOCK));
    for (DevBuf<double>* t : {&b.x[0], &b.xavg, &b.xrst})
        hipLaunchKernelGGL((k_convert<float, double>), dim3(gn), dim3(BLOCK), 0, s, n, (const float*)a.tmp_n.p, t->p);
    for (DevBuf<double>* t : {&b.y[0], &b.yavg, &b.yrst})
        hipLaunchKernelGGL((k_convert<float, double>), dim3(gm), dim3(BLOCK), 0, s, m, (const float*)a.tmp_m.p, t->p);
    hipLaunchKernelGGL(k_fill_t<double>, dim3(gn), dim3(BLOCK), 0, s, n, b.x[1].p, 0.0);
    hipLaunchKernelGGL(k_fill_t<double>, dim3(gm), dim3(BLOCK), 0, s, m, b.y[1].p, 0.0);
    hipLaunchKernelGGL(k_fill_t<double>, dim3(gn), dim3(BLOCK), 0, s, n, b.xavgl.p, 0.0);
    hipLaunchKernelGGL(k_fill_t<double>, dim3(gm), dim3(BLOCK), 0, s, m, b.yavgl.p, 0.0);
    b.cur = 0;
}

// start of solve (common.jl:186-201) up to the loop: preprocess + the start point on the
// device + fresh solver scalars.  Returns true when the early exit of common.jl:194-198
// (no constraints and c == 0) has already produced the solution.
bool Handle::begin(const double* x0, const double* y0) {
    HIP_CHECK(hipSetDevice(device));
    t_start = now_s();
    last_progress = t_start;
    stats = cpdlp_stats{};
    history.clear();
    stats.err = cpdlp_kkt{NAN, NAN, NAN, NAN, NAN, NAN};
    stats.termination_status = CPDLP_STILL_RUNNING;
    stepping = false;
    prepare();   // solve = preprocess + initialize + solve!, every call (common.jl:186-201)
    // start point (unscaled, original order) on the device
    if (x0d.n != (size_t)std::max<int64_t>(n, 1)) x0d.alloc(std::max<int64_t>(n, 1));
    if (y0d.n != (size_t)std::max<int64_t>(m, 1)) y0d.alloc(std::max<int64_t>(m, 1));
    if (x0) x0d.upload(x0, n, stream);
    else hipLaunchKernelGGL(k_fill, dim3(grid_n()), dim3(BLOCK), 0, stream, (int64_t)n, x0d.p, 0.0);
    if (y0) y0d.upload(y0, m, stream);
    else hipLaunchKernelGGL(k_fill, dim3(grid_m()), dim3(BLOCK), 0, stream, (int64_t)m, y0d.p, 0.0);
    omega = omega0;
    eta_sum = 0.0;
    inner = total = 0;
    restarts = 0;
    step_ms = 0.0;
    step_launches = 0;
    fp32_passes = 0;
    for (auto& g : graphs) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        g.exec = nullptr;
    }
    t_loop = now_s();
    // early exit for no constraints and c == 0 (common.jl:194-198)
    bool c_zero = true;
    {
        std::vector<double> hc(n);
        if (n) {
            d_c.download(hc.data(), n, stream);
            sync();
        }
        for (double v : hc) if (v != 0.0) { c_zero = false; break; }
    }
    if (!(m == 0 && c_zero)) return false;
    init_state<double>(st64, x0d.p, y0d.p);
    std::vector<double> lvh(n), uvh(n);
    if (n) {
        lv_s.download(lvh.data(), n, stream);
        uv_s.download(uvh.data(), n, stream);
        sync();
        std::vector<double> xs(n);
        for (int64_t j = 0; j < n; ++j) xs[j] = proj_box(0.0, lvh[j], uvh[j]);
        st64.x[0].upload(xs.data(), n, stream);
        sync();
    }
    active64 = true;
    stats.termination_status = CPDLP_OPTIMAL;
    return true;
}

void Handle::finish_stats() {
    sync();
    loop_seconds = now_s() - t_loop;
    stats.eta = eta;
    stats.omega = omega;
    stats.powm_iters = powm_iters;
    stats.restarts = restarts;
    stats.n_history = (int64_t)history.size();
    stats.fp32_passes = fp32_passes;
    stats.setup_seconds = setup_seconds;
    stats.loop_seconds = loop_seconds;
    stats.step_kernel_ms = step_ms;
    stats.step_kernel_launches = step_launches;
    stats.format_A = A.format;
    stats.format_At = At.format;
    auto lg_used = [](const DevSide& S) { return S.lg_built && S.lg_on && S.format == CPDLP_FMT_CSR; };
    stats.line_tiles = (lg_used(A) ? 1 : 0) | (lg_used(At) ? 2 : 0);
    solved = true;
}

// FP32 phase -> FP64 phase by continuation: every iterate of the state (sol, sol_last,
// avg, avg_last, restart) converted, same buffer parity; omega, eta_sum and the restart
// counters carry over unchanged
static void continue_fp64(IterState<float>& a, IterState<double>& b, int64_t n, int64_t m, hipStream_t s) {
    const int gn = std::max(1, ceil_div(n, BLOCK)), gm = std::max(1, ceil_div(m, BLOCK));
    const std::pair<DevBuf<float>*, DevBuf<double>*> xs[] = {
        {&a.x[0], &b.x[0]}, {&a.x[1], &b.x[1]}, {&a.xavg, &b.xavg}, {&a.xavgl, &b.xavgl}, {&a.xrst, &b.xrst}};
    const std::pair<DevBuf<float>*, DevBuf<double>*> ys[] = {
        {&a.y[0], &b.y[0]}, {&a.y[1], &b.y[1]}, {&a.yavg, &b.yavg}, {&a.yavgl, &b.yavgl}, {&a.yrst, &b.yrst}};
    for (const auto& pr : xs)
        hipLaunchKernelGGL((k_convert<float, double>), dim3(gn), dim3(BLOCK), 0, s, n, (const float*)pr.first->p,
                           pr.second->p);
    for (const auto& pr : ys)
        hipLaunchKernelGGL((k_convert<float, double>), dim3(gm), dim3(BLOCK), 0, s, m, (const float*)pr.first->p,
                           pr.second->p);
    b.cur = a.cur;
}

void Handle::solve(const double* x0, const double* y0) {
    if (begin(x0, y0)) {
        finish_stats();
        return;
    }
    const bool f32 = P.precision != CPDLP_PREC_FP64;
    if (!f32) {
        init_state<double>(st64, x0d.p, y0d.p);
        active64 = true;
        loop<double>(st64, P.termination_reltol, true);
    } else {
        init_state<float>(st32, x0d.p, y0d.p);
        active64 = false;
        const bool polish = P.precision == CPDLP_PREC_MIXED;
        const double r32 = polish ? std::max(P.polish_reltol, P.termination_reltol) : P.termination_reltol;
        best_rel = INFINITY;
        best_omega = omega;
        stall = 0;
        phase_switch = false;
        blowup = false;
        loop<float>(st32, r32, !polish);
        fp32_passes = stats.kkt_passes;
        const bool reached = stats.termination_status == CPDLP_OPTIMAL && r32 > P.termination_reltol;
        if (polish && (reached || phase_switch)) {
            hand_over_fp64();
            loop<double>(st64, P.termination_reltol, true);
        }
    }
    finish_stats();
}

// initialize(milp, sol, algo) (common.jl:176; pdlp.jl:45-67, pdhg.jl:40-54): everything
// of solve before the loop; the loop then runs in pieces through run_checks.
void Handle::initialize(const double* x0, const double* y0) {
    if (begin(x0, y0)) {
        finish_stats();
        return;
    }
    if (P.precision == CPDLP_PREC_FP64) {
        init_state<double>(st64, x0d.p, y0d.p);
        active64 = true;
    } else {
        init_state<float>(st32, x0d.p, y0d.p);
        active64 = false;
        best_rel = INFINITY;
        best_omega = omega;
        stall = 0;
        phase_switch = false;
        blowup = false;
    }
    stepping = true;
    finish_stats();
}

// mixed precision: the FP32 phase has ended (its tolerance reached, or switched early);
// the rest of the solve runs in FP64 from its state, as solve does
void Handle::hand_over_fp64() {
    st64.alloc(n, m);
    rst_err_valid = false;   // the restart point in FP64 has FP64 errors
    if (blowup) {   // restart from the best FP32 iterate with its primal weight
        omega = best_omega;
        upcast(st32, st64, n, m, stream);
        eta_sum = 0.0;
        inner = 0;
        total = 0;
    } else {        // continue: the whole PDLP state in FP64, same scalars
        continue_fp64(st32, st64, n, m, stream);
    }
    active64 = true;
    stats.termination_status = CPDLP_STILL_RUNNING;
}

// the five iterates of a state (sol, sol_last, avg, avg_last, restart: PDLPState,
// pdlp.jl:19-43), scaled, in the original order: k_from_internal with D = 1 only permutes
template <class V>
void Handle::state_to(IterState<V>& S, double* xs, double* ys, const double* ones) {
    const int c = S.cur, l = 1 - S.cur;
    const V* X[CPDLP_STATE_VECTORS] = {S.x[c].p, S.x[l].p, S.xavg.p, S.xavgl.p, S.xrst.p};
    const V* Y[CPDLP_STATE_VECTORS] = {S.y[c].p, S.y[l].p, S.yavg.p, S.yavgl.p, S.yrst.p};
    for (int k = 0; k < CPDLP_STATE_VECTORS; ++k) {
        if (n)
            hipLaunchKernelGGL(k_from_internal<V>, dim3(grid_n()), dim3(BLOCK), 0, stream, (int32_t)n,
                               d_int_of_col.p, ones, X[k], xs + k * n);
        if (m)
            hipLaunchKernelGGL(k_from_internal<V>, dim3(grid_m()), dim3(BLOCK), 0, stream, (int32_t)m,
                               d_int_of_row.p, ones, Y[k], ys + k * m);
    }
    HIP_CHECK(hipGetLastError());
}

void Handle::get_state(cpdlp_state* st, double* x, double* y, double* D2, double* D1) {
    HIP_CHECK(hipSetDevice(device));
    if (!solved) throw Error(CPDLP_ERR_STATE, "no state: call cpdlp_initialize or cpdlp_solve first");
    if ((n && (!x || !D2)) || (m && (!y || !D1))) throw Error(CPDLP_ERR_ARGUMENT, "NULL state buffer");
    const int64_t nn = std::max<int64_t>(n, 1), mm = std::max<int64_t>(m, 1);
    DevBuf<double> ones(std::max(nn, mm)), xs(CPDLP_STATE_VECTORS * nn), ys(CPDLP_STATE_VECTORS * mm);
    hipLaunchKernelGGL(k_fill, dim3(std::max(1, ceil_div(std::max(nn, mm), BLOCK))), dim3(BLOCK), 0, stream,
                       std::max(nn, mm), ones.p, 1.0);
    if (active64) state_to(st64, xs.p, ys.p, ones.p);
    else state_to(st32, xs.p, ys.p, ones.p);
    if (n) {
        xs.download(x, CPDLP_STATE_VECTORS * n, stream);
        d_D2o.download(D2, n, stream);
    }
    if (m) {
        ys.download(y, CPDLP_STATE_VECTORS * m, stream);
        d_D1o.download(D1, m, stream);
    }
    sync();
    if (st) {
        st->eta = eta;
        st->omega = omega;
        st->eta_sum = eta_sum;
        st->kkt_passes = stats.kkt_passes;
        st->inner = inner;
        st->total = total;
        st->restarts = restarts;
        st->fp32_passes = active64 ? fp32_passes : stats.kkt_passes;
    }
}

// initialize, then the given state in place of the start point: its vectors relative to
// this handle's scaling (as given when D2 / D1 equal this handle's bit for bit, else
// through the unscaled point), its scalars and counters
void Handle::resume(const cpdlp_state& st, const double* x, const double* y, const double* D2,
                    const double* D1) {
    if (P.precision == CPDLP_PREC_MIXED)
        throw Error(CPDLP_ERR_ARGUMENT, "mixed precision: a loop state carries no FP32/FP64 phase; resume in fp64 or fp32");
    if ((n && (!x || !D2)) || (m && (!y || !D1))) throw Error(CPDLP_ERR_ARGUMENT, "NULL state buffer");
    if (!(std::isfinite(st.eta) && st.eta > 0) || !(std::isfinite(st.omega) && st.omega > 0) ||
        !(std::isfinite(st.eta_sum) && st.eta_sum >= 0))
        throw Error(CPDLP_ERR_ARGUMENT, "state: eta and omega must be positive, eta_sum non-negative");
    if (st.kkt_passes < 0 || st.inner < 0 || st.total < st.inner || st.restarts < 0 || st.fp32_passes < 0)
        throw Error(CPDLP_ERR_ARGUMENT, "state: inconsistent counters");
    if (begin(nullptr, nullptr)) {   // no constraints and c == 0: solved at once
        finish_stats();
        return;
    }
    std::vector<double> own2(n), own1(m);
    if (n) d_D2o.download(own2.data(), n, stream);
    if (m) d_D1o.download(own1.data(), m, stream);
    sync();
    const bool same2 = n == 0 || std::memcmp(own2.data(), D2, (size_t)n * 8) == 0;
    const bool same1 = m == 0 || std::memcmp(own1.data(), D1, (size_t)m * 8) == 0;
    const int64_t nn = std::max<int64_t>(n, 1), mm = std::max<int64_t>(m, 1);
    std::vector<double> hx, hy;
    if (!same2) {
        hx.resize((size_t)CPDLP_STATE_VECTORS * n);
        for (int k = 0; k < CPDLP_STATE_VECTORS; ++k)
            for (int64_t j = 0; j < n; ++j) hx[k * n + j] = x[k * n + j] * D2[j] / own2[j];
    }
    if (!same1) {
        hy.resize((size_t)CPDLP_STATE_VECTORS * m);
        for (int k = 0; k < CPDLP_STATE_VECTORS; ++k)
            for (int64_t i = 0; i < m; ++i) hy[k * m + i] = y[k * m + i] * D1[i] / own1[i];
    }
    DevBuf<double> ones(std::max(nn, mm)), xs(CPDLP_STATE_VECTORS * nn), ys(CPDLP_STATE_VECTORS * mm);
    hipLaunchKernelGGL(k_fill, dim3(std::max(1, ceil_div(std::max(nn, mm), BLOCK))), dim3(BLOCK), 0, stream,
                       std::max(nn, mm), ones.p, 1.0);
    if (n) xs.upload(same2 ? x : hx.data(), CPDLP_STATE_VECTORS * n, stream);
    if (m) ys.upload(same1 ? y : hy.data(), CPDLP_STATE_VECTORS * m, stream);
    auto put = [&](auto& S) {
        using V = std::remove_reference_t<decltype(*S.x[0].p)>;
        S.alloc(n, m);
        V* X[CPDLP_STATE_VECTORS] = {S.x[0].p, S.x[1].p, S.xavg.p, S.xavgl.p, S.xrst.p};
        V* Y[CPDLP_STATE_VECTORS] = {S.y[0].p, S.y[1].p, S.yavg.p, S.yavgl.p, S.yrst.p};
        for (int k = 0; k < CPDLP_STATE_VECTORS; ++k) {   // v / 1.0: the same bits
            if (n)
                hipLaunchKernelGGL(k_to_internal<V>, dim3(grid_n()), dim3(BLOCK), 0, stream, (int32_t)n,
                                   d_col_of_int.p, xs.p + k * n, ones.p, X[k]);
            if (m)
                hipLaunchKernelGGL(k_to_internal<V>, dim3(grid_m()), dim3(BLOCK), 0, stream, (int32_t)m,
                                   d_row_of_int.p, ys.p + k * m, ones.p, Y[k]);
        }
        S.cur = 0;
        HIP_CHECK(hipGetLastError());
    };
    active64 = P.precision == CPDLP_PREC_FP64;
    if (active64) put(st64);
    else put(st32);
    sync();   // the staging buffers are freed on return
    eta = st.eta;
    omega = st.omega;
    eta_sum = st.eta_sum;
    inner = st.inner;
    total = st.total;
    restarts = st.restarts;
    fp32_passes = st.fp32_passes;
    stats.kkt_passes = st.kkt_passes;
    rst_err_valid = false;
    stepping = true;
    finish_stats();
}

// solve!'s loop (pdlp.jl:69-94, pdhg.jl:58-74) for at most max_checks chunks of
// check_every steps; the reference yields once per chunk (pdlp.jl:77).
int32_t Handle::run_checks(int64_t max_checks) {
    HIP_CHECK(hipSetDevice(device));
    if (!stepping) {
        if (solved) return stats.termination_status;
        throw Error(CPDLP_ERR_STATE, "call cpdlp_initialize first");
    }
    if (max_checks <= 0) return stats.termination_status;
    if (active64) {
        loop<double>(st64, P.termination_reltol, true, max_checks);
    } else if (P.precision != CPDLP_PREC_MIXED) {
        loop<float>(st32, P.termination_reltol, true, max_checks);
    } else {   // the FP32 phase, then (within the same max_checks) the FP64 one, as solve
        const double r32 = std::max(P.polish_reltol, P.termination_reltol);
        const int64_t p0 = stats.kkt_passes;
        loop<float>(st32, r32, false, max_checks);
        fp32_passes = stats.kkt_passes;
        const bool reached = stats.termination_status == CPDLP_OPTIMAL && r32 > P.termination_reltol;
        if (reached || phase_switch) {
            hand_over_fp64();
            const int64_t left = max_checks - (stats.kkt_passes - p0) / P.check_every;
            if (left > 0) loop<double>(st64, P.termination_reltol, true, left);
        }
    }
    if (stats.termination_status != CPDLP_STILL_RUNNING) stepping = false;
    finish_stats();
    return stats.termination_status;
}

// unscaled solution in the original order into device buffers (NULL skips one)
void Handle::solution_to(double* xo, double* yo) {
    if (!solved) throw Error(CPDLP_ERR_STATE, "no solution: call cpdlp_solve first");
    const int32_t nn = (int32_t)n, mm = (int32_t)m;
    if (active64) {
        if (xo && n)
            hipLaunchKernelGGL(k_from_internal<double>, dim3(grid_n()), dim3(BLOCK), 0, stream, nn,
                               d_int_of_col.p, d_D2o.p, (const double*)st64.x[st64.cur].p, xo);
        if (yo && m)
            hipLaunchKernelGGL(k_from_internal<double>, dim3(grid_m()), dim3(BLOCK), 0, stream, mm,
                               d_int_of_row.p, d_D1o.p, (const double*)st64.y[st64.cur].p, yo);
    } else {
        if (xo && n)
            hipLaunchKernelGGL(k_from_internal<float>, dim3(grid_n()), dim3(BLOCK), 0, stream, nn,
                               d_int_of_col.p, d_D2o.p, (const float*)st32.x[st32.cur].p, xo);
        if (yo && m)
            hipLaunchKernelGGL(k_from_internal<float>, dim3(grid_m()), dim3(BLOCK), 0, stream, mm,
                               d_int_of_row.p, d_D1o.p, (const float*)st32.y[st32.cur].p, yo);
    }
    HIP_CHECK(hipGetLastError());
}

void Handle::get_solution(double* x, double* y) {
    HIP_CHECK(hipSetDevice(device));
    DevBuf<double> xo(std::max<int64_t>(n, 1)), yo(std::max<int64_t>(m, 1));
    solution_to(xo.p, yo.p);
    if (x) xo.download(x, n, stream);
    if (y) yo.download(y, m, stream);
    sync();
}

// the reference keeps sol on the device (common.jl:200): same, into caller buffers
void Handle::get_solution_device(double* x_dev, double* y_dev) {
    HIP_CHECK(hipSetDevice(device));
    solution_to(x_dev, y_dev);
    sync();
}

// Device time per launch of the two step kernels, each timed alone with HIP events
// on the handle's stream, on a scratch copy of the current iterates (the solve's
// state is not touched).  Uses the coefficients of the last chunk.
template <class V>
void Handle::profile_steps_t(IterState<V>& S, int reps, double* ms_primal, double* ms_dual) {
    IterState<V> T;
    T.alloc(n, m);
    const int c = S.cur;
    if (n) {
        HIP_CHECK(hipMemcpyAsync(T.x[0].p, S.x[c].p, n * sizeof(V), hipMemcpyDeviceToDevice, stream));
        HIP_CHECK(hipMemcpyAsync(T.xavg.p, S.xavg.p, n * sizeof(V), hipMemcpyDeviceToDevice, stream));
    }
    if (m) {
        HIP_CHECK(hipMemcpyAsync(T.y[0].p, S.y[c].p, m * sizeof(V), hipMemcpyDeviceToDevice, stream));
        HIP_CHECK(hipMemcpyAsync(T.yavg.p, S.yavg.p, m * sizeof(V), hipMemcpyDeviceToDevice, stream));
    }
    if (d_coef.n == 0) throw Error(CPDLP_ERR_STATE, "no step coefficients: call cpdlp_solve first");
    const bool pdlp = P.algorithm == CPDLP_ALGO_PDLP;
    const MatView<V> vA = A.view<V>(), vAt = At.view<V>();
    const bool f64 = std::is_same<V, double>::value;
    EpiPrimal<V> ep;
    ep.x = T.x[0].p;
    ep.c = f64 ? (const V*)c_s.p : (const V*)c_32.p;
    ep.lv = f64 ? (const V*)lv_s.p : (const V*)lv_32.p;
    ep.uv = f64 ? (const V*)uv_s.p : (const V*)uv_32.p;
    ep.xn = T.x[1].p;
    ep.xbar = T.xbar.p;
    ep.xavg = T.xavg.p;
    ep.xavgl = T.xavgl.p;
    ep.coef = d_coef.p;
    ep.step = 0;
    ep.pdlp = pdlp;
    EpiDual<V> ed;
    ed.y = T.y[0].p;
    ed.lc = f64 ? (const V*)lc_s.p : (const V*)lc_32.p;
    ed.uc = f64 ? (const V*)uc_s.p : (const V*)uc_32.p;
    ed.yn = T.y[1].p;
    ed.yavg = T.yavg.p;
    ed.yavgl = T.yavgl.p;
    ed.coef = d_coef.p;
    ed.step = 0;
    ed.pdlp = pdlp;
    launch_spmv<V, V, 1>(vAt, XPack<V, 1>{{T.y[0].p}}, ep, stream);   // warm-up
    launch_spmv<V, V, 1>(vA, XPack<V, 1>{{T.xbar.p}}, ed, stream);
    float ms = 0.f;
    HIP_CHECK(hipEventRecord(ev0, stream));
    for (int r = 0; r < reps; ++r) launch_spmv<V, V, 1>(vAt, XPack<V, 1>{{T.y[0].p}}, ep, stream);
    HIP_CHECK(hipEventRecord(ev1, stream));
    HIP_CHECK(hipEventSynchronize(ev1));
    HIP_CHECK(hipEventElapsedTime(&ms, ev0, ev1));
    *ms_primal = ms / reps;
    HIP_CHECK(hipEventRecord(ev0, stream));
    for (int r = 0; r < reps; ++r) launch_spmv<V, V, 1>(vA, XPack<V, 1>{{T.xbar.p}}, ed, stream);
    HIP_CHECK(hipEventRecord(ev1, stream));
    HIP_CHECK(hipEventSynchronize(ev1));
    HIP_CHECK(hipEventElapsedTime(&ms, ev0, ev1));
    *ms_dual = ms / reps;
    HIP_CHECK(hipGetLastError());
}

void Handle::profile_steps(int reps, double* ms_primal, double* ms_dual) {
    HIP_CHECK(hipSetDevice(device));
    if (!solved) throw Error(CPDLP_ERR_STATE, "call cpdlp_solve first");
    if (reps <= 0) throw Error(CPDLP_ERR_ARGUMENT, "reps must be positive");
    if (active64) profile_steps_t<double>(st64, reps, ms_primal, ms_dual);
    else profile_steps_t<float>(st32, reps, ms_primal, ms_dual);
}

cpdlp_kkt Handle::kkt_host(const double* x, const double* y) {
    HIP_CHECK(hipSetDevice(device));
    if (!prepared) prepare();
    DevBuf<double> xo(std::max<int64_t>(n, 1)), yo(std::max<int64_t>(m, 1));
    DevBuf<double> xi(std::max<int64_t>(n, 1)), yi(std::max<int64_t>(m, 1));
    xo.upload(x, n, stream);
    yo.upload(y, m, stream);
    hipLaunchKernelGGL(k_to_internal<double>, dim3(grid_n()), dim3(BLOCK), 0, stream, (int32_t)n,
                       d_col_of_int.p, xo.p, d_D2o.p, xi.p);
    hipLaunchKernelGGL(k_to_internal<double>, dim3(grid_m()), dim3(BLOCK), 0, stream, (int32_t)m,
                       d_row_of_int.p, yo.p, d_D1o.p, yi.p);
    const double* X[1] = {xi.p};
    const double* Y[1] = {yi.p};
    double sums[16];
    kkt_launch<double, 1>(X, Y, nullptr, nullptr, false, sums);
    return kkt_from_sums(sums, 1, 0);
}

// layout of kkt_device's scratch, in doubles, each part 32-double (256-byte) aligned:
// x (n, internal order, scaled), y (m), partials of A's rows, partials of A''s rows, sums
static size_t round32(size_t k) { return (k + 31) & ~(size_t)31; }

size_t Handle::kkt_scratch_doubles() const {
    const int gA = A.view<double>().partial_blocks(), gAt = At.view<double>().partial_blocks();
    constexpr int NQR = EpiKktRows<double, 1>::NQ, NQC = EpiKktCols<double, 1>::NQ;
    return round32((size_t)n) + round32((size_t)m) + round32((size_t)gA * NQR + 16) +
           round32((size_t)gAt * NQC + 16) + round32(NQR + NQC);
}

// kkt_errors! (errors.jl:65-109) of an unscaled device point, as kkt_host, with every
// intermediate in the caller-given device scratch: no device memory is allocated.
// Requires the preprocessing (cpdlp_prepare or a solve), which allocates.
cpdlp_kkt Handle::kkt_device(const double* x_dev, const double* y_dev, double* scratch, size_t scratch_len) {
    HIP_CHECK(hipSetDevice(device));
    // the preprocessing also fixes the formats, on which the partials' size depends
    if (!prepared) throw Error(CPDLP_ERR_STATE, "call cpdlp_prepare (or a solve) first: preprocessing allocates");
    const size_t need = kkt_scratch_doubles();
    if (need > scratch_len)
        throw Error(CPDLP_ERR_ARGUMENT, "LP needs " + std::to_string(need) + " doubles of scratch; the "
                                        "allocation-free KKT evaluation has " + std::to_string(scratch_len));
    const int gA = A.view<double>().partial_blocks(), gAt = At.view<double>().partial_blocks();
    constexpr int NQR = EpiKktRows<double, 1>::NQ, NQC = EpiKktCols<double, 1>::NQ;
    double* xi = scratch;
    double* yi = xi + round32((size_t)n);
    double* pa = yi + round32((size_t)m);
    double* pat = pa + round32((size_t)gA * NQR + 16);
    double* sums_dev = pat + round32((size_t)gAt * NQC + 16);
    hipLaunchKernelGGL(k_to_internal<double>, dim3(grid_n()), dim3(BLOCK), 0, stream, (int32_t)n,
                       d_col_of_int.p, x_dev, d_D2o.p, xi);
    hipLaunchKernelGGL(k_to_internal<double>, dim3(grid_m()), dim3(BLOCK), 0, stream, (int32_t)m,
                       d_row_of_int.p, y_dev, d_D1o.p, yi);
    const double* X[1] = {xi};
    const double* Y[1] = {yi};
    double sums[16];
    kkt_launch_into<double, 1>(X, Y, nullptr, nullptr, false, sums, pa, pat, sums_dev);
    return kkt_from_sums(sums, 1, 0);
}

}  // namespace cpdlp

// ================================================================ C ABI
using namespace cpdlp;

struct cpdlp_handle {
    std::unique_ptr<Handle> h;
};

extern "C" {

int cpdlp_default_params(cpdlp_params* p, int32_t algorithm) {
    return guarded([&] {
        if (!p) throw Error(CPDLP_ERR_ARGUMENT, "NULL params");
        if (algorithm != CPDLP_ALGO_PDLP && algorithm != CPDLP_ALGO_PDHG)
            throw Error(CPDLP_ERR_ARGUMENT, "unknown algorithm");
        std::memset(p, 0, sizeof(*p));
        const bool pdlp = algorithm == CPDLP_ALGO_PDLP;
        p->algorithm = algorithm;
        p->ruiz_iter = pdlp ? 10 : 0;
        p->chambolle_pock_alpha = 1.0;
        p->invnorm_scaling = 0.9;
        p->primal_weight_damping = pdlp ? 0.5 : NAN;
        p->zero_tol = 1e-8;
        p->sufficient_decay = pdlp ? 0.2 : NAN;
        p->necessary_decay = pdlp ? 0.8 : NAN;
        p->artificial_decay = pdlp ? 0.36 : NAN;
        p->show_progress = 0;
        p->check_every = 100;
        p->record_error_history = 1;
        p->precision = CPDLP_PREC_FP64;
        p->termination_reltol = 1e-4;
        p->max_kkt_passes = 100000;
        p->time_limit = 100.0;
        p->format = CPDLP_FMT_AUTO;
        p->device = 0;
        p->powm_tol = NAN;
        p->powm_maxiter = -1;
        p->powm_seed = 0;
        p->polish_reltol = 1e-5;
        p->use_graph = 1;
        p->long_row = 64;
        p->batch_streams = 4;
        p->detect_infeasibility = 0;
        p->infeasibility_tol = 1e-8;
    });
}

int cpdlp_create(const cpdlp_lp* lp, const cpdlp_params* params, cpdlp_handle** out) {
    return guarded([&] {
        if (!lp || !out) throw Error(CPDLP_ERR_ARGUMENT, "NULL argument");
        cpdlp_params p;
        if (params) p = *params;
        else cpdlp_default_params(&p, CPDLP_ALGO_PDLP);
        auto* h = new cpdlp_handle;
        try {
            h->h = std::make_unique<Handle>(*lp, p);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

#define HANDLE_OR_THROW(hp)                                                      \
    if (!(hp) || !(hp)->h) throw Error(CPDLP_ERR_ARGUMENT, "NULL handle");

int cpdlp_solve(cpdlp_handle* h, const double* x0, const double* y0) {
    return guarded([&] { HANDLE_OR_THROW(h); h->h->solve(x0, y0); });
}
int cpdlp_initialize(cpdlp_handle* h, const double* x0, const double* y0) {
    return guarded([&] { HANDLE_OR_THROW(h); h->h->initialize(x0, y0); });
}
int cpdlp_run_checks(cpdlp_handle* h, int64_t max_checks, int32_t* status) {
    return guarded([&] {
        HANDLE_OR_THROW(h);
        const int32_t st = h->h->run_checks(max_checks);
        if (status) *status = st;
    });
}
int cpdlp_get_state(cpdlp_handle* h, cpdlp_state* st, double* x, double* y, double* D2, double* D1) {
    return guarded([&] { HANDLE_OR_THROW(h); h->h->get_state(st, x, y, D2, D1); });
}
int cpdlp_resume(cpdlp_handle* h, const cpdlp_state* st, const double* x, const double* y, const double* D2,
                 const double* D1) {
    return guarded([&] {
        HANDLE_OR_THROW(h);
        if (!st) throw Error(CPDLP_ERR_ARGUMENT, "NULL state");
        h->h->resume(*st, x, y, D2, D1);
    });
}
int cpdlp_get_solution_device(cpdlp_handle* h, double* x_dev, double* y_dev) {
    return guarded([&] { HANDLE_OR_THROW(h); h->h->get_solution_device(x_dev, y_dev); });
}
int cpdlp_prepare(cpdlp_handle* h) {
    return guarded([&] { HANDLE_OR_THROW(h); HIP_CHECK(hipSetDevice(h->h->device)); h->h->prepare(); });
}
int cpdlp_get_solution(cpdlp_handle* h, double* x, double* y) {
    return guarded([&] { HANDLE_OR_THROW(h); h->h->get_solution(x, y); });
}
int cpdlp_get_stats(cpdlp_handle* h, cpdlp_stats* out) {
    return guarded([&] {
        HANDLE_OR_THROW(h);
        if (!out) throw Error(CPDLP_ERR_ARGUMENT, "NULL stats");
        *out = h->h->stats;
    });
}
int cpdlp_get_history(cpdlp_handle* h, int64_t* passes, cpdlp_kkt* errs, int64_t cap, int64_t* count) {
    return guarded([&] {
        HANDLE_OR_THROW(h);
        const auto& H = h->h->history;
        const int64_t k = std::min<int64_t>(cap, (int64_t)H.size());
        for (int64_t i = 0; i < k; ++i) {
            if (passes) passes[i] = H[i].first;
            if (errs) errs[i] = H[i].second;
        }
        if (count) *count = (int64_t)H.size();
    });
}
int cpdlp_kkt_errors(cpdlp_handle* h, const double* x, const double* y, cpdlp_kkt* out) {
    return guarded([&] {
        HANDLE_OR_THROW(h);
        if ((!x && h->h->n) || (!y && h->h->m) || !out) throw Error(CPDLP_ERR_ARGUMENT, "NULL argument");
        *out = h->h->kkt_host(x, y);
    });
}
// ---- allocation-free KKT evaluation (cpdlp_kkt_errors_device_noalloc) ----------------
// Its intermediates live in this file-scope device array: each GPU holds its own copy,
// 32 MiB, from the library's load until the process ends.  Host code never uses the
// array's name as an address: the per-device address comes from hipGetSymbolAddress,
// taken once per device with that device current.
constexpr size_t NOALLOC_SCRATCH_DOUBLES = (size_t)1 << 22;
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

static std::mutex scratch_mu;
static std::vector<double*> scratch_ptr;   // per device ordinal; nullptr until looked up

static double* scratch_on(int device) {   // the calling thread's current device is `device`
    std::lock_guard<std::mutex> g(scratch_mu);
    if ((int)scratch_ptr.size() <= device) scratch_ptr.resize((size_t)device + 1, nullptr);
    if (!scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (rc != hipSuccess || !p)
            throw Error(CPDLP_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(rc));
        scratch_ptr[device] = (double*)p;
    }
    return scratch_ptr[device];
}

int cpdlp_kkt_errors_device_noalloc(cpdlp_handle* h, const double* x_dev, const double* y_dev, cpdlp_kkt* out) {
    return guarded([&] {
        HANDLE_OR_THROW(h);
        Handle& H = *h->h;
        if ((!x_dev && H.n) || (!y_dev && H.m) || !out) throw Error(CPDLP_ERR_ARGUMENT, "NULL argument");
        HIP_CHECK(hipSetDevice(H.device));
        *out = H.kkt_device(x_dev, y_dev, scratch_on(H.device), NOALLOC_SCRATCH_DOUBLES);
    });
}

int cpdlp_get_scaling(cpdlp_handle* h, double* D1, double* D2, double* a_csc, double* at_csc) {
    return guarded([&] {
        HANDLE_OR_THROW(h);
        Handle& H = *h->h;
        HIP_CHECK(hipSetDevice(H.device));
        if (!H.prepared) H.prepare();
        if (D1) H.d_D1o.download(D1, H.m, H.stream);
        if (D2) H.d_D2o.download(D2, H.n, H.stream);
        if (a_csc) H.d_a_csc.download(a_csc, H.nnz, H.stream);
        if (at_csc) H.d_at_csc.download(at_csc, H.nnz, H.stream);
        H.sync();
    });
}
int cpdlp_get_csr(cpdlp_handle* h, int64_t* rowptr, int32_t* colval) {
    return guarded([&] {
        HANDLE_OR_THROW(h);
        const Handle& H = *h->h;
        if (rowptr)
            for (int64_t i = 0; i <= H.m; ++i) rowptr[i] = H.L.csr_rowptr[i];
        if (colval) std::memcpy(colval, H.L.csr_col.data(), H.nnz * sizeof(int32_t));
    });
}
int cpdlp_get_step_size(cpdlp_handle* h, double* eta, double* omega, int64_t* powm_iters) {
    return guarded([&] {
        HANDLE_OR_THROW(h);
        Handle& H = *h->h;
        if (!H.prepared) {
            HIP_CHECK(hipSetDevice(H.device));
            H.prepare();
        }
        if (eta) *eta = H.eta;
        if (omega) *omega = H.omega0;
        if (powm_iters) *powm_iters = H.powm_iters;
    });
}
int cpdlp_profile_steps(cpdlp_handle* h, int32_t reps, double* ms_primal, double* ms_dual) {
    return guarded([&] {
        HANDLE_OR_THROW(h);
        if (!ms_primal || !ms_dual) throw Error(CPDLP_ERR_ARGUMENT, "NULL output");
        h->h->profile_steps(reps, ms_primal, ms_dual);
    });
}

void cpdlp_destroy(cpdlp_handle* h) { delete h; }

const char* cpdlp_last_error(void) { return g_last_error.c_str(); }

const char* cpdlp_version(void) { return "coolpdlp_hip 0.1.0 (gfx950)"; }

}  // extern "C"
