// spmv.hpp -- sparse matrix-vector kernels with fused per-row epilogues.
//
// One kernel template serves every product of the PDHG loop: A*xbar and A'*y in the
// half-steps, the K-candidate products of the KKT check, and the power method.  The
// matrix is a CSR in an internal row order where the rows longer than `long_row`
// come first (a stable partition by length, the reference's density permutation
// idea, src/components/permutation.jl:1-46).  Two block kinds share one launch:
//
//  * long rows: one 64-lane wavefront per row, four rows per 256-thread block.  Lane l
//    accumulates k = l, l+64, ... in order, then a butterfly (__shfl_xor) reduces the
//    wavefront.  Deterministic; the order differs from a sequential sum.
//  * short tiles: a contiguous run of at most 256 short rows and at most TILE_NNZ
//    nonzeros.  The block loads the tile's (col, val) pairs coalesced into registers,
//    forms products val*x[col] into LDS, and each thread sums its own row's products
//    sequentially from zero in ascending k -- the same operations, in the same
//    order, as the reference's thread-per-row kernel (src/utils/mat_csr.jl:63-79),
//    so these rows are bit-identical to it.
//
// Epilogues (struct with NQ and row()) consume the K row sums; when NQ > 0 each block
// reduces NQ doubles deterministically into partials[blockIdx.x * NQ + q].
#pragma once
#include "common.hpp"

namespace cpdlp {

constexpr int TILE_NNZ = 2048;
constexpr int PER_T = TILE_NNZ / BLOCK;   // nonzeros held per thread in a short tile
constexpr int LONG_ROWS_PER_BLOCK = BLOCK / WAVE;

template <class V>
struct MatView {
    int32_t rows = 0;
    const int32_t* rowptr = nullptr;
    const int32_t* col = nullptr;
    const V* val = nullptr;
    int32_t n_long = 0;           // internal rows [0, n_long) are long
    int32_t n_long_blocks = 0;    // ceil(n_long / 4)
    const int32_t* tile_ptr = nullptr;   // short tiles: rows [tile_ptr[t], tile_ptr[t+1])
    int32_t n_tiles = 0;
    // column-major ELL (format 2): rows x width, col -1 = padding
    int32_t format = CPDLP_FMT_CSR;
    int32_t ell_width = 0;
    const int32_t* ell_col = nullptr;
    const V* ell_val = nullptr;

    int grid() const {
        return format == CPDLP_FMT_ELL ? ceil_div(rows, BLOCK) : n_long_blocks + n_tiles;
    }
};

template <class V, int K>
struct XPack {
    const V* p[K];
};

template <class V, class Acc, int K, class Epi>
__global__ __launch_bounds__(BLOCK) void k_spmv_csr(MatView<V> A, XPack<V, K> X, Epi epi) {
    __shared__ Acc prod[TILE_NNZ];
    double q[Epi::NQ > 0 ? Epi::NQ : 1];
#pragma unroll
    for (int i = 0; i < (Epi::NQ > 0 ? Epi::NQ : 1); ++i) q[i] = 0.0;
    if (epi.skip()) return;
    const int tid = threadIdx.x;
    if ((int)blockIdx.x < A.n_long_blocks) {
        const int lane = tid & (WAVE - 1);
        const int r = blockIdx.x * LONG_ROWS_PER_BLOCK + tid / WAVE;
        if (r < A.n_long) {
            Acc acc[K];
#pragma unroll
            for (int kv = 0; kv < K; ++kv) acc[kv] = Acc(0);
            const int k1 = A.rowptr[r + 1];
            for (int k = A.rowptr[r] + lane; k < k1; k += WAVE) {
                const int c = A.col[k];
                const Acc v = (Acc)A.val[k];
#pragma unroll
                for (int kv = 0; kv < K; ++kv) acc[kv] += v * (Acc)X.p[kv][c];
            }
#pragma unroll
            for (int kv = 0; kv < K; ++kv) acc[kv] = wave_sum(acc[kv]);
            if (lane == 0) epi.row(r, acc, q);
        }
    } else {
        const int t = blockIdx.x - A.n_long_blocks;
        const int r0 = A.tile_ptr[t], r1 = A.tile_ptr[t + 1];
        const int base = A.rowptr[r0];
        const int cnt = A.rowptr[r1] - base;
        const int r = r0 + tid;
        const bool has = r < r1;
        int lo = 0, hi = 0;
        if (has) {
            lo = A.rowptr[r] - base;
            hi = A.rowptr[r + 1] - base;
        }
        int cc[PER_T];
        V vv[PER_T];
#pragma unroll
        for (int j = 0; j < PER_T; ++j) {
            const int e = tid + j * BLOCK;
            cc[j] = 0;
            vv[j] = V(0);
            if (e < cnt) {
                cc[j] = A.col[base + e];
                vv[j] = A.val[base + e];
            }
        }
        Acc acc[K];
#pragma unroll
        for (int kv = 0; kv < K; ++kv) {
#pragma unroll
            for (int j = 0; j < PER_T; ++j) {
                const int e = tid + j * BLOCK;
                if (e < cnt) prod[e] = (Acc)vv[j] * (Acc)X.p[kv][cc[j]];
            }
            __syncthreads();
            Acc s = Acc(0);
            for (int e = lo; e < hi; ++e) s += prod[e];
            acc[kv] = s;
            __syncthreads();
        }
        if (has) epi.row(r, acc, q);
    }
    if constexpr (Epi::NQ > 0) block_sum<Epi::NQ>(q, epi.partials + (size_t)blockIdx.x * Epi::NQ);
}

// Column-major ELL: thread per row; entry s of rows i..i+63 is contiguous, so each
// load instruction is coalesced.  Sum over s ascending, skipping padding, as the
// reference's spmv_ell! (src/utils/mat_ell.jl:72-89) -- bit-identical to it.
template <class V, class Acc, int K, class Epi>
__global__ __launch_bounds__(BLOCK) void k_spmv_ell(MatView<V> A, XPack<V, K> X, Epi epi) {
    double q[Epi::NQ > 0 ? Epi::NQ : 1];
#pragma unroll
    for (int i = 0; i < (Epi::NQ > 0 ? Epi::NQ : 1); ++i) q[i] = 0.0;
    if (epi.skip()) return;
    const int r = blockIdx.x * BLOCK + threadIdx.x;
    if (r < A.rows) {
        Acc acc[K];
#pragma unroll
        for (int kv = 0; kv < K; ++kv) acc[kv] = Acc(0);
        const size_t stride = (size_t)A.rows;
        for (int s = 0; s < A.ell_width; ++s) {
            const int c = A.ell_col[s * stride + r];
            if (c < 0) break;   // left-packed: padding only after the row's entries
            const Acc v = (Acc)A.ell_val[s * stride + r];
#pragma unroll
            for (int kv = 0; kv < K; ++kv) acc[kv] += v * (Acc)X.p[kv][c];
        }
        epi.row(r, acc, q);
    }
    if constexpr (Epi::NQ > 0) block_sum<Epi::NQ>(q, epi.partials + (size_t)blockIdx.x * Epi::NQ);
}

template <class V, class Acc, int K, class Epi>
inline void launch_spmv(const MatView<V>& A, const XPack<V, K>& X, const Epi& epi, hipStream_t s) {
    const int g = A.grid();
    if (g <= 0) return;
    if (A.format == CPDLP_FMT_ELL)
        hipLaunchKernelGGL((k_spmv_ell<V, Acc, K, Epi>), dim3(g), dim3(BLOCK), 0, s, A, X, epi);
    else
        hipLaunchKernelGGL((k_spmv_csr<V, Acc, K, Epi>), dim3(g), dim3(BLOCK), 0, s, A, X, epi);
}

}  // namespace cpdlp
