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
// Epilogues consume the K row sums: load(r) returns the row's epilogue operands and is
// issued before the gathers (so its latency overlaps them), row(r, acc, q, pre)
// finishes the row; when NQ > 0 each block reduces NQ doubles deterministically into
// partials[blockIdx.x * NQ + q].
#pragma once
#include "common.hpp"

namespace cpdlp {

constexpr int TILE_NNZ = 2048;
constexpr int PER_T = TILE_NNZ / BLOCK;   // nonzeros held per thread in a short tile
constexpr int LONG_ROWS_PER_BLOCK = BLOCK / WAVE;

// one short tile: rows [r0, r1), nonzeros [base, base + cnt); read with one 16-byte
// scalar load at block start
struct alignas(16) TileMeta {
    int32_t r0, r1, base, cnt;
};

template <class V>
struct MatView {
    int32_t rows = 0;
    const int32_t* rowptr = nullptr;
    const int32_t* col = nullptr;
    const V* val = nullptr;
    int32_t n_long = 0;           // internal rows [0, n_long) are long
    int32_t n_long_blocks = 0;    // ceil(n_long / 4)
    const TileMeta* tiles = nullptr;     // short tiles
    int32_t n_tiles = 0;
    // column-major ELL (format 2): rows x width, col -1 = padding
    int32_t format = CPDLP_FMT_CSR;
    int32_t ell_width = 0;
    const int32_t* ell_col = nullptr;
    const V* ell_val = nullptr;
    // XCD column slices (format 3): NS sub-CSRs over all rows, one per column slice;
    // slice s owns the blocks with blockIdx % NS == s (workgroups are dispatched to
    // the XCDs round-robin, so each XCD's L2 holds one slice of the gathered vector)
    int32_t ns = 0;
    const int32_t* sl_rowptr = nullptr;   // ns * (rows + 1), absolute offsets
    const int32_t* sl_col = nullptr;
    const V* sl_val = nullptr;
    const TileMeta* sl_tiles = nullptr;   // tiles of slice s: sl_tiles[sl_tile_off[s] ...]
    const int32_t* sl_tile_off = nullptr; // ns + 1; slice s has off[s+1] - off[s] tiles
    int32_t sl_blocks = 0;                // blocks per slice (long blocks + max tiles)
    void* sl_part = nullptr;              // K * ns * rows partial sums (Acc)

    int grid() const {
        if (format == CPDLP_FMT_ELL) return ceil_div(rows, BLOCK);
        if (format == CPDLP_FMT_XSLICE) return ns * sl_blocks;
        return n_long_blocks + n_tiles;
    }
};

template <class V, int K>
struct XPack {
    const V* p[K];
};

// Long row r (one wavefront): lane l takes k = l, l+64, ...; U entries are loaded
// (clamped, so every load is unconditional and all are in flight together) before any
// is accumulated; accumulation stays in ascending k per lane; then a butterfly.
template <class V, class Acc, int K>
__device__ __forceinline__ void long_row_sums(const int32_t* rowptr, const int32_t* col,
                                              const V* val, int r, const XPack<V, K>& X,
                                              Acc (&acc)[K]) {
    const int lane = threadIdx.x & (WAVE - 1);
#pragma unroll
    for (int kv = 0; kv < K; ++kv) acc[kv] = Acc(0);
    const int k0 = rowptr[r], k1 = rowptr[r + 1];
    constexpr int U = 4;
    for (int kb = k0 + lane; kb < k1; kb += U * WAVE) {
        int c[U];
        Acc v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = min(kb + u * WAVE, k1 - 1);
            c[u] = col[k];
            v[u] = (Acc)val[k];
        }
#pragma unroll
        for (int kv = 0; kv < K; ++kv) {
            Acc xv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) xv[u] = (Acc)X.p[kv][c[u]];
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (kb + u * WAVE < k1) acc[kv] += v[u] * xv[u];
        }
    }
#pragma unroll
    for (int kv = 0; kv < K; ++kv) acc[kv] = wave_sum(acc[kv]);
}

// Short tile rows [r0, r1): thread tid owns row r0 + tid.  Every load unconditional
// (index clamped to the tile), so all PER_T (col, val) pairs and then all PER_T
// gathers are in flight together; products go to LDS and each thread sums its row
// sequentially from zero in ascending k.  Slots past the tile's nnz are written but
// never summed.  All threads of the block must call this (it synchronises).
template <class V, class Acc, int K>
__device__ __forceinline__ void tile_sums(const int32_t* rowptr, const int32_t* col,
                                          const V* val, const TileMeta& tm, const XPack<V, K>& X,
                                          Acc* prod, Acc (&acc)[K]) {
    const int tid = threadIdx.x;
    const int base = tm.base, cnt = tm.cnt;
    const int r = tm.r0 + tid;
    int lo = 0, hi = 0;
    if (r < tm.r1) {
        lo = rowptr[r] - base;
        hi = rowptr[r + 1] - base;
    }
#pragma unroll
    for (int kv = 0; kv < K; ++kv) acc[kv] = Acc(0);
    if (cnt <= 0) return;   // uniform per block
    int cc[PER_T];
    V vv[PER_T];
#pragma unroll
    for (int j = 0; j < PER_T; ++j) {
        const int e = min(tid + j * BLOCK, cnt - 1);
        cc[j] = col[base + e];
        vv[j] = val[base + e];
    }
#pragma unroll
    for (int kv = 0; kv < K; ++kv) {
        Acc xv[PER_T];
#pragma unroll
        for (int j = 0; j < PER_T; ++j) xv[j] = (Acc)X.p[kv][cc[j]];
#pragma unroll
        for (int j = 0; j < PER_T; ++j) prod[tid + j * BLOCK] = (Acc)vv[j] * xv[j];
        __syncthreads();
        Acc sum = Acc(0);
        for (int e = lo; e < hi; ++e) sum += prod[e];
        acc[kv] = sum;
        __syncthreads();
    }
}

template <class V, class Acc, int K, class Epi>
__global__ __launch_bounds__(BLOCK) void k_spmv_csr(MatView<V> A, XPack<V, K> X, Epi epi) {
    __shared__ Acc prod[TILE_NNZ];
    double q[Epi::NQ > 0 ? Epi::NQ : 1];
#pragma unroll
    for (int i = 0; i < (Epi::NQ > 0 ? Epi::NQ : 1); ++i) q[i] = 0.0;
    if (epi.skip()) return;
    Acc acc[K];
    if ((int)blockIdx.x < A.n_long_blocks) {
        const int r = blockIdx.x * LONG_ROWS_PER_BLOCK + threadIdx.x / WAVE;
        if (r < A.n_long) {
            typename Epi::Pre pre{};
            if ((threadIdx.x & (WAVE - 1)) == 0) pre = epi.load(r);
            long_row_sums<V, Acc, K>(A.rowptr, A.col, A.val, r, X, acc);
            if ((threadIdx.x & (WAVE - 1)) == 0) epi.row(r, acc, q, pre);
        }
    } else {
        const TileMeta tm = A.tiles[blockIdx.x - A.n_long_blocks];
        const int r = tm.r0 + (int)threadIdx.x;
        typename Epi::Pre pre{};
        if (r < tm.r1) pre = epi.load(r);   // issued before the gathers
        tile_sums<V, Acc, K>(A.rowptr, A.col, A.val, tm, X, prod, acc);
        if (r < tm.r1) epi.row(r, acc, q, pre);
    }
    if constexpr (Epi::NQ > 0) block_sum<Epi::NQ>(q, epi.partials + (size_t)blockIdx.x * Epi::NQ);
}

// XCD-sliced product, pass 1: block b works on slice s = b % ns and writes the
// slice's partial row sums part[(kv * ns + s) * rows + r].
template <class V, class Acc, int K>
__global__ __launch_bounds__(BLOCK) void k_spmv_slices(MatView<V> A, XPack<V, K> X) {
    __shared__ Acc prod[TILE_NNZ];
    const int s = blockIdx.x % A.ns;
    const int j = blockIdx.x / A.ns;
    const int32_t* rp = A.sl_rowptr + (size_t)s * (A.rows + 1);
    Acc* part = (Acc*)A.sl_part;
    Acc acc[K];
    if (j < A.n_long_blocks) {
        const int r = j * LONG_ROWS_PER_BLOCK + threadIdx.x / WAVE;
        if (r < A.n_long) {
            long_row_sums<V, Acc, K>(rp, A.sl_col, A.sl_val, r, X, acc);
            if ((threadIdx.x & (WAVE - 1)) == 0)
#pragma unroll
                for (int kv = 0; kv < K; ++kv) part[((size_t)kv * A.ns + s) * A.rows + r] = acc[kv];
        }
    } else {
        const int t = A.sl_tile_off[s] + (j - A.n_long_blocks);
        if (t >= A.sl_tile_off[s + 1]) return;   // uniform per block: slice has fewer tiles
        const TileMeta tm = A.sl_tiles[t];
        tile_sums<V, Acc, K>(rp, A.sl_col, A.sl_val, tm, X, prod, acc);
        const int r = tm.r0 + threadIdx.x;
        if (r < tm.r1)
#pragma unroll
            for (int kv = 0; kv < K; ++kv) part[((size_t)kv * A.ns + s) * A.rows + r] = acc[kv];
    }
}

// XCD-sliced product, pass 2: row sum = sum of the slice partials in slice order,
// then the epilogue (thread per row, coalesced).
template <class V, class Acc, int K, class Epi>
__global__ __launch_bounds__(BLOCK) void k_slices_combine(MatView<V> A, Epi epi) {
    double q[Epi::NQ > 0 ? Epi::NQ : 1];
#pragma unroll
    for (int i = 0; i < (Epi::NQ > 0 ? Epi::NQ : 1); ++i) q[i] = 0.0;
    if (epi.skip()) return;
    const int r = blockIdx.x * BLOCK + threadIdx.x;
    if (r < A.rows) {
        const typename Epi::Pre pre = epi.load(r);
        const Acc* part = (const Acc*)A.sl_part;
        Acc acc[K];
#pragma unroll
        for (int kv = 0; kv < K; ++kv) {
            Acc sum = part[((size_t)kv * A.ns) * A.rows + r];
            for (int s = 1; s < A.ns; ++s) sum += part[((size_t)kv * A.ns + s) * A.rows + r];
            acc[kv] = sum;
        }
        epi.row(r, acc, q, pre);
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
        const typename Epi::Pre pre = epi.load(r);
        Acc acc[K];
#pragma unroll
        for (int kv = 0; kv < K; ++kv) acc[kv] = Acc(0);
        const size_t stride = (size_t)A.rows;
        constexpr int U = 4;
        for (int s0 = 0; s0 < A.ell_width; s0 += U) {
            int c[U];
            Acc v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int sl = min(s0 + u, A.ell_width - 1);
                c[u] = A.ell_col[sl * stride + r];
                v[u] = (Acc)A.ell_val[sl * stride + r];
            }
#pragma unroll
            for (int kv = 0; kv < K; ++kv) {
                Acc xv[U];
#pragma unroll
                for (int u = 0; u < U; ++u) xv[u] = (Acc)X.p[kv][c[u] < 0 ? 0 : c[u]];
#pragma unroll
                for (int u = 0; u < U; ++u)
                    if (s0 + u < A.ell_width && c[u] >= 0) acc[kv] += v[u] * xv[u];
            }
            if (c[U - 1] < 0) break;   // left-packed: the rest of the row is padding
        }
        epi.row(r, acc, q, pre);
    }
    if constexpr (Epi::NQ > 0) block_sum<Epi::NQ>(q, epi.partials + (size_t)blockIdx.x * Epi::NQ);
}

template <class V, class Acc, int K, class Epi>
inline void launch_spmv(const MatView<V>& A, const XPack<V, K>& X, const Epi& epi, hipStream_t s) {
    const int g = A.grid();
    if (g <= 0) return;
    if (A.format == CPDLP_FMT_XSLICE) {
        hipLaunchKernelGGL((k_spmv_slices<V, Acc, K>), dim3(g), dim3(BLOCK), 0, s, A, X);
        hipLaunchKernelGGL((k_slices_combine<V, Acc, K, Epi>), dim3(ceil_div(A.rows, BLOCK)),
                           dim3(BLOCK), 0, s, A, epi);
    } else if (A.format == CPDLP_FMT_ELL)
        hipLaunchKernelGGL((k_spmv_ell<V, Acc, K, Epi>), dim3(g), dim3(BLOCK), 0, s, A, X, epi);
    else
        hipLaunchKernelGGL((k_spmv_csr<V, Acc, K, Epi>), dim3(g), dim3(BLOCK), 0, s, A, X, epi);
}

}  // namespace cpdlp
