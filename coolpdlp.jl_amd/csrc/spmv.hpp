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
// The matrix entries are read once per product: their loads carry the non-temporal hint,
// so that they do not displace the gathered vector's lines from L2 (3-4% faster on the
// primal step of a 2M-row, 20M-nonzero product than plain loads).
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
// line-grouped tiles: LG_NT threads, LG_PT entries each, at most LG_NT rows and
// LG_TILE_NNZ entries; an entry packs its slot in the tile (12 bits) above its column
// (20 bits).  Measured on the large power-law LP's primal step (threads x entries per
// thread): 256 x 8 110.1 us, 512 x 8 108.5 us, 1024 x 4 105.8 us (46 VGPRs).
constexpr int LG_NT = 1024;
constexpr int LG_PT = 4;
constexpr int LG_TILE_NNZ = LG_NT * LG_PT;
constexpr int LG_COL_BITS = 20;
constexpr uint32_t LG_COL_MASK = (1u << LG_COL_BITS) - 1u;
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
    // line-grouped tiles of the short rows (LG_NT rows and LG_TILE_NNZ entries at most,
    // entries column-sorted within each tile, packed slot | column, and their values), or
    // null; the single-vector products without block partials (the steps) use them
    const TileMeta* lg_tiles = nullptr;
    int32_t lg_n_tiles = 0;
    const uint32_t* lg_col = nullptr;
    const V* lg_val = nullptr;
    // column-major ELL (format 2): rows x width, col -1 = padding
    int32_t format = CPDLP_FMT_CSR;
    int32_t ell_width = 0;
    const int32_t* ell_col = nullptr;
    const V* ell_val = nullptr;
    // XCD column slices (format 3): the first sl_rows rows (the rows longer than the
    // slice threshold -- the long rows and the longest short ones --, or every row when
    // sl_all) are also stored cut by column slice, NS slices (one per XCD); the slice-s
    // part of every such row is summed by a block with blockIdx % NS == s (workgroups are
    // dispatched to the XCDs round-robin), so each XCD gathers from one slice of the
    // vector, which stays in that XCD's L2.  The other rows run as the short tiles xs_tiles.
    int32_t ns = 0;
    int32_t sl_rows = 0;                  // rows [0, sl_rows) are sliced
    int32_t sl_n_long = 0;                // rows [0, sl_n_long): parts summed by 16-lane groups
    bool sl_all = false;                  // sl_rows == rows: no short-tile blocks
    int32_t sl_stride = 0;                // per-slice rowptr length (built rows + 1)
    const int32_t* sl_rowptr = nullptr;   // ns * sl_stride, absolute offsets
    const int32_t* sl_col = nullptr;
    const V* sl_val = nullptr;
    const TileMeta* sl_tiles = nullptr;   // tiles of slice s: sl_tiles[sl_tile_off[s] ...]
    const int32_t* sl_tile_off = nullptr; // ns + 1: first tile of each slice
    const int32_t* sl_tile_n = nullptr;   // ns: tiles of each slice covering [0, sl_rows)
    int32_t sl_max_tiles = 0;             // most of any slice
    void* sl_part = nullptr;              // ns * sl_rows slice sums (Acc) per vector
    const TileMeta* xs_tiles = nullptr;   // short tiles of rows [sl_rows, rows)
    int32_t n_xs_tiles = 0;

    // CSR short tiles: they start at a block index that is a multiple of 8 (blocks
    // [n_long_blocks, csr_short0()) are idle), and XCD x takes the contiguous tiles
    // [x * C, (x + 1) * C), C = ceil(n_tiles / 8), in order.  (The sliced format keeps
    // its short tiles interleaved over the XCDs: there each XCD's L2 holds its column
    // slice, and a contiguous range of short rows measured 26% slower on the dual step.)
    __host__ __device__ int csr_short0() const { return (n_long_blocks + 7) & ~7; }
    __host__ __device__ int csr_per_xcd() const { return (n_tiles + 7) / 8; }
    __host__ __device__ int short_tile_of(int b) const {   // b: block index from the first short block
        const int t = (b % 8) * csr_per_xcd() + b / 8;
        return t < n_tiles ? t : -1;
    }
    // the sliced format's blocks that run the side's own short tiles
    __host__ __device__ int sl_short_tiles() const { return sl_all ? 0 : n_xs_tiles; }
    int grid() const {
        if (format == CPDLP_FMT_ELL) return ceil_div(rows, BLOCK);
        if (format == CPDLP_FMT_XSLICE) return ns * sl_max_tiles + sl_short_tiles();
        return csr_short0() + 8 * csr_per_xcd();
    }
    // blocks that write an epilogue's block partials (NQ per block): the sliced format's
    // short tiles, then its combine blocks
    int partial_blocks() const {
        if (format == CPDLP_FMT_XSLICE) return sl_short_tiles() + ceil_div(sl_rows, BLOCK);
        return grid();
    }
};

constexpr int SL_KMAX = 5;   // most vectors per product in the sliced format
constexpr int SL_NS_MAX = 8; // most column slices (one per XCD)
constexpr int SL_RPT = 8;    // short rows per thread in a sliced tile

template <class V, int K>
struct XPack {   // K separate vectors
    static constexpr bool aos = false;
    const V* p[K];
};

// K vectors interleaved: vector kv's entry c at p[c * K + kv] (K a power of two, p aligned
// to K * sizeof(V)), so one wide load brings an entry's K values: one L2 request per
// gather instead of K (the K-candidate check)
template <class V, int K>
struct XPackAoS {
    static constexpr bool aos = true;
    const V* p;
};
template <class V, int K>
struct alignas(K * sizeof(V)) VecK {
    V v[K];
};
template <class V, int K>
__device__ __forceinline__ VecK<V, K> load_k(const XPackAoS<V, K>& X, int c) {
    return *(const VecK<V, K>*)(X.p + (size_t)c * K);
}

// Long row r (one wavefront): lane l takes k = l, l+64, ...; U entries are loaded
// (clamped, so every load is unconditional and all are in flight together) before any
// is accumulated; accumulation stays in ascending k per lane; then a butterfly.
template <class V, class Acc, int K, class XP>
__device__ __forceinline__ void long_row_sums(const int32_t* rowptr, const int32_t* col,
                                              const V* val, int r, const XP& X,
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
            c[u] = __builtin_nontemporal_load(col + k);
            v[u] = (Acc)__builtin_nontemporal_load(val + k);
        }
        if constexpr (XP::aos) {
            VecK<V, K> xk[U];
#pragma unroll
            for (int u = 0; u < U; ++u) xk[u] = load_k(X, c[u]);
#pragma unroll
            for (int kv = 0; kv < K; ++kv)
#pragma unroll
                for (int u = 0; u < U; ++u)
                    if (kb + u * WAVE < k1) acc[kv] += v[u] * (Acc)xk[u].v[kv];
        } else {
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
    }
#pragma unroll
    for (int kv = 0; kv < K; ++kv) acc[kv] = wave_sum(acc[kv]);
}

// Short tile rows [r0, r1): thread tid owns row r0 + tid.  Every load unconditional
// (index clamped to the tile), so all PT (col, val) pairs and then all PT
// gathers are in flight together; products go to LDS and each thread sums its row
// sequentially from zero in ascending k.  Slots past the tile's nnz are written but
// never summed.  NT threads, NT * PT entries at most.  LG: the tile's line-grouped
// copy (lgc, lgv): entries in column order, each product written to its own slot, so the
// sums are the same.  (A separate instantiation: the row-ordered kernel keeps its
// registers and paired LDS stores.)
// All threads of the block must call this (it synchronises).
template <class V, class Acc, int K, class XP, bool LG, int NT, int PT = PER_T>
__device__ __forceinline__ void tile_sums(const int32_t* rowptr, const int32_t* col,
                                          const V* val, const uint32_t* lgc, const V* lgv,
                                          const TileMeta& tm, const XP& X,
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
    int cc[PT];
    int slot[PT];   // LDS slot of product j
    V vv[PT];
#pragma unroll
    for (int j = 0; j < PT; ++j) {
        const int i = tid + j * NT;
        const int e = min(i, cnt - 1);
        if constexpr (LG) {
            const uint32_t p = __builtin_nontemporal_load(lgc + base + e);
            cc[j] = (int)(p & LG_COL_MASK);
            slot[j] = i < cnt ? (int)(p >> LG_COL_BITS) : i;
            vv[j] = __builtin_nontemporal_load(lgv + base + e);
        } else {
            cc[j] = __builtin_nontemporal_load(col + base + e);
            slot[j] = i;
            vv[j] = __builtin_nontemporal_load(val + base + e);
        }
    }
    if constexpr (XP::aos) {   // an entry's K values in one load, then the K products in turn
        VecK<V, K> xk[PT];
#pragma unroll
        for (int j = 0; j < PT; ++j) xk[j] = load_k(X, cc[j]);
#pragma unroll
        for (int kv = 0; kv < K; ++kv) {
#pragma unroll
            for (int j = 0; j < PT; ++j) prod[slot[j]] = (Acc)vv[j] * (Acc)xk[j].v[kv];
            __syncthreads();
            acc[kv] = seq_sum(prod, lo, hi);
            __syncthreads();
        }
    } else {
#pragma unroll
        for (int kv = 0; kv < K; ++kv) {
            Acc xv[PT];
#pragma unroll
            for (int j = 0; j < PT; ++j) xv[j] = (Acc)X.p[kv][cc[j]];
#pragma unroll
            for (int j = 0; j < PT; ++j) prod[slot[j]] = (Acc)vv[j] * xv[j];
            __syncthreads();
            acc[kv] = seq_sum(prod, lo, hi);
            __syncthreads();
        }
    }
}

// (65 VGPRs, 7 waves per SIMD for the FP64 primal step.  Forcing 8 waves cut it to 49
// VGPRs but fewer loads in flight: 145 us instead of 130 us on the large LP.)
// NT threads per block: NT / 64 long rows or one tile of NT rows and NT * PT entries;
// LG: the view's tiles are the line-grouped ones.
template <class V, class Acc, int K, class Epi, class XP = XPack<V, K>, bool LG = false, int NT = BLOCK,
          int PT = PER_T>
__global__ __launch_bounds__(NT) void k_spmv_csr(MatView<V> A, XP X, Epi epi) {
    __shared__ Acc prod[NT * PT];
    double q[Epi::NQ > 0 ? Epi::NQ : 1];
#pragma unroll
    for (int i = 0; i < (Epi::NQ > 0 ? Epi::NQ : 1); ++i) q[i] = 0.0;
    if (epi.skip()) return;
    Acc acc[K];
    if ((int)blockIdx.x < A.n_long_blocks) {
        const int r = blockIdx.x * (NT / WAVE) + threadIdx.x / WAVE;
        if (r < A.n_long) {
            typename Epi::Pre pre{};
            if ((threadIdx.x & (WAVE - 1)) == 0) pre = epi.load(r);
            long_row_sums<V, Acc, K, XP>(A.rowptr, A.col, A.val, r, X, acc);
            if ((threadIdx.x & (WAVE - 1)) == 0) epi.row(r, acc, q, pre);
        }
    } else if ((int)blockIdx.x >= A.csr_short0()) {
        // XCD blockIdx % 8 walks its own contiguous range of tiles, so entries first
        // touched by neighbouring tiles share lines in that XCD's L2
        const int t = A.short_tile_of((int)blockIdx.x - A.csr_short0());
        TileMeta tm{0, 0, 0, 0};
        if (t >= 0) tm = A.tiles[t];
        const int r = tm.r0 + (int)threadIdx.x;
        typename Epi::Pre pre{};
        if (r < tm.r1) pre = epi.load(r);   // issued before the gathers
        tile_sums<V, Acc, K, XP, LG, NT, PT>(A.rowptr, A.col, A.val, A.lg_col, A.lg_val, tm, X, prod, acc);
        if (r < tm.r1) epi.row(r, acc, q, pre);
    }
    if constexpr (Epi::NQ > 0) block_sum<Epi::NQ, NT>(q, epi.partials + (size_t)blockIdx.x * Epi::NQ);
}

// Sum of a value over the SEG lanes of an aligned lane group (butterfly; fixed order).
template <int SEG, class Acc>
__device__ __forceinline__ Acc group_sum(Acc v) {
#pragma unroll
    for (int o = SEG / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, SEG);
    return v;
}

// XCD-sliced product, pass 1.  Blocks [0, ns * sl_max_tiles): block b sums slice
// s = b % ns of the rows of tile b / ns of that slice, for each of the K vectors,
// into sl_part[(kv * ns + s) * sl_rows + r]: the tile's entries are loaded coalesced
// once and the products of vector kv formed in LDS.  A long row's part is summed by a
// group of 16 lanes (lane l takes entries l, l+16, ... in order, then a butterfly); a
// short row's part (a few entries) by one thread, sequentially; a row part longer than
// TILE_NNZ by the whole block (strided, then the four wave sums in wave order).  Blocks
// past those (when the short rows are not sliced): the side's short tiles, summed and
// finished by the epilogue exactly as in k_spmv_csr, block partials at index
// b - ns * sl_max_tiles.  Deterministic; the sliced rows' order of summation differs
// from a sequential sum (as the wavefront path's does).
template <class V, class Acc, int K, class Epi, class XP = XPack<V, K>>
__global__ __launch_bounds__(BLOCK) void k_spmv_xslice(MatView<V> A, XP X, Epi epi) {
    __shared__ Acc prod[TILE_NNZ];
    __shared__ Acc red[K][BLOCK / WAVE];
    if (epi.skip()) return;
    const int nsb = A.ns * A.sl_max_tiles;
    const int tid = threadIdx.x;
    if ((int)blockIdx.x >= nsb) {
        double q[Epi::NQ > 0 ? Epi::NQ : 1];
#pragma unroll
        for (int i = 0; i < (Epi::NQ > 0 ? Epi::NQ : 1); ++i) q[i] = 0.0;
        const TileMeta tm = A.xs_tiles[blockIdx.x - nsb];
        const int r = tm.r0 + tid;
        typename Epi::Pre pre{};
        if (r < tm.r1) pre = epi.load(r);
        Acc acc[K];
        tile_sums<V, Acc, K, XP, false, BLOCK>(A.rowptr, A.col, A.val, nullptr, nullptr, tm, X, prod, acc);
        if (r < tm.r1) epi.row(r, acc, q, pre);
        if constexpr (Epi::NQ > 0)
            block_sum<Epi::NQ>(q, epi.partials + (size_t)(blockIdx.x - nsb) * Epi::NQ);
        return;
    }
    const int s = blockIdx.x % A.ns;
    const int tl = (int)blockIdx.x / A.ns;
    if (tl >= A.sl_tile_n[s]) return;   // uniform per block: this slice has fewer tiles
    const TileMeta tm = A.sl_tiles[A.sl_tile_off[s] + tl];
    const int32_t* rp = A.sl_rowptr + (size_t)s * A.sl_stride;
    Acc* part = (Acc*)A.sl_part;
    auto part_at = [&](int kv, int r) -> Acc& { return part[((size_t)kv * A.ns + s) * A.sl_rows + r]; };
    if (tm.cnt > TILE_NNZ) {   // one long row part, the whole block
        constexpr int U = 4;
        Acc acc[K];
#pragma unroll
        for (int kv = 0; kv < K; ++kv) acc[kv] = Acc(0);
        const int k1 = tm.base + tm.cnt;
        for (int kb = tm.base + tid; kb < k1; kb += U * BLOCK) {
            int c[U];
            Acc v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = min(kb + u * BLOCK, k1 - 1);
                c[u] = __builtin_nontemporal_load(A.sl_col + k);
                v[u] = (Acc)__builtin_nontemporal_load(A.sl_val + k);
            }
            if constexpr (XP::aos) {
                VecK<V, K> xk[U];
#pragma unroll
                for (int u = 0; u < U; ++u) xk[u] = load_k(X, c[u]);
#pragma unroll
                for (int kv = 0; kv < K; ++kv)
#pragma unroll
                    for (int u = 0; u < U; ++u)
                        if (kb + u * BLOCK < k1) acc[kv] += v[u] * (Acc)xk[u].v[kv];
            } else {
#pragma unroll
                for (int kv = 0; kv < K; ++kv) {
                    Acc xv[U];
#pragma unroll
                    for (int u = 0; u < U; ++u) xv[u] = (Acc)X.p[kv][c[u]];
#pragma unroll
                    for (int u = 0; u < U; ++u)
                        if (kb + u * BLOCK < k1) acc[kv] += v[u] * xv[u];
                }
            }
        }
#pragma unroll
        for (int kv = 0; kv < K; ++kv) {
            const Acc w = wave_sum(acc[kv]);
            if ((tid & (WAVE - 1)) == 0) red[kv][tid / WAVE] = w;
        }
        __syncthreads();
        if (tid < K) {
            Acc sum = red[tid][0];
#pragma unroll
            for (int w = 1; w < BLOCK / WAVE; ++w) sum += red[tid][w];
            part_at(tid, tm.r0) = sum;
        }
        return;
    }
    constexpr int SEG = 16;
    const int g = tid / SEG, l = tid % SEG;
    // short rows' tiles (uniform per block: tiles never span the long/short boundary):
    // thread tid sums the parts of rows r0 + tid + j * BLOCK, j < SL_RPT.  Their bounds
    // are read when summed, so that the long rows' launches do not hold registers for them.
    const bool per_thread = tm.r0 >= A.sl_n_long;
    // every lane of a group runs the same trips, so the butterfly's partners are active
    auto row_parts = [&](int kv) {
        if (per_thread) {
#pragma unroll
            for (int j = 0; j < SL_RPT; ++j) {
                const int rt = tm.r0 + tid + j * BLOCK;
                if (rt < tm.r1) part_at(kv, rt) = seq_sum(prod, rp[rt] - tm.base, rp[rt + 1] - tm.base);
            }
            return;
        }
        for (int r = tm.r0 + g; r < tm.r1; r += BLOCK / SEG) {
            const int lo = rp[r] - tm.base, hi = rp[r + 1] - tm.base;
            Acc sum = Acc(0);
            for (int e = lo + l; e < hi; e += SEG) sum += prod[e];
            sum = group_sum<SEG>(sum);
            if (l == 0) part_at(kv, r) = sum;
        }
    };
    int cc[PER_T];
    V vv[PER_T];
    if constexpr (XP::aos) {   // an entry's K values in one load, then the K products in turn
        VecK<V, K> xk[PER_T];
        if (tm.cnt > 0) {   // uniform per block
#pragma unroll
            for (int j = 0; j < PER_T; ++j) {
                const int e = min(tid + j * BLOCK, tm.cnt - 1);
                cc[j] = __builtin_nontemporal_load(A.sl_col + tm.base + e);
                vv[j] = __builtin_nontemporal_load(A.sl_val + tm.base + e);
            }
#pragma unroll
            for (int j = 0; j < PER_T; ++j) xk[j] = load_k(X, cc[j]);
        }
#pragma unroll
        for (int kv = 0; kv < K; ++kv) {   // prod is reused by the next vector
            if (kv > 0) __syncthreads();
            if (tm.cnt > 0) {
#pragma unroll
                for (int j = 0; j < PER_T; ++j) prod[tid + j * BLOCK] = (Acc)vv[j] * (Acc)xk[j].v[kv];
            }
            __syncthreads();
            row_parts(kv);
        }
    } else {
        if (tm.cnt > 0) {   // uniform per block; loads and the first vector's gathers together
#pragma unroll
            for (int j = 0; j < PER_T; ++j) {
                const int e = min(tid + j * BLOCK, tm.cnt - 1);
                cc[j] = __builtin_nontemporal_load(A.sl_col + tm.base + e);
                vv[j] = __builtin_nontemporal_load(A.sl_val + tm.base + e);
            }
            Acc xv[PER_T];
#pragma unroll
            for (int j = 0; j < PER_T; ++j) xv[j] = (Acc)X.p[0][cc[j]];
#pragma unroll
            for (int j = 0; j < PER_T; ++j) prod[tid + j * BLOCK] = (Acc)vv[j] * xv[j];
        }
        __syncthreads();
        row_parts(0);
#pragma unroll
        for (int kv = 1; kv < K; ++kv) {   // prod is reused by the next vector
            __syncthreads();
            if (tm.cnt > 0) {
                Acc xv[PER_T];
#pragma unroll
                for (int j = 0; j < PER_T; ++j) xv[j] = (Acc)X.p[kv][cc[j]];
#pragma unroll
                for (int j = 0; j < PER_T; ++j) prod[tid + j * BLOCK] = (Acc)vv[j] * xv[j];
            }
            __syncthreads();
            row_parts(kv);
        }
    }
}

// XCD-sliced product, pass 2: sliced row r's sum = its slice sums added in slice order,
// then the epilogue (thread per row); block partials at index sl_short_tiles() + block.
template <class V, class Acc, int K, class Epi>
__global__ __launch_bounds__(BLOCK) void k_xslice_combine(MatView<V> A, Epi epi) {
    if (epi.skip()) return;
    double q[Epi::NQ > 0 ? Epi::NQ : 1];
#pragma unroll
    for (int i = 0; i < (Epi::NQ > 0 ? Epi::NQ : 1); ++i) q[i] = 0.0;
    const int r = blockIdx.x * BLOCK + threadIdx.x;
    if (r < A.sl_rows) {
        const typename Epi::Pre pre = epi.load(r);
        const Acc* part = (const Acc*)A.sl_part;
        Acc acc[K];
#pragma unroll
        for (int kv = 0; kv < K; ++kv) {
            // all the slice sums in flight together, then added in slice order
            Acc pv[SL_NS_MAX];
#pragma unroll
            for (int s = 0; s < SL_NS_MAX; ++s)
                pv[s] = s < A.ns ? part[((size_t)kv * A.ns + s) * A.sl_rows + r] : Acc(0);
            Acc sum = pv[0];
#pragma unroll
            for (int s = 1; s < SL_NS_MAX; ++s)
                if (s < A.ns) sum += pv[s];
            acc[kv] = sum;
        }
        epi.row(r, acc, q, pre);
    }
    if constexpr (Epi::NQ > 0)
        block_sum<Epi::NQ>(q, epi.partials + (size_t)(A.sl_short_tiles() + blockIdx.x) * Epi::NQ);
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

template <class V, class Acc, int K, class Epi, class XP = XPack<V, K>>
inline void launch_spmv(const MatView<V>& A, const XP& X, const Epi& epi, hipStream_t s) {
    const int g = A.grid();
    if (g <= 0) return;
    if constexpr (K <= SL_KMAX) {
        if (A.format == CPDLP_FMT_XSLICE) {
            hipLaunchKernelGGL((k_spmv_xslice<V, Acc, K, Epi, XP>), dim3(g), dim3(BLOCK), 0, s, A, X, epi);
            if (A.sl_rows > 0)
                hipLaunchKernelGGL((k_xslice_combine<V, Acc, K, Epi>), dim3(ceil_div(A.sl_rows, BLOCK)),
                                   dim3(BLOCK), 0, s, A, epi);
            return;
        }
    }
    // a CSR side with line-grouped tiles: the single-vector products without block
    // partials run on those tiles (the same products and row sums, so the same bits)
    if constexpr (K == 1 && Epi::NQ == 0 && !XP::aos) {
        if (A.format == CPDLP_FMT_CSR && A.lg_col) {
            MatView<V> v = A;
            v.tiles = A.lg_tiles;
            v.n_tiles = A.lg_n_tiles;
            v.n_long_blocks = ceil_div(A.n_long, LG_NT / WAVE);
            hipLaunchKernelGGL((k_spmv_csr<V, Acc, 1, Epi, XP, true, LG_NT, LG_PT>), dim3(v.grid()), dim3(LG_NT),
                               0, s, v, X, epi);
            return;
        }
    }
    if (A.format == CPDLP_FMT_ELL) {
        if constexpr (XP::aos) throw Error(CPDLP_ERR_ARGUMENT, "interleaved vectors need a CSR or sliced format");
        else hipLaunchKernelGGL((k_spmv_ell<V, Acc, K, Epi>), dim3(g), dim3(BLOCK), 0, s, A, X, epi);
    } else {
        // the sliced format's long rows are also in the plain CSR, which serves K > 1
        // and the block-partial reductions
        MatView<V> v = A;
        v.format = CPDLP_FMT_CSR;
        hipLaunchKernelGGL((k_spmv_csr<V, Acc, K, Epi, XP>), dim3(v.grid()), dim3(BLOCK), 0, s, v, X, epi);
    }
}

}  // namespace cpdlp
