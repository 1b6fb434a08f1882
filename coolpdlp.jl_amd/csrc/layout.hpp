// layout.hpp -- host-side construction of the device matrix layouts.
//
// Input: the reference's SparseMatrixCSC (colptr n+1, rowval nnz, 0-based here, rows
// sorted within columns).  Output, all int32:
//   * CSR(A) in the original row order: the stable counting-sort transpose,
//     i.e. sparse(transpose(A)) as GPUSparseMatrixCSR(A) builds it
//     (src/utils/mat_csr.jl:58-61), with perm[k_csr] = k_csc;
//   * CSR(A') = the CSC(A) arrays themselves (src/utils/mat_csr.jl:58-61 applied to
//     MILP.At, src/problems/milp.jl:60);
//   * the internal layouts used by the kernels: rows of A (resp. A') longer than
//     `long_row` are moved to the front by a stable partition, and column indices are
//     relabelled through the other side's partition; the short rows follow in
//     first-touch order (build_layout; density or the caller's order on request).  Entries inside each row keep their original
//     order, so the sequential row sums are unchanged.
#pragma once
#include <algorithm>
#include <cstdint>
#include <limits>
#include <numeric>
#include <thread>
#include <vector>

#include "common.hpp"
#include "spmv.hpp"

namespace cpdlp {

struct HostSide {
    int32_t rows = 0, cols = 0;
    int64_t nnz = 0;
    int32_t n_long = 0;    // rows [0, n_long): long (a wavefront per row in CSR)
    int32_t n_slice = 0;   // rows [0, n_slice) (n_slice >= n_long): cut by the XCD slices
    std::vector<int32_t> rowptr, col, vmap;
    std::vector<TileMeta> tiles;      // short tiles of rows [n_long, rows)
    std::vector<TileMeta> xs_tiles;   // short tiles of rows [n_slice, rows) (sliced format)
};

struct HostLayout {
    int64_t m = 0, n = 0, nnz = 0;
    std::vector<int32_t> colptr, rowval;           // CSC(A) = CSR(A')
    std::vector<int32_t> csr_rowptr, csr_col, perm; // CSR(A), original order
    std::vector<int32_t> row_of_int, int_of_row, col_of_int, int_of_col;
    HostSide A, At;
};

// Normalise the caller's CSC (0/1-based, 32/64-bit) into 0-based int32 arrays and
// validate it (sizes, monotone colptr, row bounds, strictly increasing rows).
inline void read_csc(int64_t m, int64_t n, int32_t base, int32_t bits, const void* colptr,
                     const void* rowval, std::vector<int32_t>& cp, std::vector<int32_t>& rv) {
    if (m < 0 || n < 0) throw Error(CPDLP_ERR_DIMENSION, "negative matrix size");
    if (base != 0 && base != 1) throw Error(CPDLP_ERR_ARGUMENT, "index_base must be 0 or 1");
    if (bits != 32 && bits != 64) throw Error(CPDLP_ERR_ARGUMENT, "index_bits must be 32 or 64");
    if (m > INT32_MAX - 1 || n > INT32_MAX - 1)
        throw Error(CPDLP_ERR_ARGUMENT, "matrix dimension exceeds int32 indexing");
    if (n > 0 && !colptr) throw Error(CPDLP_ERR_ARGUMENT, "colptr is NULL");
    cp.resize(n + 1);
    auto cp_at = [&](int64_t j) -> int64_t {
        return bits == 32 ? (int64_t)((const int32_t*)colptr)[j] : ((const int64_t*)colptr)[j];
    };
    const int64_t c0 = n >= 0 && colptr ? cp_at(0) - base : 0;
    if (c0 != 0) throw Error(CPDLP_ERR_ARGUMENT, "colptr must start at index_base");
    const int64_t nnz = colptr ? cp_at(n) - base : 0;
    if (nnz < 0 || nnz > INT32_MAX) throw Error(CPDLP_ERR_ARGUMENT, "nnz out of int32 range");
    if (nnz > 0 && !rowval) throw Error(CPDLP_ERR_ARGUMENT, "rowval is NULL");
    for (int64_t j = 0; j <= n; ++j) {
        const int64_t v = colptr ? cp_at(j) - base : 0;
        if (v < 0 || v > nnz || (j > 0 && v < cp[j - 1]))
            throw Error(CPDLP_ERR_ARGUMENT, "colptr is not monotone");
        cp[j] = (int32_t)v;
    }
    rv.resize(nnz);
    for (int64_t k = 0; k < nnz; ++k) {
        const int64_t v = (bits == 32 ? (int64_t)((const int32_t*)rowval)[k]
                                      : ((const int64_t*)rowval)[k]) - base;
        if (v < 0 || v >= m) throw Error(CPDLP_ERR_DIMENSION, "row index out of range");
        rv[k] = (int32_t)v;
    }
    for (int64_t j = 0; j < n; ++j)
        for (int32_t k = cp[j] + 1; k < cp[j + 1]; ++k)
            if (rv[k] <= rv[k - 1])
                throw Error(CPDLP_ERR_ARGUMENT, "rows must be strictly increasing within a column");
}

// stable counting-sort transpose: CSR(A) with perm[k_csr] = k_csc
inline void csc_to_csr(int64_t m, int64_t n, const std::vector<int32_t>& cp,
                       const std::vector<int32_t>& rv, std::vector<int32_t>& rp,
                       std::vector<int32_t>& col, std::vector<int32_t>& perm) {
    const int64_t nnz = cp[n];
    rp.assign(m + 1, 0);
    for (int64_t k = 0; k < nnz; ++k) rp[rv[k] + 1]++;
    for (int64_t i = 0; i < m; ++i) rp[i + 1] += rp[i];
    std::vector<int32_t> next(rp.begin(), rp.end() - 1);
    col.resize(nnz);
    perm.resize(nnz);
    for (int64_t j = 0; j < n; ++j)
        for (int32_t k = cp[j]; k < cp[j + 1]; ++k) {
            const int32_t d = next[rv[k]]++;
            col[d] = (int32_t)j;
            perm[d] = k;
        }
}

// partition: long rows (length > long_row) first, longest first (ties by index), so a
// block of four wavefronts gets rows of similar length and the longest start first;
// the short rows follow in their original order, or stably ordered by key[] when given
inline void partition_rows(const std::vector<int32_t>& ptr, int32_t long_row,
                           std::vector<int32_t>& of_int, std::vector<int32_t>& int_of,
                           int32_t& n_long, const std::vector<int32_t>* key = nullptr) {
    const int64_t rows = (int64_t)ptr.size() - 1;
    of_int.clear();
    of_int.reserve(rows);
    for (int64_t i = 0; i < rows; ++i)
        if (ptr[i + 1] - ptr[i] > long_row) of_int.push_back((int32_t)i);
    std::stable_sort(of_int.begin(), of_int.end(), [&](int32_t a, int32_t b) {
        return ptr[a + 1] - ptr[a] > ptr[b + 1] - ptr[b];
    });
    n_long = (int32_t)of_int.size();
    for (int64_t i = 0; i < rows; ++i)
        if (ptr[i + 1] - ptr[i] <= long_row) of_int.push_back((int32_t)i);
    if (key)
        std::stable_sort(of_int.begin() + n_long, of_int.end(),
                         [&](int32_t a, int32_t b) { return (*key)[a] < (*key)[b]; });
    int_of.assign(rows, 0);
    for (int64_t r = 0; r < rows; ++r) int_of[of_int[r]] = (int32_t)r;
}

// greedy short tiles over rows [first, rows): <= max_rows rows, <= max_nnz nnz
inline void build_tiles(const int32_t* rowptr, int32_t first, int32_t rows,
                        std::vector<TileMeta>& tiles, int32_t max_rows = BLOCK,
                        int32_t max_nnz = TILE_NNZ) {
    int32_t r = first;
    while (r < rows) {
        const int32_t r0 = r;
        int32_t rows_in = 0;
        int64_t nz = 0;
        while (r < rows && rows_in < max_rows) {
            const int64_t L = rowptr[r + 1] - rowptr[r];
            if (rows_in > 0 && nz + L > max_nnz) break;
            nz += L;
            ++rows_in;
            ++r;
        }
        tiles.push_back(TileMeta{r0, r, rowptr[r0], rowptr[r] - rowptr[r0]});
    }
}

// Internal side: row r of the internal matrix is source row src_row[r]; its entries
// are src_ptr/src_col/src_k of that row in order, columns mapped by col_map.
inline void build_side(HostSide& S, int32_t rows, int32_t cols, const std::vector<int32_t>& src_ptr,
                       const std::vector<int32_t>& src_col, const std::vector<int32_t>* src_k,
                       const std::vector<int32_t>& row_of_int, const std::vector<int32_t>* col_map,
                       int32_t n_long, int32_t n_slice) {
    S.rows = rows;
    S.cols = cols;
    S.n_long = n_long;
    S.n_slice = std::max(n_slice, n_long);
    S.nnz = src_ptr[rows];
    S.rowptr.assign(rows + 1, 0);
    for (int32_t r = 0; r < rows; ++r) {
        const int32_t i = row_of_int[r];
        S.rowptr[r + 1] = S.rowptr[r] + (src_ptr[i + 1] - src_ptr[i]);
    }
    S.col.resize(S.nnz);
    S.vmap.resize(S.nnz);
    for (int32_t r = 0; r < rows; ++r) {
        const int32_t i = row_of_int[r];
        int32_t d = S.rowptr[r];
        for (int32_t k = src_ptr[i]; k < src_ptr[i + 1]; ++k, ++d) {
            S.col[d] = col_map ? (*col_map)[src_col[k]] : src_col[k];
            S.vmap[d] = src_k ? (*src_k)[k] : k;
        }
    }
    S.tiles.clear();
    build_tiles(S.rowptr.data(), S.n_long, S.rows, S.tiles);
    S.xs_tiles.clear();
    build_tiles(S.rowptr.data(), S.n_slice, S.rows, S.xs_tiles);
}

// Order of the short rows and columns: first touch (the default, below), density
// (increasing stored-entry count, stable: the order sort_rows_columns gives the whole MILP,
// src/components/permutation.jl:16-46) or the caller's own order.  The rows longer than
// min(long_row, slice_row) come first, longest first, in every order: those longer than
// long_row are the long rows (a prefix), the rest of that group are short rows that the
// sliced format also cuts by column slice.
enum class RowOrder { first_touch = 0, density = 1, given = 2 };

inline HostLayout build_layout(int64_t m, int64_t n, std::vector<int32_t>&& cp,
                               std::vector<int32_t>&& rv, int32_t long_row,
                               RowOrder order = RowOrder::first_touch, int32_t slice_row = INT32_MAX) {
    HostLayout L;
    L.m = m;
    L.n = n;
    L.nnz = cp[n];
    L.colptr = std::move(cp);
    L.rowval = std::move(rv);
    csc_to_csr(m, n, L.colptr, L.rowval, L.csr_rowptr, L.csr_col, L.perm);
    const int32_t split = std::min(long_row, slice_row);
    int32_t nlA = 0, nlAt = 0;   // rows longer than split (the group that comes first)
    // First-touch order of the short rows, in two passes.  Pass 1: A's short rows by the
    // first internal row of A' (column of A) that holds them; pass 2: A's short columns
    // by the first (new) row of A that holds them, then A's rows again.  Each step kernel
    // sweeps its rows in order (each XCD a contiguous range of tiles), so the first
    // gather of most entries of y and x comes in sequence and shares a line with its
    // neighbours in that XCD's L2.  Long rows and columns stay first, longest first.
    auto first_touch = [](int64_t outer, const std::vector<int32_t>& ptr, const std::vector<int32_t>& idx,
                          const std::vector<int32_t>& order_of_outer, int64_t inner) {
        std::vector<int32_t> first(inner, std::numeric_limits<int32_t>::max());
        for (int64_t o = 0; o < outer; ++o)
            for (int32_t k = ptr[o]; k < ptr[o + 1]; ++k)
                first[idx[k]] = std::min(first[idx[k]], order_of_outer[o]);
        return first;
    };
    if (order == RowOrder::first_touch) {
        partition_rows(L.colptr, split, L.col_of_int, L.int_of_col, nlAt);
        std::vector<int32_t> key = first_touch(n, L.colptr, L.rowval, L.int_of_col, m);
        partition_rows(L.csr_rowptr, split, L.row_of_int, L.int_of_row, nlA, &key);
        key = first_touch(m, L.csr_rowptr, L.csr_col, L.int_of_row, n);
        partition_rows(L.colptr, split, L.col_of_int, L.int_of_col, nlAt, &key);
        key = first_touch(n, L.colptr, L.rowval, L.int_of_col, m);
        partition_rows(L.csr_rowptr, split, L.row_of_int, L.int_of_row, nlA, &key);
    } else {
        const bool density = order == RowOrder::density;
        auto lengths = [](const std::vector<int32_t>& ptr) {
            std::vector<int32_t> len(ptr.size() - 1);
            for (size_t i = 0; i + 1 < ptr.size(); ++i) len[i] = ptr[i + 1] - ptr[i];
            return len;
        };
        const std::vector<int32_t> row_len = lengths(L.csr_rowptr), col_len = lengths(L.colptr);
        partition_rows(L.csr_rowptr, split, L.row_of_int, L.int_of_row, nlA, density ? &row_len : nullptr);
        partition_rows(L.colptr, split, L.col_of_int, L.int_of_col, nlAt, density ? &col_len : nullptr);
    }
    // the long rows: the prefix of that group longer than long_row (it is longest first)
    auto prefix_longer = [](const std::vector<int32_t>& ptr, const std::vector<int32_t>& of_int, int32_t n_first,
                            int32_t len) {
        int32_t k = 0;
        while (k < n_first && ptr[of_int[k] + 1] - ptr[of_int[k]] > len) ++k;
        return k;
    };
    const int32_t longA = prefix_longer(L.csr_rowptr, L.row_of_int, nlA, long_row);
    const int32_t longAt = prefix_longer(L.colptr, L.col_of_int, nlAt, long_row);
    build_side(L.A, (int32_t)m, (int32_t)n, L.csr_rowptr, L.csr_col, &L.perm, L.row_of_int,
               &L.int_of_col, longA, nlA);
    build_side(L.At, (int32_t)n, (int32_t)m, L.colptr, L.rowval, nullptr, L.col_of_int,
               &L.int_of_row, longAt, nlAt);
    return L;
}

// Line-grouped tiles of a side's short rows: tiles of at most LG_NT rows and LG_TILE_NNZ
// entries (the short rows [n_long, rows) cut greedily, like the short tiles), whose
// entries are stored a second time sorted by column (stably), each packed with its
// position in the tile: col[k] = (slot << LG_COL_BITS) | column.  A 64-lane load
// instruction of the tile then holds neighbouring columns, so lanes whose gathers fall in
// one 128-byte line of the vector share one L2 request (the larger the tile, the more
// often: 0.87 requests per entry in row order, 0.62 in column order at 2,048 entries and
// 0.51 at 4,096 for the large power-law LP's A').  The kernel writes each product to LDS at its slot, so every row is summed
// from the same products in the same order as from the row-ordered tile (bit-identical).
// Only the short rows' range [rowptr[n_long], nnz) of col / vmap is filled.  The estimate
// counts, per 64-entry window, the distinct lines of `line_entries` vector entries the
// gathers touch, in the row-ordered short tiles and in the line-grouped tiles.
struct HostLineTiles {
    std::vector<TileMeta> tiles;
    std::vector<uint32_t> col;
    std::vector<int32_t> vmap;
    int64_t req_rows = 0, req_lines = 0;   // estimated gather requests, row / column order
};

inline bool line_tiles_packable(const HostSide& S) {
    return S.cols <= (int64_t(1) << LG_COL_BITS) && LG_TILE_NNZ <= (int64_t(1) << (32 - LG_COL_BITS));
}

// estimate only (T.col / T.vmap stay empty) when build == false; sample > 1 estimates from
// every sample-th tile of each order.  Tiles are independent: the work is split over up to
// 16 host threads.
inline void line_group_tiles(const HostSide& S, int32_t line_entries, bool build, HostLineTiles& T,
                             int32_t sample = 1) {
    T.req_rows = T.req_lines = 0;
    T.tiles.clear();
    build_tiles(S.rowptr.data(), S.n_long, S.rows, T.tiles, LG_NT, LG_TILE_NNZ);
    if (build) {
        T.col.assign((size_t)std::max<int64_t>(S.nnz, 1), 0u);
        T.vmap.assign((size_t)std::max<int64_t>(S.nnz, 1), 0);
    }
    sample = std::max(sample, 1);
    // distinct lines per 64-entry window of colv[0, cnt)
    auto lines_in_windows = [line_entries](const int32_t* colv, int32_t cnt, std::vector<int32_t>& win) {
        int64_t req = 0;
        for (int32_t w0 = 0; w0 < cnt; w0 += WAVE) {
            win.assign(colv + w0, colv + std::min(cnt, w0 + WAVE));
            for (int32_t& c : win) c /= line_entries;
            std::sort(win.begin(), win.end());
            req += std::unique(win.begin(), win.end()) - win.begin();
        }
        return req;
    };
    // worker w takes tiles [w * n / nw, (w + 1) * n / nw) of both lists
    auto work = [&](size_t w, size_t nw, int64_t& req_rows, int64_t& req_lines) {
        std::vector<int32_t> idx, win, sorted_cols;
        const size_t nr = S.tiles.size(), nt = T.tiles.size();
        for (size_t i = w * nr / nw; i < (w + 1) * nr / nw; ++i) {
            const TileMeta& t = S.tiles[i];
            if (t.cnt > 0 && i % sample == 0) req_rows += lines_in_windows(S.col.data() + t.base, t.cnt, win);
        }
        for (size_t i = w * nt / nw; i < (w + 1) * nt / nw; ++i) {
            const TileMeta& t = T.tiles[i];
            const bool est = i % sample == 0;
            if (t.cnt <= 0 || (!build && !est)) continue;
            const int32_t* c = S.col.data() + t.base;
            idx.resize(t.cnt);
            std::iota(idx.begin(), idx.end(), 0);
            std::stable_sort(idx.begin(), idx.end(), [&](int32_t a, int32_t b) { return c[a] < c[b]; });
            if (est) {
                sorted_cols.resize(t.cnt);
                for (int32_t e = 0; e < t.cnt; ++e) sorted_cols[e] = c[idx[e]];
                req_lines += lines_in_windows(sorted_cols.data(), t.cnt, win);
            }
            if (build)
                for (int32_t e = 0; e < t.cnt; ++e) {
                    T.col[t.base + e] = ((uint32_t)idx[e] << LG_COL_BITS) | (uint32_t)c[idx[e]];
                    T.vmap[t.base + e] = S.vmap[t.base + idx[e]];
                }
        }
    };
    const size_t hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t nw = std::max<size_t>(1, std::min<size_t>({(size_t)16, hw, T.tiles.size() / 64 + 1}));
    std::vector<int64_t> rr(nw, 0), rl(nw, 0);
    std::vector<std::thread> pool;
    for (size_t w = 1; w < nw; ++w) pool.emplace_back(work, w, nw, std::ref(rr[w]), std::ref(rl[w]));
    work(0, nw, rr[0], rl[0]);
    for (std::thread& th : pool) th.join();
    for (size_t w = 0; w < nw; ++w) {
        T.req_rows += rr[w];
        T.req_lines += rl[w];
    }
}

// ELL (column-major rows x width, -1 padding) of an internal side, or width -1 if the
// padding would exceed `max_fill` times nnz.
inline int32_t build_ell(const HostSide& S, double max_fill, std::vector<int32_t>& ecol,
                         std::vector<int32_t>& evmap) {
    int32_t w = 0;
    for (int32_t r = 0; r < S.rows; ++r) w = std::max(w, S.rowptr[r + 1] - S.rowptr[r]);
    if (S.rows == 0 || (double)w * S.rows > max_fill * std::max<int64_t>(S.nnz, 1)) return -1;
    ecol.assign((size_t)w * S.rows, -1);
    evmap.assign((size_t)w * S.rows, -1);
    for (int32_t r = 0; r < S.rows; ++r)
        for (int32_t k = S.rowptr[r]; k < S.rowptr[r + 1]; ++k) {
            const size_t s = (size_t)(k - S.rowptr[r]) * S.rows + r;
            ecol[s] = S.col[k];
            evmap[s] = S.vmap[k];
        }
    return w;
}

// XCD column slices of the first `nrows` rows of an internal side (its rows [0, n_slice),
// or all of its rows).  Column slice s holds a contiguous range of internal columns (ranges
// balanced by entry count, slice_of_columns); slice s's part of row r (r < nrows) is the
// run of that row's entries whose column lies in the slice, in the row's entry order.  rowptr:
// ns * (nrows + 1) absolute offsets, slices stored one after the other.  Each slice's
// rows are cut into tiles of at most TILE_NNZ entries and BLOCK rows (rows whose parts
// are summed by lane groups: the first n_group rows, by default the long rows n_long) or
// SL_RPT * BLOCK rows (the others, a thread per row part) that never span that boundary
// nor n_slice (so the tiles of rows [0, n_slice) are a prefix of every slice's tiles,
// tile_n_pre[s] of them); a row part longer than TILE_NNZ is a tile of its own (r1 = r0 +
// 1, cnt > TILE_NNZ).  Rows past nrows are not sliced.
struct HostSlices {
    int32_t ns = 0, rows = 0, max_tiles = 0, max_tiles_pre = 0;
    int32_t n_group = 0;   // rows [0, n_group): parts summed by lane groups, the rest a thread each
    std::vector<int32_t> rowptr, col, vmap, tile_off, tile_n_pre;
    std::vector<TileMeta> tiles;
};

// Column slice boundaries: contiguous column ranges holding about nnz / ns entries of
// rows [0, nrows) each (a skewed column distribution -- e.g. the other side's long rows,
// which come first -- would otherwise load one XCD with most of the work).
inline std::vector<int8_t> slice_of_columns(const HostSide& S, int32_t ns, int32_t nrows) {
    std::vector<int64_t> per_col(S.cols, 0);
    for (int64_t k = 0; k < S.rowptr[nrows]; ++k) per_col[S.col[k]]++;
    const int64_t total = S.rowptr[nrows];
    std::vector<int8_t> sl(S.cols, 0);
    int64_t acc = 0;
    int32_t cur = 0;
    for (int32_t j = 0; j < S.cols; ++j) {
        // move to the next slice once this one holds its share
        while (cur < ns - 1 && acc * ns >= (int64_t)(cur + 1) * total) ++cur;
        sl[j] = (int8_t)cur;
        acc += per_col[j];
    }
    return sl;
}

inline void build_slices(const HostSide& S, int32_t ns, int32_t nrows, HostSlices& H, int32_t n_group = -1) {
    const int32_t nr = nrows, np = std::min(std::max(S.n_slice, S.n_long), nrows);
    const int32_t nl = std::min(n_group < 0 ? S.n_long : std::min(n_group, S.n_long), nrows);
    H.n_group = nl;
    const std::vector<int8_t> slice_of = slice_of_columns(S, ns, nr);
    H.ns = ns;
    H.rows = nr;
    H.rowptr.assign((size_t)ns * (nr + 1), 0);
    std::vector<int32_t> cnt((size_t)ns * nr, 0);
    for (int32_t r = 0; r < nr; ++r)
        for (int32_t k = S.rowptr[r]; k < S.rowptr[r + 1]; ++k)
            cnt[(size_t)slice_of[S.col[k]] * nr + r]++;
    int64_t off = 0;
    for (int32_t sl = 0; sl < ns; ++sl) {
        int32_t* rp = H.rowptr.data() + (size_t)sl * (nr + 1);
        for (int32_t r = 0; r < nr; ++r) {
            rp[r] = (int32_t)off;
            off += cnt[(size_t)sl * nr + r];
        }
        rp[nr] = (int32_t)off;
    }
    std::vector<int32_t>().swap(cnt);
    H.col.resize(off);
    H.vmap.resize(off);
    std::vector<int32_t> next(H.rowptr);
    for (int32_t r = 0; r < nr; ++r)
        for (int32_t k = S.rowptr[r]; k < S.rowptr[r + 1]; ++k) {
            const int32_t sl = slice_of[S.col[k]];
            const int32_t d = next[(size_t)sl * (nr + 1) + r]++;
            H.col[d] = S.col[k];
            H.vmap[d] = S.vmap[k];
        }
    H.tiles.clear();
    H.tile_off.assign(ns + 1, 0);
    H.tile_n_pre.assign(ns, 0);
    H.max_tiles = H.max_tiles_pre = 0;
    for (int32_t sl = 0; sl < ns; ++sl) {
        H.tile_off[sl] = (int32_t)H.tiles.size();
        const int32_t* rp = H.rowptr.data() + (size_t)sl * (nr + 1);
        int32_t r = 0;
        while (r < nr) {
            if (r == np) H.tile_n_pre[sl] = (int32_t)H.tiles.size() - H.tile_off[sl];
            const int32_t r0 = r, lim = r0 < nl ? nl : r0 < np ? np : nr;
            // short rows' parts are summed a thread per row, SL_RPT rows per thread
            const int32_t max_rows = r0 < nl ? BLOCK : SL_RPT * BLOCK;
            if (rp[r + 1] - rp[r] > TILE_NNZ) {
                ++r;
            } else {
                while (r < lim && r - r0 < max_rows && rp[r + 1] - rp[r] <= TILE_NNZ &&
                       rp[r + 1] - rp[r0] <= TILE_NNZ)
                    ++r;
            }
            H.tiles.push_back(TileMeta{r0, r, rp[r0], rp[r] - rp[r0]});
        }
        if (np == nr) H.tile_n_pre[sl] = (int32_t)H.tiles.size() - H.tile_off[sl];
        H.max_tiles = std::max(H.max_tiles, (int32_t)H.tiles.size() - H.tile_off[sl]);
        H.max_tiles_pre = std::max(H.max_tiles_pre, H.tile_n_pre[sl]);
    }
    H.tile_off[ns] = (int32_t)H.tiles.size();
}

}  // namespace cpdlp
