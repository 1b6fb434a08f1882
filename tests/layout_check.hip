// Host-only checks of the device layouts built by csrc/layout.hpp (no GPU needed):
// partition_rows, build_tiles, build_layout, build_slices and line_group_tiles on random matrices with
// long and short rows.  Prints "ok" and exits 0, or prints the first violated invariant
// and exits 1.  Built and run by tests/test_layout.py.
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <random>

#include "../coolpdlp.jl_amd/csrc/layout.hpp"

using namespace cpdlp;

static void fail(const char* what, long a, long b) {
    std::printf("FAIL %s (%ld, %ld)\n", what, a, b);
    std::exit(1);
}

// a side whose first n_long rows are long (65..maxlong entries), the rest 0..64
static HostSide make_side(std::mt19937& rng, int rows, int cols, int n_long, int maxlong, bool skew) {
    HostSide S;
    S.rows = rows;
    S.cols = cols;
    S.n_long = n_long;
    S.rowptr.assign(rows + 1, 0);
    std::vector<int> len(rows);
    for (int r = 0; r < rows; ++r)
        len[r] = std::min(cols, r < n_long ? 65 + (int)(rng() % (maxlong - 64)) : (int)(rng() % 65));
    for (int r = 0; r < rows; ++r) S.rowptr[r + 1] = S.rowptr[r] + len[r];
    S.nnz = S.rowptr[rows];
    S.col.resize(S.nnz);
    S.vmap.resize(S.nnz);
    std::vector<int> pick(cols);
    for (int r = 0; r < rows; ++r) {
        for (int j = 0; j < cols; ++j) pick[j] = j;
        // skew: the first tenth of the columns is drawn 20x as often (a hot head)
        if (skew) {
            for (int k = 0; k < len[r]; ++k) {
                const int lim = (rng() % 3 == 0) ? cols : std::max(1, cols / 10);
                std::swap(pick[k], pick[k + rng() % (lim > k ? lim - k : 1)]);
            }
        } else {
            std::shuffle(pick.begin(), pick.end(), rng);
        }
        std::sort(pick.begin(), pick.begin() + len[r]);
        for (int k = 0; k < len[r]; ++k) {
            S.col[S.rowptr[r] + k] = pick[k];
            S.vmap[S.rowptr[r] + k] = S.rowptr[r] + k;
        }
    }
    return S;
}

static void check_slices(const HostSide& S, int nr, int n_group = -1) {
    const int ns = 8, np = std::min(std::max(S.n_slice, S.n_long), nr);
    const int nl = std::min(n_group < 0 ? S.n_long : std::min(n_group, S.n_long), nr);
    HostSlices H;
    build_slices(S, ns, nr, H, n_group);
    if (H.n_group != nl) fail("lane-group rows", H.n_group, nl);
    const std::vector<int8_t> so = slice_of_columns(S, ns, nr);
    // contiguous column ranges, in slice order
    for (int j = 1; j < S.cols; ++j)
        if (so[j] < so[j - 1]) fail("slice ranges not contiguous", j, so[j]);
    // balance: a slice holds at most its share plus the largest column's count
    std::vector<long> per_col(S.cols, 0), per_slice(ns, 0);
    for (long k = 0; k < S.rowptr[nr]; ++k) per_col[S.col[k]]++;
    long maxcol = 0;
    for (int j = 0; j < S.cols; ++j) {
        per_slice[so[j]] += per_col[j];
        maxcol = std::max(maxcol, per_col[j]);
    }
    const long share = (S.rowptr[nr] + ns - 1) / ns;
    for (int s = 0; s < ns; ++s)
        if (per_slice[s] > share + maxcol) fail("slice over its share", s, per_slice[s]);
    // every row's parts, in slice order, are the row's entries in their order
    for (int r = 0; r < nr; ++r) {
        long at = S.rowptr[r];
        for (int s = 0; s < ns; ++s) {
            const int32_t* rp = H.rowptr.data() + (size_t)s * (nr + 1);
            for (int k = rp[r]; k < rp[r + 1]; ++k, ++at) {
                if (at >= S.rowptr[r + 1]) fail("row part too long", r, s);
                if (H.vmap[k] != S.vmap[at] || H.col[k] != S.col[at]) fail("row part order", r, k);
                if (so[H.col[k]] != s) fail("entry in the wrong slice", r, k);
            }
        }
        if (at != S.rowptr[r + 1]) fail("row part missing entries", r, at);
    }
    // tiles: cover [0, nr) per slice, limits, never span the long/short boundary nor
    // n_slice, and the tiles of rows [0, n_slice) are the first tile_n_pre of the slice
    int maxt = 0, maxl = 0;
    for (int s = 0; s < ns; ++s) {
        const int32_t* rp = H.rowptr.data() + (size_t)s * (nr + 1);
        int expect = 0;
        for (int t = H.tile_off[s]; t < H.tile_off[s + 1]; ++t) {
            const TileMeta& tm = H.tiles[t];
            if (tm.r0 != expect || tm.r1 <= tm.r0 || tm.r1 > nr) fail("tile cover", s, t);
            const int cap = tm.r0 < nl ? BLOCK : SL_RPT * BLOCK;
            if (tm.r1 - tm.r0 > cap) fail("tile rows", s, t);
            if (tm.base != rp[tm.r0] || tm.cnt != rp[tm.r1] - rp[tm.r0]) fail("tile base", s, t);
            if (tm.cnt > TILE_NNZ && tm.r1 - tm.r0 != 1) fail("oversized tile of several rows", s, t);
            if (tm.r0 < nl && tm.r1 > nl) fail("tile spans the long/short boundary", s, t);
            if (tm.r0 < np && tm.r1 > np) fail("tile spans n_slice", s, t);
            const bool in_prefix = t - H.tile_off[s] < H.tile_n_pre[s];
            if (in_prefix != (tm.r0 < np)) fail("prefix tile count", s, t);
            expect = tm.r1;
        }
        if (expect != nr) fail("tiles end early", s, expect);
        maxt = std::max(maxt, H.tile_off[s + 1] - H.tile_off[s]);
        maxl = std::max(maxl, H.tile_n_pre[s]);
    }
    if (maxt != H.max_tiles || maxl != H.max_tiles_pre) fail("max tiles", maxt, maxl);
}

static void check_partition_and_tiles(std::mt19937& rng) {
    // partition_rows: long rows first, longest first (stable), then the short rows in order
    const int rows = 5000;
    std::vector<int32_t> ptr(rows + 1, 0);
    for (int r = 0; r < rows; ++r) ptr[r + 1] = ptr[r] + (int)(rng() % 10 == 0 ? 60 + rng() % 300 : rng() % 40);
    std::vector<int32_t> of_int, int_of;
    int32_t n_long = 0;
    partition_rows(ptr, 64, of_int, int_of, n_long);
    auto L = [&](int i) { return ptr[i + 1] - ptr[i]; };
    for (int r = 0; r < rows; ++r) {
        if (int_of[of_int[r]] != r) fail("partition not a permutation", r, of_int[r]);
        if ((r < n_long) != (L(of_int[r]) > 64)) fail("long rows not first", r, L(of_int[r]));
        if (r > 0 && r < n_long && L(of_int[r]) > L(of_int[r - 1])) fail("long rows not longest first", r, 0);
        if (r > 0 && r < n_long && L(of_int[r]) == L(of_int[r - 1]) && of_int[r] < of_int[r - 1])
            fail("long rows not stable", r, 0);
        if (r > n_long && of_int[r] < of_int[r - 1]) fail("short rows out of order", r, 0);
    }
    // build_tiles over the short rows: <= BLOCK rows, <= TILE_NNZ entries, contiguous
    std::vector<int32_t> rp(rows + 1, 0);
    for (int r = 0; r < rows; ++r) rp[r + 1] = rp[r] + L(of_int[r]);
    std::vector<TileMeta> tiles;
    build_tiles(rp.data(), n_long, rows, tiles);
    int expect = n_long;
    for (size_t t = 0; t < tiles.size(); ++t) {
        const TileMeta& tm = tiles[t];
        if (tm.r0 != expect || tm.r1 - tm.r0 > BLOCK || tm.cnt > TILE_NNZ) fail("short tile", (long)t, tm.r0);
        if (tm.base != rp[tm.r0] || tm.cnt != rp[tm.r1] - rp[tm.r0]) fail("short tile base", (long)t, 0);
        expect = tm.r1;
    }
    if (expect != rows) fail("short tiles end early", expect, rows);
}

// build_layout: both orders are permutations with the long rows (columns) first; A's
// short rows are in first-touch order of the final column order; every internal row is
// its source row with columns relabelled, entry order kept (so row sums are unchanged)
static void check_build_layout(std::mt19937& rng) {
    const int m = 3000, n = 7000;
    std::vector<std::vector<int32_t>> colrows(n);
    for (int j = 0; j < n; ++j) {
        const int len = j % 97 == 0 ? 80 + (int)(rng() % 200) : (int)(rng() % 12);   // a few long columns
        std::vector<int32_t> rs;
        for (int k = 0; k < len; ++k) rs.push_back((int32_t)(rng() % (k % 5 == 0 ? 50 : m)));   // hot rows
        std::sort(rs.begin(), rs.end());
        rs.erase(std::unique(rs.begin(), rs.end()), rs.end());
        colrows[j] = rs;
    }
    std::vector<int32_t> cp(n + 1, 0), rv;
    for (int j = 0; j < n; ++j) {
        rv.insert(rv.end(), colrows[j].begin(), colrows[j].end());
        cp[j + 1] = (int32_t)rv.size();
    }
    const std::vector<int32_t> cp0 = cp, rv0 = rv;
    const HostLayout L = build_layout(m, n, std::move(cp), std::move(rv), 64);
    auto is_perm = [](const std::vector<int32_t>& of, const std::vector<int32_t>& inv, int len) {
        if ((int)of.size() != len || (int)inv.size() != len) return false;
        for (int i = 0; i < len; ++i)
            if (of[i] < 0 || of[i] >= len || inv[of[i]] != i) return false;
        return true;
    };
    if (!is_perm(L.row_of_int, L.int_of_row, m)) fail("row order not a permutation", 0, 0);
    if (!is_perm(L.col_of_int, L.int_of_col, n)) fail("column order not a permutation", 0, 0);
    auto rlen = [&](int i) { return L.csr_rowptr[i + 1] - L.csr_rowptr[i]; };
    auto clen = [&](int j) { return cp0[j + 1] - cp0[j]; };
    for (int r = 0; r < m; ++r)
        if ((r < L.A.n_long) != (rlen(L.row_of_int[r]) > 64)) fail("long rows not first", r, 0);
    for (int c = 0; c < n; ++c)
        if ((c < L.At.n_long) != (clen(L.col_of_int[c]) > 64)) fail("long columns not first", c, 0);
    // first touch of each row in the final column order, nondecreasing over the short rows
    std::vector<int32_t> first(m, INT32_MAX);
    for (int j = 0; j < n; ++j)
        for (int k = cp0[j]; k < cp0[j + 1]; ++k) first[rv0[k]] = std::min(first[rv0[k]], L.int_of_col[j]);
    for (int r = L.A.n_long + 1; r < m; ++r)
        if (first[L.row_of_int[r]] < first[L.row_of_int[r - 1]]) fail("short rows not in first-touch order", r, 0);
    // internal rows of A and A' are the source rows, relabelled, in entry order
    for (int r = 0; r < m; ++r) {
        const int i = L.row_of_int[r];
        if (L.A.rowptr[r + 1] - L.A.rowptr[r] != rlen(i)) fail("A row length", r, i);
        for (int k = 0; k < rlen(i); ++k) {
            const int src = L.csr_rowptr[i] + k;
            if (L.A.col[L.A.rowptr[r] + k] != L.int_of_col[L.csr_col[src]]) fail("A row entry", r, k);
            if (L.A.vmap[L.A.rowptr[r] + k] != L.perm[src]) fail("A value map", r, k);
        }
    }
    for (int c = 0; c < n; ++c) {
        const int j = L.col_of_int[c];
        if (L.At.rowptr[c + 1] - L.At.rowptr[c] != clen(j)) fail("A' row length", c, j);
        for (int k = 0; k < clen(j); ++k) {
            if (L.At.col[L.At.rowptr[c] + k] != L.int_of_row[rv0[cp0[j] + k]]) fail("A' row entry", c, k);
            if (L.At.vmap[L.At.rowptr[c] + k] != cp0[j] + k) fail("A' value map", c, k);
        }
    }
}

// line-grouped tiles: the short rows cut into tiles of <= LG_NT rows and <= LG_TILE_NNZ
// entries, contiguous; each tile's copy holds every slot of the tile once, sorted by column
// (stable: equal columns in slot order), with the column and value map of the entry at that
// slot; the request estimates count distinct lines per 64-entry window (row order: the
// short tiles; column order: the line-grouped tiles)
static long windows_lines(const int32_t* c, int cnt, int line) {
    long req = 0;
    for (int w0 = 0; w0 < cnt; w0 += 64) {
        std::vector<int> ls;
        for (int e = w0; e < std::min(cnt, w0 + 64); ++e) ls.push_back(c[e] / line);
        std::sort(ls.begin(), ls.end());
        req += std::unique(ls.begin(), ls.end()) - ls.begin();
    }
    return req;
}

static void check_line_tiles(HostSide S, int line) {
    S.tiles.clear();
    build_tiles(S.rowptr.data(), S.n_long, S.rows, S.tiles);
    HostLineTiles T;
    line_group_tiles(S, line, true, T);
    if ((long)T.col.size() != std::max<long>(S.nnz, 1)) fail("line tiles size", (long)T.col.size(), S.nnz);
    int expect = S.n_long;
    long rows_req = 0, lines_req = 0;
    for (const TileMeta& tm : S.tiles) rows_req += windows_lines(S.col.data() + tm.base, tm.cnt, line);
    for (const TileMeta& tm : T.tiles) {
        if (tm.r0 != expect || tm.r1 <= tm.r0 || tm.r1 - tm.r0 > LG_NT) fail("line tile rows", tm.r0, tm.r1);
        if (tm.base != S.rowptr[tm.r0] || tm.cnt != S.rowptr[tm.r1] - S.rowptr[tm.r0]) fail("line tile base", tm.r0, 0);
        if (tm.cnt > LG_TILE_NNZ && tm.r1 - tm.r0 > 1) fail("line tile entries", tm.r0, tm.cnt);
        expect = tm.r1;
        std::vector<int> seen(tm.cnt, 0), sorted_cols(tm.cnt);
        for (int e = 0; e < tm.cnt; ++e) {
            const uint32_t p = T.col[tm.base + e];
            const int slot = (int)(p >> LG_COL_BITS), c = (int)(p & LG_COL_MASK);
            if (slot < 0 || slot >= tm.cnt || seen[slot]++) fail("line tile slots not a permutation", tm.base, e);
            if (c != S.col[tm.base + slot] || T.vmap[tm.base + e] != S.vmap[tm.base + slot]) fail("line tile entry", tm.base, e);
            if (e > 0) {
                const uint32_t q = T.col[tm.base + e - 1];
                const int cq = (int)(q & LG_COL_MASK), sq = (int)(q >> LG_COL_BITS);
                if (cq > c || (cq == c && sq > slot)) fail("line tile not sorted stably", tm.base, e);
            }
            sorted_cols[e] = c;
        }
        lines_req += windows_lines(sorted_cols.data(), tm.cnt, line);
    }
    if (expect != S.rows) fail("line tiles end early", expect, S.rows);
    if (T.req_rows != rows_req) fail("row-order request estimate", T.req_rows, rows_req);
    if (T.req_lines != lines_req) fail("line-order request estimate", T.req_lines, lines_req);
    HostLineTiles E;
    line_group_tiles(S, line, false, E);
    if (E.req_rows != T.req_rows || E.req_lines != T.req_lines || !E.col.empty()) fail("estimate-only pass", 0, 0);
}

// a hot head of columns: the line-grouped tiles need fewer requests than the rows
static void check_line_tiles_gain(std::mt19937& rng) {
    HostSide S = make_side(rng, 3000, 4000, 0, 100, true);
    S.tiles.clear();
    build_tiles(S.rowptr.data(), 0, S.rows, S.tiles);
    HostLineTiles T;
    line_group_tiles(S, 16, false, T);
    if (!(T.req_lines * 10 < T.req_rows * 9)) fail("no gain on a hot head", T.req_lines, T.req_rows);
}

// slice_row below long_row: the rows longer than slice_row come first, longest first; the
// long rows (longer than long_row) are their prefix; the short tiles cover [n_long, rows)
// and the sliced format's short tiles [n_slice, rows)
static void check_slice_split(std::mt19937& rng) {
    const int m = 2500, n = 6000;
    std::vector<int32_t> cp(n + 1, 0), rv;
    for (int j = 0; j < n; ++j) {
        std::vector<int32_t> rs;
        const int len = (int)(rng() % 14);
        for (int k = 0; k < len; ++k) rs.push_back((int32_t)(rng() % (k % 3 == 0 ? 200 : m)));
        std::sort(rs.begin(), rs.end());
        rs.erase(std::unique(rs.begin(), rs.end()), rs.end());
        rv.insert(rv.end(), rs.begin(), rs.end());
        cp[j + 1] = (int32_t)rv.size();
    }
    const HostLayout L = build_layout(m, n, std::move(cp), std::move(rv), 64, RowOrder::first_touch, 8);
    for (const HostSide* S : {&L.A, &L.At}) {
        auto len = [&](int r) { return S->rowptr[r + 1] - S->rowptr[r]; };
        if (S->n_slice < S->n_long) fail("n_slice below n_long", S->n_slice, S->n_long);
        for (int r = 0; r < S->rows; ++r) {
            if ((r < S->n_long) != (len(r) > 64)) fail("long rows not the prefix", r, len(r));
            if ((r < S->n_slice) != (len(r) > 8)) fail("sliced rows not the prefix", r, len(r));
            if (r > 0 && r < S->n_slice && len(r) > len(r - 1)) fail("sliced rows not longest first", r, 0);
        }
        int expect = S->n_slice;
        for (const TileMeta& tm : S->xs_tiles) {
            if (tm.r0 != expect || tm.r1 - tm.r0 > BLOCK || tm.cnt > TILE_NNZ) fail("xs tile", tm.r0, tm.r1);
            expect = tm.r1;
        }
        if (expect != S->rows) fail("xs tiles end early", expect, S->rows);
        expect = S->n_long;
        for (const TileMeta& tm : S->tiles) {
            if (tm.r0 != expect) fail("short tile", tm.r0, expect);
            expect = tm.r1;
        }
        if (expect != S->rows) fail("short tiles end early", expect, S->rows);
    }
    if (L.A.n_slice == L.A.n_long) fail("no sliced short rows in the test matrix", L.A.n_slice, 0);
}

int main() {
    std::mt19937 rng(12345);
    check_slice_split(rng);
    check_partition_and_tiles(rng);
    check_build_layout(rng);
    for (int trial = 0; trial < 6; ++trial) {
        const int rows = 2000 + 700 * trial, cols = 3000 + 1100 * trial;
        const int n_long = trial == 0 ? 0 : 20 + 15 * trial;
        const int maxlong = trial == 5 ? 9000 : 2500;   // trial 5: row parts longer than a tile
        const HostSide S = make_side(rng, rows, cols, n_long, maxlong, trial % 2 == 1);
        check_slices(S, S.n_long);   // xslice: long rows
        check_slices(S, S.rows);     // xslice_all: every row
        HostSide S2 = S;             // sliced prefix beyond the long rows
        S2.n_slice = std::min(S.rows, S.n_long + 300);
        check_slices(S2, S2.n_slice);
        check_slices(S2, S2.rows);
        for (int ng : {0, S.n_long / 3}) {   // fewer rows summed by lane groups
            check_slices(S2, S2.n_slice, ng);
            check_slices(S2, S2.rows, ng);
        }
        check_line_tiles(S, 16);
        check_line_tiles(S, 4);
    }
    check_line_tiles_gain(rng);
    std::printf("ok\n");
    return 0;
}
