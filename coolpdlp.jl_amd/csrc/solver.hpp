// solver.hpp -- host-side objects of libcoolpdlp_hip.so.
#pragma once
#include <cfloat>
#include <utility>
#include <vector>

#include "kernels.hpp"
#include "layout.hpp"

namespace cpdlp {

// one matrix (A or A') on the device, internal layout; FP64 values always, FP32 copy
// for the FP32 / mixed loop; optional column-major ELL of the same rows
struct DevSide {
    int32_t rows = 0, cols = 0, n_long = 0, n_slice = 0, n_tiles = 0, n_xs_tiles = 0;
    int64_t nnz = 0;
    int32_t format = CPDLP_FMT_CSR;
    int32_t ell_w = 0;
    DevBuf<int32_t> rowptr, col, vmap;
    DevBuf<TileMeta> tiles, xs_tiles;
    DevBuf<double> val64;
    DevBuf<float> val32;
    DevBuf<int32_t> ell_col, ell_vmap;
    DevBuf<double> ell_val64;
    DevBuf<float> ell_val32;
    // XCD column slices of rows [0, sl_built) (rows [0, n_slice), or all rows)
    int32_t sl_ns = 0, sl_built = 0, sl_max_tiles = 0, sl_max_tiles_pre = 0;
    int32_t sl_n_group = 0;   // sliced rows whose parts are summed by lane groups (a prefix)
    int64_t sl_nnz = 0;
    DevBuf<int32_t> sl_rowptr, sl_col, sl_vmap, sl_tile_off, sl_tile_n, sl_tile_n_pre;
    DevBuf<TileMeta> sl_tiles;
    DevBuf<double> sl_val64, sl_part;
    DevBuf<float> sl_val32;
    // line-grouped copy of the short tiles (built when it saves gather requests; used
    // while lg_on)
    bool lg_built = false, lg_on = false;
    int32_t lg_n_tiles = 0;
    DevBuf<TileMeta> lg_tiles;
    DevBuf<uint32_t> lg_col;
    DevBuf<int32_t> lg_vmap;
    DevBuf<double> lg_val64;
    DevBuf<float> lg_val32;

    template <class V>
    const V* vals() const;
    template <class V>
    const V* ell_vals() const;
    template <class V>
    const V* sl_vals() const;
    template <class V>
    const V* lg_vals() const;
    // allow_slices = false: the plain CSR copy instead of the XCD slices
    template <class V>
    MatView<V> view(bool allow_slices = true) const;
    void upload_structure(const HostSide& S, hipStream_t s);
    void build_slices(const HostSide& S, int32_t nrows, int32_t ns, hipStream_t s, int32_t n_group = -1);
    bool slices_all() const { return sl_ns > 0 && sl_built == rows; }
    void build_ell(const HostSide& S, hipStream_t s, double max_fill);
    void build_line_tiles(const HostLineTiles& T, hipStream_t s);
    void fill_values(const double* src_csc, bool want32, hipStream_t s);
};

// PDLPState's solutions (pdlp.jl:20-43): x[cur], y[cur] = sol; x[1-cur], y[1-cur] =
// sol_last (the two swap every step, pdlp.jl:101); xbar = scratch.x of the step.
template <class V>
struct IterState {
    DevBuf<V> x[2], y[2];
    DevBuf<V> xbar, xavg, yavg, xavgl, yavgl, xrst, yrst, tmp_n, tmp_m;
    int cur = 0;
    void alloc(int64_t n, int64_t m);
};

struct GraphCache {
    hipGraphExec_t exec = nullptr;
    int end_parity = 0;
};

struct Handle {
    cpdlp_params P;
    int64_t m = 0, n = 0, nnz = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    HostLayout L;
    // inputs, original order
    DevBuf<double> d_nzval, d_c, d_lv, d_uv, d_lc, d_uc;
    DevBuf<int32_t> d_colptr, d_rowval, d_csr_rowptr, d_perm;
    DevBuf<int32_t> d_row_of_int, d_col_of_int, d_int_of_row, d_int_of_col;
    // scaling results, original order / CSC(A) layout
    DevBuf<double> d_D1o, d_D2o, d_a_csc, d_at_csc;
    // scaled problem, internal order
    DevBuf<double> c_s, lv_s, uv_s, lc_s, uc_s, D1_s, D2_s, invD1_s, invD2_s;
    DevBuf<float> c_32, lv_32, uv_32, lc_32, uc_32;
    DevSide A, At;
    // scratch
    DevBuf<double> d_partials, d_partials2, d_sums;
    // the 4-candidate check's vectors interleaved (x: 4 n, y: 4 m entries of the check's
    // precision), so that a gather brings all four with one request
    DevBuf<double> aos_x, aos_y;
    DevBuf<StepCoef> d_coef;
    GraphCache graphs[4];
    // solver scalars
    bool prepared = false, solved = false, active64 = true, stepping = false;
    double eta = 0, omega0 = 1, omega = 1, eta_sum = 0;
    double primal_scale = 1, dual_scale = 1, c_norm = 0, comb_norm = 0;
    int64_t powm_iters = 0, inner = 0, total = 0, restarts = 0, step_launches = 0, fp32_passes = 0;
    double setup_seconds = 0, loop_seconds = 0, step_ms = 0, t_start = 0, t_loop = 0;
    DevBuf<double> x0d, y0d;   // the start point of the current solve
    // mixed precision hand-over (FP32 phase -> FP64 polish)
    static constexpr int STALL_CHECKS = 10;
    double best_rel = INFINITY, best_omega = 1.0, last_progress = 0.0;
    int stall = 0;
    bool phase_switch = false, blowup = false;
    // KKT errors of the restart point (PDLP): it changes only at a restart, where it becomes
    // the candidate, so the check evaluates 4 candidates and reuses these; invalid after
    // the start and after a change of precision
    cpdlp_kkt rst_err{};
    bool rst_err_valid = false;
    bool no_rst_cache = false;   // CPDLP_KKT_FULL_CHECK set: evaluate all 5 every check
    bool no_aos = false;         // CPDLP_KKT_SEPARATE set: the 4 candidates as separate vectors
    int lg_mode = -1;            // CPDLP_LINE_TILES: -1 by estimate and timing, 0 never, 1 always
    IterState<double> st64;
    IterState<float> st32;
    cpdlp_stats stats{};
    std::vector<std::pair<int64_t, cpdlp_kkt>> history;

    Handle(const cpdlp_lp& lp, const cpdlp_params& prm);
    ~Handle();
    int grid_n() const;
    int grid_m() const;
    void sync();
    void ensure_partials(size_t count);
    void reduce_to_host(int nb, int nq, const double* partials, double* host_out);
    void scale();
    void choose_formats();
    template <class V>
    void power_method(double tol, int64_t maxiter);
    void prepare();
    template <class V, int K>
    void kkt_launch(const V* const* X, const V* const* Y, const V* xr, const V* yr, bool dist,
                    double* host_sums);
    // kkt_launch on caller-given device buffers: partials (rows of A), partials2 (rows of
    // A'), sums_dev (the reduced sums); allocates nothing
    template <class V, int K>
    void kkt_launch_into(const V* const* X, const V* const* Y, const V* xr, const V* yr, bool dist,
                         double* host_sums, double* partials, double* partials2, double* sums_dev);
    // doubles of device scratch kkt_device needs (x and y in internal order, block
    // partials of both products, the sums)
    size_t kkt_scratch_doubles() const;
    cpdlp_kkt kkt_device(const double* x_dev, const double* y_dev, double* scratch, size_t scratch_len);
    cpdlp_kkt kkt_from_sums(const double* sums, int K, int k) const;
    template <class V>
    int32_t ray_check(IterState<V>& S);
    template <class V>
    void launch_steps(IterState<V>& S, int nsteps, bool pdlp);
    template <class V>
    void run_chunk(IterState<V>& S, bool pdlp);
    template <class V>
    void loop(IterState<V>& S, double phase_reltol, bool final_phase,
              int64_t max_checks = INT64_MAX);
    template <class V>
    void init_state(IterState<V>& S, const double* x0_dev, const double* y0_dev);
    bool begin(const double* x0, const double* y0);
    void finish_stats();
    void solve(const double* x0, const double* y0);
    void initialize(const double* x0, const double* y0);
    void hand_over_fp64();
    int32_t run_checks(int64_t max_checks);
    template <class V>
    void state_to(IterState<V>& S, double* xs, double* ys, const double* ones);
    void get_state(cpdlp_state* st, double* x, double* y, double* D2, double* D1);
    void resume(const cpdlp_state& st, const double* x, const double* y, const double* D2, const double* D1);
    void solution_to(double* xo, double* yo);
    void get_solution_device(double* x_dev, double* y_dev);
    void progress(bool final_line);
    void get_solution(double* x, double* y);
    cpdlp_kkt kkt_host(const double* x, const double* y);
    template <class V>
    void profile_steps_t(IterState<V>& S, int reps, double* ms_primal, double* ms_dual);
    void profile_steps(int reps, double* ms_primal, double* ms_dual);
};

}  // namespace cpdlp
