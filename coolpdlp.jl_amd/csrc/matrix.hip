// matrix.hip -- the sparse-matrix plug-in of the C ABI: GPUSparseMatrixCSR /
// GPUSparseMatrixELL construction from a SparseMatrixCSC and the 5-argument mul!
// (src/utils/mat_csr.jl:58-92, src/utils/mat_ell.jl:54-98) on device pointers.
#include <memory>

#include "kernels.hpp"
#include "layout.hpp"
#include "solver.hpp"

using namespace cpdlp;

namespace cpdlp {
template <class T>
__global__ void k_gather_vals_m(int64_t nnz, const int32_t* vmap, const double* src, T* out) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nnz) return;
    const int32_t s = vmap[k];
    out[k] = s >= 0 ? (T)src[s] : T(0);
}
__global__ void k_fill_m(int64_t n, double* v, double a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = a;
}
}  // namespace cpdlp

struct cpdlp_matrix {
    int64_t m = 0, n = 0, nnz = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    DevSide S;
    DevBuf<int32_t> out_index;
    std::vector<int32_t> row_of_int, int_of_row;
    std::vector<int32_t> ell_col_orig;
    std::vector<int32_t> ell_vmap_orig;
    std::vector<double> nzval;

    ~cpdlp_matrix() {
        (void)hipSetDevice(device);
        // products may still be queued on the caller's streams: wait for the device
        // before the buffers go back to the block cache
        (void)hipDeviceSynchronize();
        if (stream) (void)hipStreamDestroy(stream);
    }

    void launch(double alpha, const double* b, double beta, double* c, hipStream_t s) {
        EpiAxpby e;
        e.out = c;
        e.out_index = out_index.p;
        e.alpha = alpha;
        e.beta = beta;
        launch_spmv<double, double, 1>(S.view<double>(), XPack<double, 1>{{b}}, e, s);
        HIP_CHECK(hipGetLastError());
    }
};

extern "C" {

int cpdlp_matrix_create(int64_t m, int64_t n, int32_t index_base, int32_t index_bits,
                        const void* colptr, const void* rowval, const double* nzval,
                        int32_t format, int32_t transpose, int32_t device, cpdlp_matrix** out) {
    return guarded([&] {
        if (!out) throw Error(CPDLP_ERR_ARGUMENT, "NULL output");
        if (format < CPDLP_FMT_AUTO || format > CPDLP_FMT_ELL)
            throw Error(CPDLP_ERR_ARGUMENT, "unknown format");
        std::vector<int32_t> cp, rv;
        read_csc(m, n, index_base, index_bits, colptr, rowval, cp, rv);
        const int64_t nnz = cp[n];
        if (nnz > 0 && !nzval) throw Error(CPDLP_ERR_ARGUMENT, "nzval is NULL");
        int ndev = 0;
        HIP_CHECK(hipGetDeviceCount(&ndev));
        if (device < 0 || device >= ndev) throw Error(CPDLP_ERR_ARGUMENT, "invalid device ordinal");
        HIP_CHECK(hipSetDevice(device));
        auto M = std::make_unique<cpdlp_matrix>();
        M->device = device;
        HIP_CHECK(hipStreamCreateWithFlags(&M->stream, hipStreamNonBlocking));
        M->nnz = nnz;
        // rows of the stored matrix: CSR(A) (transpose = 0) or CSR(A') = CSC(A) arrays
        std::vector<int32_t> rp, col, perm;
        int32_t rows, cols;
        if (transpose) {
            rows = (int32_t)n;
            cols = (int32_t)m;
            rp = cp;
            col = rv;
        } else {
            rows = (int32_t)m;
            cols = (int32_t)n;
            csc_to_csr(m, n, cp, rv, rp, col, perm);
        }
        M->m = rows;
        M->n = cols;
        int32_t n_long = 0;
        partition_rows(rp, 64, M->row_of_int, M->int_of_row, n_long);
        HostSide H;
        build_side(H, rows, cols, rp, col, transpose ? nullptr : &perm, M->row_of_int, nullptr, n_long, n_long);
        M->S.upload_structure(H, M->stream);
        // an explicit ELL is built whatever its padding (as GPUSparseMatrixELL does);
        // AUTO only considers it within 1.5x nnz
        if (format != CPDLP_FMT_CSR) M->S.build_ell(H, M->stream, format == CPDLP_FMT_ELL ? 1e30 : 1.5);
        DevBuf<double> src(std::max<int64_t>(nnz, 1));
        src.upload(nzval, nnz, M->stream);
        M->S.val64.alloc(std::max<int64_t>(nnz, 1));
        if (nnz)
            hipLaunchKernelGGL(k_gather_vals_m<double>, dim3(ceil_div(nnz, BLOCK)), dim3(BLOCK), 0,
                               M->stream, nnz, M->S.vmap.p, src.p, M->S.val64.p);
        if (M->S.ell_w > 0) {
            const int64_t len = (int64_t)M->S.ell_w * rows;
            M->S.ell_val64.alloc(len);
            hipLaunchKernelGGL(k_gather_vals_m<double>, dim3(ceil_div(len, BLOCK)), dim3(BLOCK), 0,
                               M->stream, len, M->S.ell_vmap.p, src.p, M->S.ell_val64.p);
            // ELL in the original row order for cpdlp_matrix_get_ell
            std::vector<int32_t> ecol, evmap;
            cpdlp::build_ell(H, 1e30, ecol, evmap);
            M->ell_col_orig.assign(ecol.size(), -1);
            M->ell_vmap_orig.assign(ecol.size(), -1);
            for (int32_t i = 0; i < rows; ++i)
                for (int32_t s = 0; s < M->S.ell_w; ++s) {
                    const size_t src_i = (size_t)s * rows + M->int_of_row[i];
                    const size_t dst_i = (size_t)s * rows + i;
                    M->ell_col_orig[dst_i] = ecol[src_i];
                    M->ell_vmap_orig[dst_i] = evmap[src_i];
                }
            M->nzval.assign(nzval, nzval + nnz);
        }
        M->out_index.upload(M->row_of_int, M->stream);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipStreamSynchronize(M->stream));
        // format choice
        M->S.format = CPDLP_FMT_CSR;
        if (format == CPDLP_FMT_ELL && M->S.ell_w > 0) M->S.format = CPDLP_FMT_ELL;
        if (format == CPDLP_FMT_AUTO && M->S.ell_w > 0) {
            DevBuf<double> b(std::max<int64_t>(cols, 1)), c(std::max<int64_t>(rows, 1));
            hipLaunchKernelGGL(k_fill_m, dim3(std::max(1, ceil_div(cols, BLOCK))), dim3(BLOCK), 0,
                               M->stream, (int64_t)cols, b.p, 1.0);
            hipEvent_t e0, e1;
            HIP_CHECK(hipEventCreate(&e0));
            HIP_CHECK(hipEventCreate(&e1));
            float t[2];
            for (int f = 0; f < 2; ++f) {
                M->S.format = f ? CPDLP_FMT_ELL : CPDLP_FMT_CSR;
                for (int r = 0; r < 3; ++r) M->launch(1.0, b.p, 0.0, c.p, M->stream);
                HIP_CHECK(hipEventRecord(e0, M->stream));
                for (int r = 0; r < 10; ++r) M->launch(1.0, b.p, 0.0, c.p, M->stream);
                HIP_CHECK(hipEventRecord(e1, M->stream));
                HIP_CHECK(hipEventSynchronize(e1));
                HIP_CHECK(hipEventElapsedTime(&t[f], e0, e1));
            }
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            M->S.format = t[1] < t[0] ? CPDLP_FMT_ELL : CPDLP_FMT_CSR;
        }
        *out = M.release();
    });
}

int cpdlp_matrix_mul(cpdlp_matrix* A, double alpha, const double* b_dev, double beta,
                     double* c_dev, void* stream) {
    return guarded([&] {
        if (!A) throw Error(CPDLP_ERR_ARGUMENT, "NULL matrix");
        if ((!b_dev && A->n) || (!c_dev && A->m)) throw Error(CPDLP_ERR_ARGUMENT, "NULL vector");
        HIP_CHECK(hipSetDevice(A->device));
        // NULL is HIP's null stream (a caller's default stream), not the matrix's own: the
        // product must be ordered with the caller's work on b and c
        A->launch(alpha, b_dev, beta, c_dev, (hipStream_t)stream);
    });
}

int cpdlp_matrix_mul_host(cpdlp_matrix* A, double alpha, const double* b, double beta, double* c) {
    return guarded([&] {
        if (!A) throw Error(CPDLP_ERR_ARGUMENT, "NULL matrix");
        if ((!b && A->n) || (!c && A->m)) throw Error(CPDLP_ERR_ARGUMENT, "NULL vector");
        HIP_CHECK(hipSetDevice(A->device));
        DevBuf<double> bd(std::max<int64_t>(A->n, 1)), cd(std::max<int64_t>(A->m, 1));
        bd.upload(b, A->n, A->stream);
        if (beta != 0.0) cd.upload(c, A->m, A->stream);
        A->launch(alpha, bd.p, beta, cd.p, A->stream);
        cd.download(c, A->m, A->stream);
        HIP_CHECK(hipStreamSynchronize(A->stream));
    });
}

int cpdlp_matrix_info(cpdlp_matrix* A, int64_t* m, int64_t* n, int64_t* nnz, int32_t* format,
                      int64_t* ell_width) {
    return guarded([&] {
        if (!A) throw Error(CPDLP_ERR_ARGUMENT, "NULL matrix");
        if (m) *m = A->m;
        if (n) *n = A->n;
        if (nnz) *nnz = A->nnz;
        if (format) *format = A->S.format;
        if (ell_width) *ell_width = A->S.ell_w;
    });
}

int cpdlp_matrix_get_ell(cpdlp_matrix* A, int32_t* colval, double* nzval) {
    return guarded([&] {
        if (!A) throw Error(CPDLP_ERR_ARGUMENT, "NULL matrix");
        if (A->S.ell_w <= 0) throw Error(CPDLP_ERR_STATE, "matrix has no ELL form");
        const size_t len = A->ell_col_orig.size();
        for (size_t t = 0; t < len; ++t) {
            if (colval) colval[t] = A->ell_col_orig[t];
            if (nzval) nzval[t] = A->ell_vmap_orig[t] >= 0 ? A->nzval[A->ell_vmap_orig[t]] : 0.0;
        }
    });
}

void cpdlp_matrix_destroy(cpdlp_matrix* A) { delete A; }

}  // extern "C"
