"""Sparse-matrix plug-ins: GPUSparseMatrixCSR / GPUSparseMatrixELL + mul!.

Mirrors src/utils/mat_csr.jl:8-92 and src/utils/mat_ell.jl:8-102: construction from
a SparseMatrixCSC (here: any scipy.sparse matrix) and the 5-argument
``mul!(c, A, b, alpha, beta)`` (c = alpha*A*b + beta*c; beta == 0 never reads c).
The matrix lives in HBM; ``mul`` works on torch device tensors (zero copy) or host
numpy arrays.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import scipy.sparse as sp

from . import _lib
from ._lib import ArgumentError, check, lib


class _GPUSparseMatrix:
    _format = _lib.FMT_CSR

    def __init__(self, A, *, transpose=False, device=0, format=None):
        A = sp.csc_matrix(A)
        A.sort_indices()
        self._cp = np.ascontiguousarray(A.indptr, dtype=np.int64)
        self._rv = np.ascontiguousarray(A.indices, dtype=np.int64)
        self._nz = np.ascontiguousarray(A.data, dtype=np.float64)
        h = C.c_void_p()
        fmt = self._format if format is None else format
        check(lib.cpdlp_matrix_create(A.shape[0], A.shape[1], 0, 64,
                                      self._cp.ctypes.data_as(C.c_void_p),
                                      self._rv.ctypes.data_as(C.c_void_p),
                                      self._nz.ctypes.data_as(C.c_void_p), fmt,
                                      1 if transpose else 0, device, C.byref(h)))
        self.h = h
        m, n, nnz, f, w = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int32(), C.c_int64()
        check(lib.cpdlp_matrix_info(h, C.byref(m), C.byref(n), C.byref(nnz), C.byref(f),
                                    C.byref(w)))
        self.shape = (m.value, n.value)
        self._nnz = nnz.value
        self.format = f.value
        self.ell_width = w.value

    def close(self):
        if getattr(self, "h", None):
            lib.cpdlp_matrix_destroy(self.h)
            self.h = None

    __del__ = close

    @property
    def nnz(self):
        return self._nnz

    def size(self, dim=None):
        return self.shape if dim is None else self.shape[dim - 1]

    def mul(self, c, b, alpha=1.0, beta=0.0, stream=None):
        """mul!(c, A, b, alpha, beta); c and b: torch tensors on the GPU (float64,
        contiguous) or numpy float64 arrays.  Returns c."""
        if hasattr(b, "data_ptr"):
            import torch
            if b.dtype != torch.float64 or c.dtype != torch.float64:
                raise ArgumentError("float64 tensors required")
            if not (b.is_cuda and c.is_cuda):
                raise ArgumentError("There are several different backends among the arguments")
            if b.numel() != self.shape[1] or c.numel() != self.shape[0]:
                raise _lib.DimensionMismatch("vector sizes do not match the matrix")
            s = stream if stream is not None else torch.cuda.current_stream(b.device).cuda_stream
            check(lib.cpdlp_matrix_mul(self.h, float(alpha), C.c_void_p(b.data_ptr()),
                                       float(beta), C.c_void_p(c.data_ptr()), C.c_void_p(s)))
            return c
        if hasattr(c, "data_ptr"):
            raise ArgumentError("There are several different backends among the arguments")
        b = np.ascontiguousarray(b, dtype=np.float64)
        if c.dtype != np.float64 or not c.flags.c_contiguous:
            raise ArgumentError("c must be a contiguous float64 array")
        if len(b) != self.shape[1] or len(c) != self.shape[0]:
            raise _lib.DimensionMismatch("vector sizes do not match the matrix")
        check(lib.cpdlp_matrix_mul_host(self.h, float(alpha), b.ctypes.data_as(C.c_void_p),
                                        float(beta), c.ctypes.data_as(C.c_void_p)))
        return c

    def __matmul__(self, b):
        c = np.zeros(self.shape[0])
        return self.mul(c, b)

    def ell(self):
        """(colval, nzval) of the column-major ELL form, shape (d, m) in numpy order
        (entry s of row i at [s, i]); padding column -1, value 0."""
        if self.ell_width <= 0:
            raise ArgumentError("no ELL form")
        d, m = self.ell_width, self.shape[0]
        col = np.zeros(d * m, np.int32)
        val = np.zeros(d * m)
        check(lib.cpdlp_matrix_get_ell(self.h, col.ctypes.data_as(C.c_void_p),
                                       val.ctypes.data_as(C.c_void_p)))
        return col.reshape(d, m), val.reshape(d, m)


class GPUSparseMatrixCSR(_GPUSparseMatrix):
    """GPUSparseMatrixCSR(A::SparseMatrixCSC), src/utils/mat_csr.jl:58-61"""
    _format = _lib.FMT_CSR


class GPUSparseMatrixELL(_GPUSparseMatrix):
    """GPUSparseMatrixELL(A::SparseMatrixCSC), src/utils/mat_ell.jl:54-70"""
    _format = _lib.FMT_ELL


def format_of(matrix_type) -> int:
    """Matrix type of Algorithm(T, Ti, M): None / "auto" lets the engine's autotuner pick
    per matrix (CSR, ELL, XCD column slices of the long rows or of every row); the others
    force one format."""
    if matrix_type is None or matrix_type in ("auto", "AUTO"):
        return _lib.FMT_AUTO
    if matrix_type in ("csr", "CSR"):
        return _lib.FMT_CSR
    if matrix_type in ("ell", "ELL"):
        return _lib.FMT_ELL
    if matrix_type in ("xslice", "XSLICE"):
        return _lib.FMT_XSLICE
    if matrix_type in ("xslice_all", "XSLICE_ALL"):
        return _lib.FMT_XSLICE_ALL
    if isinstance(matrix_type, type) and issubclass(matrix_type, _GPUSparseMatrix):
        return matrix_type._format
    if matrix_type in ("SparseMatrixCSC",):
        return _lib.FMT_CSR
    raise ArgumentError(f"unsupported matrix type {matrix_type!r}")
