"""ctypes binding of libcoolpdlp_hip.so (include/coolpdlp_hip.h).

The library is required: importing this module raises if the built .so is absent
(no CPU fallback exists).  Build it with ``python -m coolpdlp_hip._build`` or
``__graft_entry__.build()``.
"""
from __future__ import annotations

import ctypes as C
import os

from ._build import LIB

CPDLP_OK = 0
ERR_DIMENSION, ERR_ARGUMENT, ERR_MEMORY, ERR_HIP, ERR_STATE = 1, 2, 3, 4, 5

ALGO_PDLP, ALGO_PDHG = 0, 1
PREC_FP64, PREC_FP32, PREC_MIXED = 0, 1, 2
FMT_AUTO, FMT_CSR, FMT_ELL, FMT_XSLICE, FMT_XSLICE_ALL = 0, 1, 2, 3, 4


class DimensionMismatch(ValueError):
    """Julia's DimensionMismatch (src/problems/milp.jl:71-76)."""


class ArgumentError(ValueError):
    """Julia's ArgumentError (src/problems/milp.jl:83-92, src/utils/device.jl:6-12)."""


class HIPError(RuntimeError):
    pass


class CLP(C.Structure):
    _fields_ = [("m", C.c_int64), ("n", C.c_int64), ("index_base", C.c_int32),
                ("index_bits", C.c_int32), ("colptr", C.c_void_p), ("rowval", C.c_void_p),
                ("nzval", C.c_void_p), ("c", C.c_void_p), ("lv", C.c_void_p),
                ("uv", C.c_void_p), ("lc", C.c_void_p), ("uc", C.c_void_p)]


class CParams(C.Structure):
    _fields_ = [
        ("algorithm", C.c_int32), ("ruiz_iter", C.c_int32),
        ("chambolle_pock_alpha", C.c_double), ("invnorm_scaling", C.c_double),
        ("primal_weight_damping", C.c_double), ("zero_tol", C.c_double),
        ("sufficient_decay", C.c_double), ("necessary_decay", C.c_double),
        ("artificial_decay", C.c_double), ("show_progress", C.c_int32),
        ("check_every", C.c_int32), ("record_error_history", C.c_int32),
        ("precision", C.c_int32), ("termination_reltol", C.c_double),
        ("max_kkt_passes", C.c_int64), ("time_limit", C.c_double),
        ("format", C.c_int32), ("device", C.c_int32), ("powm_tol", C.c_double),
        ("powm_maxiter", C.c_int64), ("powm_seed", C.c_uint64),
        ("polish_reltol", C.c_double), ("use_graph", C.c_int32), ("long_row", C.c_int32),
        ("batch_streams", C.c_int32), ("detect_infeasibility", C.c_int32),
        ("infeasibility_tol", C.c_double),
    ]


SERVER_MAX_LANES = 4


class CServerConfig(C.Structure):
    _fields_ = [
        ("batch_bytes", C.c_int64), ("plain_bytes", C.c_int64), ("batch_limit", C.c_int32),
        ("checks_per_round", C.c_int32), ("n_lanes", C.c_int32),
        ("lane_workers", C.c_int32 * SERVER_MAX_LANES), ("lane_max_nnz", C.c_double * SERVER_MAX_LANES),
        ("batch2_bytes", C.c_int64), ("batch2_limit", C.c_int32), ("batch2_checks", C.c_int32),
    ]


class CKKT(C.Structure):
    _fields_ = [(f, C.c_double) for f in
                ("primal", "primal_scale", "dual", "dual_scale", "gap", "gap_scale")]


class CStats(C.Structure):
    _fields_ = [("err", CKKT), ("time_elapsed", C.c_double), ("kkt_passes", C.c_int64),
                ("termination_status", C.c_int32), ("format_A", C.c_int32),
                ("format_At", C.c_int32), ("line_tiles", C.c_int32), ("eta", C.c_double),
                ("omega", C.c_double), ("powm_iters", C.c_int64), ("restarts", C.c_int64),
                ("n_history", C.c_int64), ("fp32_passes", C.c_int64),
                ("setup_seconds", C.c_double), ("loop_seconds", C.c_double),
                ("step_kernel_ms", C.c_double), ("step_kernel_launches", C.c_int64)]


STATE_VECTORS = 5


class CState(C.Structure):
    _fields_ = [("eta", C.c_double), ("omega", C.c_double), ("eta_sum", C.c_double),
                ("kkt_passes", C.c_int64), ("inner", C.c_int64), ("total", C.c_int64),
                ("restarts", C.c_int64), ("fp32_passes", C.c_int64)]


# every symbol the header declares, with its signature
_P = C.POINTER
_SIGS = {
    "cpdlp_default_params": (C.c_int, [_P(CParams), C.c_int32]),
    "cpdlp_create": (C.c_int, [_P(CLP), _P(CParams), _P(C.c_void_p)]),
    "cpdlp_solve": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "cpdlp_initialize": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "cpdlp_run_checks": (C.c_int, [C.c_void_p, C.c_int64, _P(C.c_int32)]),
    "cpdlp_get_state": (C.c_int, [C.c_void_p, _P(CState), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "cpdlp_resume": (C.c_int, [C.c_void_p, _P(CState), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "cpdlp_get_solution_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "cpdlp_get_solution": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "cpdlp_get_stats": (C.c_int, [C.c_void_p, _P(CStats)]),
    "cpdlp_get_history": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                    _P(C.c_int64)]),
    "cpdlp_kkt_errors": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, _P(CKKT)]),
    "cpdlp_kkt_errors_device_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, _P(CKKT)]),
    "cpdlp_prepare": (C.c_int, [C.c_void_p]),
    "cpdlp_get_scaling": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p]),
    "cpdlp_get_csr": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "cpdlp_get_step_size": (C.c_int, [C.c_void_p, _P(C.c_double), _P(C.c_double),
                                      _P(C.c_int64)]),
    "cpdlp_profile_steps": (C.c_int, [C.c_void_p, C.c_int32, _P(C.c_double), _P(C.c_double)]),
    "cpdlp_destroy": (None, [C.c_void_p]),
    "cpdlp_last_error": (C.c_char_p, []),
    "cpdlp_version": (C.c_char_p, []),
    "cpdlp_batch_create": (C.c_int, [_P(CParams), C.c_int64, _P(C.c_void_p)]),
    "cpdlp_batch_add": (C.c_int, [C.c_void_p, _P(CLP), _P(C.c_int64)]),
    "cpdlp_batch_add_start": (C.c_int, [C.c_void_p, _P(CLP), C.c_void_p, C.c_void_p, _P(C.c_int64)]),
    "cpdlp_server_default_config": (C.c_int, [_P(CServerConfig)]),
    "cpdlp_server_create": (C.c_int, [_P(CParams), _P(CServerConfig), _P(C.c_void_p)]),
    "cpdlp_server_submit": (C.c_int, [C.c_void_p, _P(CLP), C.c_void_p, C.c_void_p, _P(C.c_int64)]),
    "cpdlp_server_submit_borrowed": (C.c_int, [C.c_void_p, _P(CLP), C.c_void_p, C.c_void_p, _P(C.c_int64)]),
    "cpdlp_server_pump": (C.c_int, [C.c_void_p, _P(C.c_int64)]),
    "cpdlp_server_start": (C.c_int, [C.c_void_p]),
    "cpdlp_server_poll": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, _P(C.c_int64)]),
    "cpdlp_server_get": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, _P(CStats),
                                   _P(C.c_double), _P(C.c_int32)]),
    "cpdlp_server_destroy": (None, [C.c_void_p]),
    "cpdlp_batch_footprint": (C.c_int64, [C.c_int64, C.c_int64, C.c_int64]),
    "cpdlp_batch_run": (C.c_int, [C.c_void_p, C.c_int32, _P(C.c_int64)]),
    "cpdlp_batch_poll": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, _P(C.c_int64)]),
    "cpdlp_batch_get": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, _P(CStats)]),
    "cpdlp_batch_release": (C.c_int, [C.c_void_p, C.c_int64]),
    "cpdlp_batch_take_state": (C.c_int, [C.c_void_p, C.c_int64, _P(CState), C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_void_p]),
    "cpdlp_batch_memory": (C.c_int, [C.c_void_p, _P(C.c_int64), _P(C.c_int64)]),
    "cpdlp_batch_destroy": (None, [C.c_void_p]),
    "cpdlp_matrix_create": (C.c_int, [C.c_int64, C.c_int64, C.c_int32, C.c_int32, C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32,
                                      _P(C.c_void_p)]),
    "cpdlp_matrix_mul": (C.c_int, [C.c_void_p, C.c_double, C.c_void_p, C.c_double, C.c_void_p,
                                   C.c_void_p]),
    "cpdlp_matrix_mul_host": (C.c_int, [C.c_void_p, C.c_double, C.c_void_p, C.c_double,
                                        C.c_void_p]),
    "cpdlp_matrix_info": (C.c_int, [C.c_void_p, _P(C.c_int64), _P(C.c_int64), _P(C.c_int64),
                                    _P(C.c_int32), _P(C.c_int64)]),
    "cpdlp_matrix_get_ell": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "cpdlp_matrix_destroy": (None, [C.c_void_p]),
}

if not os.path.exists(LIB):
    raise ImportError(
        f"libcoolpdlp_hip.so not found at {LIB}: the HIP engine is required (there is no "
        "CPU fallback); build it with __graft_entry__.build()")

# One HIP runtime per process: PyTorch ships its own libamdhip64 (same SONAME as
# ROCm's).  Loading torch first makes this library bind to that copy, so device
# pointers and streams from torch tensors are valid here; loading ours first would
# put two runtimes in the process and torch's would then fail to initialise.
if os.environ.get("COOLPDLP_NO_TORCH") != "1":
    try:
        import torch  # noqa: F401
    except ImportError:
        pass

lib = C.CDLL(LIB)
for _name, (_res, _args) in _SIGS.items():
    _f = getattr(lib, _name)
    _f.restype = _res
    _f.argtypes = _args


def check(code: int):
    if code == CPDLP_OK:
        return
    msg = lib.cpdlp_last_error().decode(errors="replace")
    if code == ERR_DIMENSION:
        raise DimensionMismatch(msg)
    if code == ERR_ARGUMENT:
        raise ArgumentError(msg)
    if code == ERR_MEMORY:
        raise MemoryError(msg)
    raise HIPError(f"[{code}] {msg}")


def default_params(algorithm: int) -> CParams:
    p = CParams()
    check(lib.cpdlp_default_params(C.byref(p), algorithm))
    return p
