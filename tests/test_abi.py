"""The C ABI: the library loads, exports every symbol include/coolpdlp_hip.h declares,
and its structs have the layout the Python binding assumes.  CPU only: no call here
touches the GPU (cpdlp_default_params and cpdlp_version are pure host functions)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "coolpdlp_hip.h")


def declared_symbols():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(cpdlp_\w+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from coolpdlp_hip import _lib
    syms = declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(_lib.lib, s), s
        assert s in _lib._SIGS, f"{s} has no ctypes signature"
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (cpdlp_\w+)", out))
    assert exported == set(syms)


def test_struct_layout_matches_c(tmp_path):
    from coolpdlp_hip import _lib
    prog = tmp_path / "layout.c"
    structs = {"cpdlp_lp": _lib.CLP, "cpdlp_params": _lib.CParams, "cpdlp_kkt": _lib.CKKT,
               "cpdlp_stats": _lib.CStats, "cpdlp_server_config": _lib.CServerConfig,
               "cpdlp_state": _lib.CState}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0;}")
    prog.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-o", str(exe), str(prog)], check=True)
    out = dict(ln.rsplit(" ", 1) for ln in subprocess.run([str(exe)], capture_output=True,
                                                          text=True).stdout.splitlines())
    for cname, py in structs.items():
        assert int(out[cname]) == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(out[f"{cname}.{f}"]) == getattr(py, f).offset, f"{cname}.{f}"


def test_default_params_mirror_the_reference():
    # Algorithm{A} defaults, src/algorithms/common.jl:53-78; PDHG, src/algorithms/pdhg.jl:6-16
    import coolpdlp_hip as cp
    p = cp.PDLP().params
    assert (p.ruiz_iter, p.chambolle_pock_alpha, p.invnorm_scaling) == (10, 1.0, 0.9)
    assert (p.primal_weight_damping, p.zero_tol) == (0.5, 1e-8)
    assert (p.sufficient_decay, p.necessary_decay, p.artificial_decay) == (0.2, 0.8, 0.36)
    assert (p.check_every, p.record_error_history, p.show_progress) == (100, 1, 0)
    assert (p.termination_reltol, p.max_kkt_passes, p.time_limit) == (1e-4, 10**5, 100.0)
    q = cp.PDHG().params
    assert q.ruiz_iter == 0 and np.isnan(q.primal_weight_damping)
    assert np.isnan(q.sufficient_decay) and np.isnan(q.necessary_decay) and np.isnan(q.artificial_decay)
    r = cp.PDLP("float32", termination_reltol=1e-6, check_every=64).params
    assert (r.precision, r.termination_reltol, r.check_every) == (1, 1e-6, 64)
    with pytest.raises(cp.ArgumentError):
        cp.PDLP(not_an_option=1)
    with pytest.raises(cp.ArgumentError):
        cp.PDLP(backend="cpu")


def test_milp_validation():
    # src/problems/milp.jl:71-76 (DimensionMismatch)
    import coolpdlp_hip as cp
    A = sp.csc_matrix(np.ones((2, 3)))
    ok = dict(c=np.ones(3), lv=np.zeros(3), uv=np.ones(3), A=A, lc=np.zeros(2), uc=np.ones(2))
    m = cp.MILP(**ok)
    assert (cp.nbvar(m), cp.nbcons(m), cp.nbcons_eq(m), cp.nbcons_ineq(m)) == (3, 2, 0, 2)
    for k, bad in [("c", np.ones(4)), ("lv", np.zeros(2)), ("uc", np.ones(3)), ("lc", np.zeros(1))]:
        with pytest.raises(cp.DimensionMismatch):
            cp.MILP(**{**ok, k: bad})
    m2 = cp.MILP(**ok, int_var=[True, False, False], path="/x/afiro.mps")
    assert cp.nbvar_int(m2) == 1 and cp.nbvar_cont(m2) == 2 and m2.name == "afiro"


def test_is_feasible_and_objective():
    # test/problems/solution.jl:8-23 (cube MILP)
    import coolpdlp_hip as cp
    m = cp.MILP(c=[1.0, 2.0], lv=np.zeros(2), uv=2 * np.ones(2), A=np.array([[1.0, 1.0]]),
                lc=[1.0], uc=[1.0], int_var=[True, False])
    assert cp.is_feasible([1.0, 0.0], m)
    with pytest.warns(UserWarning, match="Integrality"):
        assert not cp.is_feasible([0.5, 0.5], m)
    with pytest.warns(UserWarning, match="Constraints"):
        assert not cp.is_feasible([0.0, 0.0], m)
    with pytest.warns(UserWarning, match="Variable bounds"):
        assert not cp.is_feasible([2.0, -1.0], m)
    assert cp.objective_value([1.0, 1.0], m) == 3


def test_kkt_struct_helpers():
    import coolpdlp_hip as cp
    e = cp.KKTErrors(1.0, 2.0, 3.0, 4.0, 5.0, 10.0)
    assert e.relative() == 0.75
    assert e.absolute(2.0) == np.sqrt(4 + 9 / 4 + 25)


def test_server_config_defaults_and_validation():
    """cpdlp_server_default_config, and create-time validation that fails before any
    device work (batch route off, so no arena is allocated)."""
    from coolpdlp_hip import _lib
    cfg = _lib.CServerConfig()
    _lib.check(_lib.lib.cpdlp_server_default_config(C.byref(cfg)))
    assert (cfg.batch_bytes, cfg.plain_bytes) == (16 << 30, 64 << 30)
    assert (cfg.batch_limit, cfg.checks_per_round, cfg.n_lanes) == (8192, 20, 2)
    assert list(cfg.lane_workers)[:2] == [3, 1] and cfg.lane_max_nnz[0] == 1e6
    assert cfg.lane_max_nnz[1] == float("inf")
    assert (cfg.batch2_bytes, cfg.batch2_limit, cfg.batch2_checks) == (2 << 30, 0, 0)
    import coolpdlp_hip as cp
    for field, value in [("batch_limit", 70000), ("n_lanes", 0), ("n_lanes", 5), ("checks_per_round", 0),
                         ("batch2_limit", 70000), ("batch2_limit", -1), ("batch2_checks", -1)]:
        bad = _lib.CServerConfig()
        _lib.check(_lib.lib.cpdlp_server_default_config(C.byref(bad)))
        bad.batch_limit = bad.batch2_limit = 0
        setattr(bad, field, value)
        h = C.c_void_p()
        with pytest.raises(cp.ArgumentError):
            _lib.check(_lib.lib.cpdlp_server_create(C.byref(cp.PDLP().params), C.byref(bad), C.byref(h)))
