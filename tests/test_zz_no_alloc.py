"""cpdlp_kkt_errors_device_noalloc, the allocation-free variant of cpdlp_kkt_errors
(kkt_errors!, src/components/errors.jl:65-109): called through the C ABI on device
vectors, it gives the same six KKT fields, bit for bit, as cpdlp_kkt_errors on the same
point, on two calls in a row (a larger LP in the sliced format, then a smaller one in
CSR) and again on the first LP after them (the shared scratch keeps the last call's
contents); it refuses an LP beyond its scratch and a handle whose preprocessing (which
allocates) has not run."""
import ctypes as C

import numpy as np
import pytest

from coolpdlp_hip import generators as G

pytestmark = pytest.mark.gpu

FIELDS = ("primal", "primal_scale", "dual", "dual_scale", "gap", "gap_scale")


@pytest.fixture(scope="module")
def cp():
    import coolpdlp_hip
    return coolpdlp_hip


def _noalloc(eng, xd, yd):
    from coolpdlp_hip import _lib
    e = _lib.CKKT()
    rc = _lib.lib.cpdlp_kkt_errors_device_noalloc(
        eng.h, C.c_void_p(xd.data_ptr()) if xd is not None else None,
        C.c_void_p(yd.data_ptr()) if yd is not None else None, C.byref(e))
    msg = _lib.lib.cpdlp_last_error().decode() if rc else ""
    return rc, e, msg


def _point(g, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(g.n), rng.standard_normal(g.m)


def _check(cp, eng, g, seed):
    import torch
    x, y = _point(g, seed)
    ref = eng.kkt_errors(cp.PrimalDualSolution(x, y))          # cpdlp_kkt_errors (host point)
    xd = torch.tensor(x, dtype=torch.float64, device="cuda:0")
    yd = torch.tensor(y, dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    rc, e, msg = _noalloc(eng, xd, yd)
    assert rc == 0, msg
    for f in FIELDS:
        assert getattr(e, f) == getattr(ref, f), (g.name, f, getattr(e, f), getattr(ref, f))
    assert np.isfinite([getattr(e, f) for f in FIELDS]).all()


def test_noalloc_kkt_equals_kkt_errors_on_two_calls(cp):
    big = G.power_law_lp(20_000, 60_000, 600_000, cap=8_000, seed=3)
    small = G.column_lp(300, 700, 6, seed=11)
    with cp.Engine(cp.MILP.from_generated(big), cp.PDLP("float64", None, "xslice", powm_tol=1e-6)) as e1, \
            cp.Engine(cp.MILP.from_generated(small), cp.PDLP("float64", None, "csr")) as e2:
        e1.prepare()
        e2.prepare()
        _check(cp, e1, big, 1)       # first call: the larger LP (column slices)
        _check(cp, e2, small, 2)     # second call: a smaller LP, the scratch holds the first's data
        _check(cp, e1, big, 3)       # the larger one again, over the smaller one's leftovers


def test_noalloc_kkt_refuses_what_it_cannot_do(cp):
    import torch
    from coolpdlp_hip import _lib
    g = G.column_lp(200, 500, 5, seed=12)
    with cp.Engine(cp.MILP.from_generated(g), cp.PDLP()) as eng:
        xd = torch.zeros(g.n, dtype=torch.float64, device="cuda:0")
        yd = torch.zeros(g.m, dtype=torch.float64, device="cuda:0")
        torch.cuda.synchronize()
        rc, _, msg = _noalloc(eng, xd, yd)
        assert rc == _lib.ERR_STATE and "cpdlp_prepare" in msg      # preprocessing allocates
        eng.prepare()
        rc, _, _ = _noalloc(eng, None, yd)
        assert rc == _lib.ERR_ARGUMENT
        _check(cp, eng, g, 4)
    # n + m beyond the 4,194,304-double scratch
    huge = G.column_lp(1000, 4_200_000, 1, seed=13)
    with cp.Engine(cp.MILP.from_generated(huge), cp.PDLP("float64", None, "csr", powm_tol=1e-3)) as eng:
        eng.prepare()
        xd = torch.zeros(huge.n, dtype=torch.float64, device="cuda:0")
        yd = torch.zeros(huge.m, dtype=torch.float64, device="cuda:0")
        torch.cuda.synchronize()
        rc, _, msg = _noalloc(eng, xd, yd)
        assert rc == _lib.ERR_ARGUMENT and "doubles of scratch" in msg, msg
        # the allocating entry point still takes it
        e = eng.kkt_errors(cp.PrimalDualSolution(np.zeros(huge.n), np.zeros(huge.m)))
        assert np.isfinite(e.primal_scale)
