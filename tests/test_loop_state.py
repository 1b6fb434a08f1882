"""The loop state's host side (coolpdlp_hip.LoopState, CPU only): its ctypes mirror of
cpdlp_state carries every scalar, and a checkpoint written to disk reads back the same
bits (arrays and scalars, no pickled objects)."""
import numpy as np


def test_state_to_c_and_back():
    import coolpdlp_hip as cp
    from coolpdlp_hip import _lib
    rng = np.random.default_rng(3)
    st = cp.LoopState(rng.normal(size=(5, 7)), rng.normal(size=(5, 4)), rng.uniform(size=7), rng.uniform(size=4),
                      eta=0.37, omega=1.9, eta_sum=12.5, kkt_passes=700, inner=300, total=700, restarts=2,
                      fp32_passes=400)
    c = st.to_c()
    assert isinstance(c, _lib.CState)
    back = cp.LoopState.from_c(c, st.x, st.y, st.D2, st.D1)
    for f in ("eta", "omega", "eta_sum", "kkt_passes", "inner", "total", "restarts", "fp32_passes"):
        assert getattr(back, f) == getattr(st, f), f


def test_checkpoint_file_round_trip(tmp_path):
    import coolpdlp_hip as cp
    rng = np.random.default_rng(4)
    st = cp.LoopState(rng.normal(size=(5, 9)), rng.normal(size=(5, 3)), rng.uniform(size=9), rng.uniform(size=3),
                      eta=float(np.nextafter(0.5, 1.0)), omega=3.0, eta_sum=0.0, kkt_passes=2**40, inner=0, total=5,
                      restarts=7, fp32_passes=0)
    path = tmp_path / "ckpt.npz"
    st.save(path)
    got = cp.LoopState.load(path)
    for f in ("x", "y", "D2", "D1"):
        assert np.array_equal(getattr(got, f), getattr(st, f)) and getattr(got, f).dtype == np.float64, f
    for f in ("eta", "omega", "eta_sum", "kkt_passes", "inner", "total", "restarts", "fp32_passes"):
        assert getattr(got, f) == getattr(st, f) and type(getattr(got, f)) is type(getattr(st, f)), f
