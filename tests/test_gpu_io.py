"""Snapshots taken inside the fused loop and exact restarts, on the GPU."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S():
    import swalbe
    if swalbe.device_count() < 1:
        pytest.skip("no GPU")
    return swalbe


def _state(S, T, tau, Lx=96, Ly=70):
    sysc = S.SysConst(Lx, Ly, S.Taumucs(Tmax=10, tdump=10 ** 9, kbt=1e-5, τ=tau))
    st = S.Sys(sysc, "GPU", T=T, kind="thermal")
    st.height = O.init_normal(Lx, Ly, 4, 1.0, 0.01)
    S.equilibrium(st, sysc)
    if tau != 1.0:
        st.copy("ftemp", "feq")
    return sysc, st


@pytest.mark.parametrize("T", [np.float64, np.float32], ids=["f64", "f32"])
def test_fused_snapshots_equal_downloads(S, T):
    sysc, a = _state(S, T, 1.0)
    snaps = {}
    sysc.param.Tmax = 50
    S.time_loop(sysc, a, thermal=True, seed=5, snapshots=snaps, snapshot_every=10)
    assert sorted(snaps) == sorted(f"h_{t}" for t in (10, 20, 30, 40, 50))
    sysc2, b = _state(S, T, 1.0)
    sysc2.param.Tmax = 10
    for t in (10, 20, 30, 40, 50):
        S.time_loop(sysc2, b, thermal=True, seed=5)
        assert np.array_equal(snaps[f"h_{t}"], b.get("height", np.float64)), t
    assert np.array_equal(a.get("height", np.float64), b.get("height", np.float64))


@pytest.mark.parametrize("T,tau", [(np.float64, 1.0), (np.float32, 1.0), (np.float32, 0.8)],
                         ids=["f64", "f32", "f32-tau0.8"])
def test_restart_from_saved_state_is_bit_exact(S, T, tau, tmp_path):
    sysc, a = _state(S, T, tau)
    sysc.param.Tmax = 40
    S.time_loop(sysc, a, thermal=True, seed=8)
    path = str(tmp_path / "state.npz")
    S.save_state(a, path)
    sysc.param.Tmax = 30
    S.time_loop(sysc, a, thermal=True, seed=8)
    b = S.load_state(path)
    assert b.t == 40
    S.time_loop(sysc, b, thermal=True, seed=8)
    for name in ("height", "velx", "vely", "pressure", "fout"):
        assert np.array_equal(a.get(name, np.float64), b.get(name, np.float64)), name


def test_state_on_a_caller_stream(S):
    """swalbe_set_stream: the same loop on a caller's HIP stream gives the same
    bits, and NULL restores the state's own stream."""
    import ctypes as C
    from swalbe import _lib
    hip = C.CDLL("libamdhip64.so")
    stream = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(stream)) == 0
    try:
        sysc, a = _state(S, np.float32, 1.0)
        sysc2, b = _state(S, np.float32, 1.0)
        sysc.param.Tmax = sysc2.param.Tmax = 30
        _lib.check(_lib.lib().swalbe_set_stream(b.handle, stream), "set_stream")
        S.time_loop(sysc, a, thermal=True, seed=4)
        S.time_loop(sysc2, b, thermal=True, seed=4)
        _lib.check(_lib.lib().swalbe_set_stream(b.handle, None), "set_stream")
        S.time_loop(sysc, a, thermal=True, seed=4)
        S.time_loop(sysc2, b, thermal=True, seed=4)
        assert np.array_equal(a.get("height"), b.get("height"))
        b.close()
    finally:
        assert hip.hipStreamDestroy(stream) == 0
