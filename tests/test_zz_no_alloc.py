"""GPU test of swalbe_measure_ptr_noalloc, the measurement entry point that
allocates no device memory: through the C ABI, its results equal those of
swalbe_measure_ptr (the same kernels, partials in the library's per-GPU
scratch array instead of a fresh allocation) bit for bit, on two calls in a
row, the second with a smaller film, for FP64 and FP32 heights, with a θ
field and a mask."""
import ctypes as C

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S():
    import swalbe as S
    if S.device_count() == 0:
        pytest.skip("no GPU")
    return S


def _measure(S, fn, T, h, th, thresh):
    """Runs `fn` (an entry point's name) on device copies of h and θ and
    returns (out[4], mask) downloaded after a stream synchronisation."""
    from swalbe import _lib
    from swalbe.core import DeviceArray
    L = _lib.lib()
    Lx, Ly = h.shape
    dt = _lib.F64 if T is np.float64 else _lib.F32
    dh = DeviceArray(h.shape, T, h.astype(T))
    dth = DeviceArray(th.shape, T, th.astype(T)) if th is not None else None
    dmask = DeviceArray((Lx * Ly,), np.uint8, np.full(Lx * Ly, 7, np.uint8))
    dout = DeviceArray((4,), np.float64, np.full(4, np.nan))
    _lib.check(getattr(L, fn)(dt, dh.ptr, dth.ptr if dth else None, 1 / 9, Lx, Ly, thresh,
                              dmask.ptr, C.cast(dout.ptr, C.POINTER(C.c_double)), None), fn)
    # the no-allocation form only enqueues on the null stream; the blocking
    # copy (hipMemcpy on the null stream) runs after it
    out = dout.get()
    mask = dmask.get().reshape((Lx, Ly), order="F")
    return out, mask


@pytest.mark.parametrize("T", [np.float64, np.float32])
def test_noalloc_measure_equals_the_allocating_entry_point(S, T):
    rng = np.random.default_rng(11)
    shapes = [(300, 170), (37, 23)]  # second call smaller: the scratch keeps the first's partials
    for k, (Lx, Ly) in enumerate(shapes):
        h = 0.05 + np.abs(rng.standard_normal((Lx, Ly))) * 0.3
        th = rng.uniform(1 / 18, 1 / 6, (Lx, Ly)) if k == 0 else None
        print(f"[no_alloc] T={T.__name__} shape={Lx}x{Ly} theta_field={th is not None}",
              flush=True)
        want, wmask = _measure(S, "swalbe_measure_ptr", T, h, th, 0.055)
        got, gmask = _measure(S, "swalbe_measure_ptr_noalloc", T, h, th, 0.055)
        assert np.array_equal(got, want), (got, want)
        assert np.array_equal(gmask, wmask)
        # and against the oracle (FP64 exactly for count and max)
        if T is np.float64:
            ref, rmask = (O.measure(h, th, 0.0, 0.055) if th is not None
                          else O.measure(h, None, 1 / 9, 0.055))
            assert got[0] == ref["count"] and got[1] == ref["max"]
            assert abs(got[2] - ref["area"]) <= 1e-12 * ref["area"]
            assert np.array_equal(gmask, rmask)


def test_noalloc_measure_rejects_bad_arguments(S):
    from swalbe import _lib
    rc = _lib.lib().swalbe_measure_ptr_noalloc(_lib.F64, None, None, 0.0, 4, 4, 0.0, None,
                                               None, None)
    assert rc == _lib.EINVAL
