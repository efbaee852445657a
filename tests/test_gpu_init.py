"""Seeded films made on the device against the oracle's seeded films.

The headline workload and the accuracy tests start from a film generated on
the GPU: swalbe_init_random (h = base + amp N(0,1) or base + amp U(0,1) from
the counter-based Philox stream, the film that replaces randinterface!,
src/initialvalues.jl:23-33) and swalbe_init_height with INIT_NORMAL (the
initialiser of batch films).  Both must give the oracle's film
(orc_init_normal / orc_init_uniform) bit for bit in FP64.  An FP32 state
stores h - href in float32, so its stored plane must be the oracle's film
minus href rounded to float32 (round to nearest even), read back raw.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

SHAPES = [(64, 64), (130, 19), (7, 5), (1024, 768)]


@pytest.fixture(scope="module")
def S():
    import swalbe
    if swalbe.device_count() < 1:
        pytest.skip("no GPU")
    return swalbe


def _state(S, Lx, Ly, T, href):
    st = S.Sys(S.SysConst(Lx, Ly, S.Taumucs(Tmax=1)), "GPU", T=T)
    st.set_reference_height(href)
    return st


def _raw_height(st, Lx, Ly):
    from swalbe import _lib
    out = np.empty((Lx, Ly), dtype=np.float32, order="F")
    _lib.check(_lib.lib().swalbe_download_raw(st.handle, _lib.FIELDS["height"],
                                              out.ctypes.data, out.nbytes), "download_raw")
    return out


def _init_random(st, dist, seed, base, amp):
    from swalbe import _lib
    _lib.check(_lib.lib().swalbe_init_random(st.handle, dist, C.c_uint64(seed), float(base),
                                             float(amp)), "init_random")


@pytest.mark.parametrize("Lx,Ly", SHAPES)
@pytest.mark.parametrize("dist", ["normal", "uniform"])
def test_init_random_fp64_equals_oracle(S, Lx, Ly, dist):
    seed, base, amp = 1234 + Lx, 1.0, 0.001 if dist == "normal" else 0.01
    print(f"[init] init_random FP64 {dist} {Lx}x{Ly}", flush=True)
    st = _state(S, Lx, Ly, np.float64, 0.0)
    _init_random(st, 1 if dist == "normal" else 0, seed, base, amp)
    h = st.get("height")
    st.close()
    ref = (O.init_normal if dist == "normal" else O.init_uniform)(Lx, Ly, seed, base, amp)
    assert np.array_equal(h, ref)


@pytest.mark.parametrize("Lx,Ly", SHAPES)
def test_init_random_fp32_stores_the_rounded_deviation(S, Lx, Ly):
    seed, base, amp, href = 77, 1.0, 0.001, 1.0
    print(f"[init] init_random FP32 normal {Lx}x{Ly}", flush=True)
    st = _state(S, Lx, Ly, np.float32, href)
    _init_random(st, 1, seed, base, amp)
    raw = _raw_height(st, Lx, Ly)
    st.close()
    ref = O.init_normal(Lx, Ly, seed, base, amp)
    assert np.array_equal(raw, (ref - href).astype(np.float32))


@pytest.mark.parametrize("T", [np.float64, np.float32])
def test_film_init_normal_equals_oracle(S, T):
    Lx, Ly, seed, href = 200, 100, 11, 1.0
    print(f"[init] init_height INIT_NORMAL {np.dtype(T).name} {Lx}x{Ly}", flush=True)
    st = _state(S, Lx, Ly, T, href)
    st.init_height(S.film_init(S.INIT_NORMAL, seed=seed, base=1.0, amp=1e-3, href=href))
    ref = O.init_normal(Lx, Ly, seed, 1.0, 1e-3)
    if T == np.float64:
        h = st.get("height")
        st.close()
        # an FP64 state stores h itself (its reference height is not used)
        assert np.array_equal(h, ref)
    else:
        raw = _raw_height(st, Lx, Ly)
        st.close()
        assert np.array_equal(raw, (ref - href).astype(np.float32))


def test_seeds_and_bases_change_the_film(S):
    """Different seeds give different films; base and amp enter linearly
    (h - base = amp z with the same z for every amp)."""
    Lx, Ly = 96, 40
    st = _state(S, Lx, Ly, np.float64, 0.0)
    _init_random(st, 1, 5, 0.0, 1.0)
    z = st.get("height")
    _init_random(st, 1, 6, 0.0, 1.0)
    z6 = st.get("height")
    _init_random(st, 1, 5, 2.0, 0.25)
    h = st.get("height")
    st.close()
    assert not np.array_equal(z, z6)
    assert np.array_equal(h, 2.0 + 0.25 * z)
    assert abs(z.mean()) < 0.05 and abs(z.std() - 1.0) < 0.05
