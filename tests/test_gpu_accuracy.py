"""Accuracy of the FP32 path at full size: the headline rupture film
(4096^2, thermal noise, 10^4 steps) and the sessile droplet (1024^2,
5 10^4 steps), FP32 against the FP64 path from the same initial height.
The FP64 path is the oracle's arithmetic bit for bit (test_gpu_parity.py),
so it stands in for the oracle at sizes the CPU cannot run."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S():
    import swalbe
    if swalbe.device_count() < 1:
        pytest.skip("no GPU")
    return swalbe


def relL2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def run(S, Lx, Ly, T, param, init, nsteps, thermal, seed, variant=0):
    from swalbe import _lib
    sysc = S.SysConst(Lx, Ly, param)
    st = S.Sys(sysc, "GPU", T=T, kind="thermal" if thermal else "simple")
    st.set_reference_height(init.href)
    st.init_height(init)
    S.equilibrium(st, sysc)
    m0 = st.reduce()["sum"]
    o = _lib.StepOpts()
    o.thermal, o.seed, o.variant = int(thermal), seed, variant
    _lib.check(_lib.lib().swalbe_time_loop(st.handle, C.byref(param.cparams()), nsteps,
                                           C.byref(o)), "time_loop")
    h = st.get("height", np.float64)
    ux = st.get("velx", np.float64)
    m1 = st.reduce()["sum"]
    st.close()
    return h, ux, m0, m1


def test_rupture_4096_fp32_vs_fp64_after_1e4_steps(S):
    """4096^2, h0 = 1 + 0.001 N(0,1), kbt = 1e-6 noise, θ = 1/9, n = 9, m = 3,
    γ = 0.01: rel-L2(h) of FP32 vs FP64 after 10^4 steps <= 1e-5."""
    Lx = Ly = 4096
    nsteps = 10_000
    prm = S.Taumucs(Tmax=nsteps, tdump=10 ** 9, kbt=1e-6, γ=0.01, δ=1.0, n=9, m=3, hmin=0.1,
                    hcrit=0.05, θ=1 / 9)
    init = S.film_init(S.INIT_NORMAL, seed=1234, base=1.0, amp=0.001, href=1.0)
    print("[accuracy] rupture 4096^2 FP64 ...", flush=True)
    h64, u64, m064, m164 = run(S, Lx, Ly, np.float64, prm, init, nsteps, True, 7)
    print("[accuracy] rupture 4096^2 FP32 ...", flush=True)
    h32, u32, m032, m132 = run(S, Lx, Ly, np.float32, prm, init, nsteps, True, 7)
    e = relL2(h32, h64)
    dyn = relL2(h32 - h32.mean(), h64 - h64.mean())
    print(f"[accuracy] rupture 4096^2 10^4 steps: rel-L2 h = {e:.3e}, dynamics {dyn:.3e}, "
          f"ux {relL2(u32, u64):.3e}; mass drift FP64 {abs(m164 - m064) / m064:.2e}, "
          f"FP32 {abs(m132 - m032) / m032:.2e}; h range FP64 [{h64.min():.4f}, {h64.max():.4f}]",
          flush=True)
    assert np.isfinite(h32).all()
    assert e <= 1e-5
    assert abs(m164 - m064) <= 1e-10 * m064
    assert abs(m132 - m064) <= 1e-6 * m064


def test_droplet_1024_fp32_and_fp64_5e4_steps(S):
    """1024^2 sessile droplet, cap radius 150 at the centre, θ = 1/9 (cap and
    substrate), precursor 0.1, γ = 0.01, 5 10^4 steps.  FP32 tracks FP64 to
    rel-L2(h) <= 1e-5 over the first 10^3 steps.  Later the film amplifies
    small differences exponentially (FP32 vs FP64 grows ~x20 per 10^3 steps
    up to O(0.1) by 2 10^4 steps, in both storage offsets), so at 5 10^4 steps
    the bounds are on conserved and bulk quantities: FP64 mass to rounding,
    FP32 mass to 1e-4, the droplet's height to 1e-3 and its volume above the
    precursor to 1e-2; the pointwise rel-L2 is reported."""
    from swalbe import _lib
    Lx = Ly = 1024
    prm = S.Taumucs(Tmax=1, tdump=10 ** 9, γ=0.01, δ=1.0, n=9, m=3, hmin=0.1, hcrit=0.05,
                    θ=1 / 9)
    init = S.film_init(S.INIT_CAP, precursor=0.1, href=0.1, clamp=True,
                       cap1=(150.0, 1 / 9, Lx / 2 + 0.5, Ly / 2 + 0.5))
    sts = {}
    for T in (np.float64, np.float32):
        sysc = S.SysConst(Lx, Ly, prm)
        st = S.Sys(sysc, "GPU", T=T)
        st.set_reference_height(init.href)
        st.init_height(init)
        S.equilibrium(st, sysc)
        sts[T] = st
    m0 = {T: st.reduce()["sum"] for T, st in sts.items()}

    def advance(n, t0):
        for st in sts.values():
            o = _lib.StepOpts()
            o.t0 = t0
            _lib.check(_lib.lib().swalbe_time_loop(st.handle, C.byref(prm.cparams()), n,
                                                   C.byref(o)), "time_loop")
        return [sts[T].get("height", np.float64) for T in (np.float64, np.float32)]

    h64, h32 = advance(1000, 0)
    e1k = relL2(h32, h64)
    h64, h32 = advance(49_000, 1000)
    e = relL2(h32, h64)
    m1 = {T: st.reduce()["sum"] for T, st in sts.items()}
    vol32 = np.clip(h32 - 0.1, 0, None).sum()
    vol64 = np.clip(h64 - 0.1, 0, None).sum()
    d64 = abs(m1[np.float64] - m0[np.float64]) / m0[np.float64]
    d32 = abs(m1[np.float32] - m0[np.float64]) / m0[np.float64]
    print(f"[accuracy] droplet 1024^2: rel-L2 h FP32 vs FP64 = {e1k:.3e} at 10^3 steps, "
          f"{e:.3e} at 5 10^4; mass drift FP64 {d64:.2e}, FP32 {d32:.2e}; max h {h64.max():.4f} / "
          f"{h32.max():.4f}; volume above precursor {vol64:.1f} / {vol32:.1f}", flush=True)
    for st in sts.values():
        st.close()
    assert np.isfinite(h32).all()
    assert e1k <= 1e-5
    assert d64 <= 1e-10
    assert d32 <= 1e-4
    assert abs(h32.max() - h64.max()) <= 1e-3 * h64.max()
    assert abs(vol32 - vol64) <= 1e-2 * vol64
