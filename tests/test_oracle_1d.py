"""The 1D (D1Q3) oracle against the reference's known answers and against
what the mathematics fixes: equilibrium moments, mass conservation (with
and without bounce-back walls), streaming as a roll, the noise amplitude."""
import numpy as np
import pytest

import oracle as O
from kat1d import load_cases, run_case


class Oracle1D:
    equilibrium = staticmethod(O.equilibrium_1d)

    @staticmethod
    def bgk_and_stream(feq, ftemp, F, tau):
        return O.bgk_and_stream_1d(feq, ftemp, F, tau)[0]

    moments = staticmethod(O.moments_1d)

    @staticmethod
    def filmpressure(h, gamma, theta, n, m, hmin, hcrit, form):
        if form == "raw":
            return O.filmpressure_raw_1d(h, gamma, theta, n, m, hmin, hcrit)
        return O.filmpressure_1d(h, gamma, theta, n, m, hmin, hcrit)

    grad = staticmethod(O.grad_1d)
    laplacian = staticmethod(O.laplacian_1d)
    slippage = staticmethod(O.slippage_1d)

    @staticmethod
    def hgradp(p, h):
        return O.hgradp_1d(p, h)

    thermal = staticmethod(O.thermal_1d)
    inclination = staticmethod(O.inclination_1d)

    @staticmethod
    def run_flat(L, Tmax, tdump):
        st = O.State1D(L)
        st.height[...] = 1.0
        O.time_loop_1d(st, O.params(tdump=tdump), Tmax)
        return st.height


CASES = load_cases()


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_oracle_1d_reference_kat(case):
    run_case(Oracle1D, case)


def test_equilibrium_moments_closed_form():
    rng = np.random.default_rng(1)
    h, u = rng.uniform(0.5, 2.0, 40), rng.uniform(-0.1, 0.1, 40)
    for g in (0.0, 0.3):
        f = O.equilibrium_1d(h, u, g)
        assert np.allclose(f.sum(axis=1), h, rtol=1e-14)
        assert np.allclose(f[:, 1] - f[:, 2], h * u, rtol=1e-13, atol=1e-16)
        # second moment: h u^2 + g h^2 / 2 (shallow water)
        assert np.allclose(f[:, 1] + f[:, 2], h * u * u + 0.5 * g * h * h, rtol=1e-12)


def test_streaming_is_a_roll_and_force_moment():
    rng = np.random.default_rng(2)
    L = 17
    feq = rng.uniform(0.1, 1.0, (L, 3))
    F = rng.uniform(-0.01, 0.01, L)
    fout, _ = O.bgk_and_stream_1d(feq, np.zeros((L, 3)), F, 1.0)
    assert np.array_equal(fout[:, 0], feq[:, 0])
    assert np.array_equal(fout[:, 1], np.roll(feq[:, 1] + 0.5 * F, 1))
    assert np.array_equal(fout[:, 2], np.roll(feq[:, 2] - 0.5 * F, -1))


def test_mass_is_conserved_with_and_without_walls():
    L = 120
    st = O.State1D(L)
    x = np.arange(L)
    st.height[...] = 1.0 + 0.05 * np.sin(2 * np.pi * 3 * x / L)
    st.feq[...] = O.equilibrium_1d(st.height, st.vel, 0.0)
    m0 = st.height.sum()
    O.time_loop_1d(st, O.params(tdump=10 ** 9), 500)
    assert abs(st.height.sum() - m0) <= 1e-12 * m0
    # bounce-back walls at both ends of the film (four-node obstacles)
    obs = np.zeros(L)
    obs[:4] = obs[-4:] = 1
    _, (b1, b2) = O.obslist_1d(obs)
    st2 = O.State1D(L)
    st2.height[...] = 1.0 + 0.05 * np.sin(2 * np.pi * x / L)
    st2.feq[...] = O.equilibrium_1d(st2.height, st2.vel, 0.0)
    m0 = st2.height.sum()
    O.time_loop_1d(st2, O.params(tdump=10 ** 9), 300, borders=(b1, b2))
    assert abs(st2.height.sum() - m0) <= 1e-10 * m0
    # the walls change the flow: the same film without them differs
    st3 = O.State1D(L)
    st3.height[...] = 1.0 + 0.05 * np.sin(2 * np.pi * x / L)
    st3.feq[...] = O.equilibrium_1d(st3.height, st3.vel, 0.0)
    O.time_loop_1d(st3, O.params(tdump=10 ** 9), 300)
    assert np.max(np.abs(st3.height - st2.height)) > 1e-6


def test_obslist_matches_the_reference_example():
    # test/collide.jl:131-138: walls at the first and last node
    obs = np.zeros(30)
    obs[0] = obs[-1] = 1
    interior, (right, left) = O.obslist_1d(obs)
    assert interior.sum() == 0
    assert right[-1] == 1 and right.sum() == 1   # obs[i] and obs[i+1] (periodic)
    assert left[0] == 1 and left.sum() == 1      # obs[i] and obs[i-1]


def test_rayleigh_taylor_1d_grows():
    # test/simulate.jl:38-40: L = 512, g = -0.002, k = 4, ε = 0.01
    L = 512
    st = O.State1D(L)
    i = np.arange(1, L + 1)
    st.height[...] = 1.0 * (1 + 0.01 * np.sin(2 * np.pi * 4 * i / (L - 1)))
    st.feq[...] = O.equilibrium_1d(st.height, st.vel, -0.002)
    out = O.time_loop_1d(st, O.params(g=-0.002, tdump=500), 1000, dh=True)
    assert out["dh"][0] < out["dh"][-1]


def test_gamma_field_pressure_reduces_to_the_scalar_form():
    # a constant γ field is the scalar γ of the KATs; a γ field scales both
    # terms node by node
    rng = np.random.default_rng(3)
    h = rng.uniform(0.2, 1.5, 50)
    p_s = O.filmpressure_1d(h, 0.01, 1 / 9, 9, 3, 0.1, 0.05)
    p_f = O.filmpressure_1d(h, 0.0, 1 / 9, 9, 3, 0.1, 0.05, gamma_field=np.full(50, 0.01))
    assert np.array_equal(p_s, p_f)
    gam = rng.uniform(0.005, 0.02, 50)
    p_g = O.filmpressure_1d(h, 0.0, 1 / 9, 9, 3, 0.1, 0.05, gamma_field=gam)
    p_1 = O.filmpressure_1d(h, 1.0, 1 / 9, 9, 3, 0.1, 0.05)
    assert np.allclose(p_g, gam * p_1, rtol=1e-13, atol=1e-18)


def test_grad_gamma_forms_on_a_ramp():
    # ∇γ!(state): -3/2 (γ[x-1] - γ[x+1]) / 2 = 3/2 · slope inside a linear ramp;
    # ∇γ!(state, sys): (2h² + 6δh + 3δ²)/(6h) · h/2 · slope
    L = 40
    x = np.arange(L, dtype=float)
    gam = 0.01 + 1e-4 * x
    g0 = O.grad_gamma_1d(gam, form=0)
    assert np.allclose(g0[1:-1], 1.5 * 1e-4, rtol=1e-9)
    # the periodic wrap sees the jump at the ends
    assert g0[0] == -1.5 * ((gam[-1] - gam[1]) / 2.0)
    assert np.all(O.grad_gamma_1d(np.full(L, 0.02), form=0) == 0.0)
    h = np.full(L, 0.7)
    d = 1.0
    g1 = O.grad_gamma_1d(gam, h, delta=d, form=1)
    want = (2 * 0.49 + 6 * d * 0.7 + 3 * d * d) / (6 * 0.7) * 0.7 / 2 * (-1e-4)
    assert np.allclose(g1[1:-1], want, rtol=1e-9)


def test_grad_gamma_force_sign_drives_the_flow():
    # uniform γ in the pressure, a γ step only in ∇γ: from rest on a flat
    # film the first step's velocity is the ∇γ force, so the two signs of
    # the force term (run_gamma subtracts it, the coalescence scans add it)
    # give opposite velocities where γ jumps and none elsewhere
    L = 64
    prm = O.params(gamma=0.01, theta=1 / 9, n=9, m=3, hmin=0.1, hcrit=0.05, tdump=1000)
    gam = np.where(np.arange(L) < L // 2, 0.01, 0.02)
    gg = O.grad_gamma_1d(gam, form=0)
    assert gg[L // 2 - 1] > 0 and gg[L // 2] > 0
    vel = {}
    for sign in (-1, 1):
        st = O.State1D(L)
        st.feq[...] = O.equilibrium_1d(st.height, st.vel, 0.0)
        st.ftemp[...] = st.feq
        st.fout[...] = st.feq
        O.time_loop_1d(st, prm, 1, grad_gamma=gg, grad_gamma_sign=sign)
        vel[sign] = st.vel.copy()
        assert abs(st.height.sum() - L) < 1e-12
    j = L // 2
    assert vel[-1][j] < 0 < vel[1][j]
    assert np.allclose(vel[-1], -vel[1], rtol=2e-2, atol=1e-15)
    far = np.r_[4:j - 4, j + 5:L - 4]
    assert np.all(vel[-1][far] == 0.0)


def test_numpy_restatement_in_float64_is_the_c_oracle():
    # the dtype-generic restatement (the FP32 reference) run in float64 equals
    # the C oracle bit for bit, with and without the γ terms
    L, n = 128, 60
    x = np.arange(L)
    h = 1.0 + 0.2 * np.sin(2 * np.pi * 3 * x / L)
    prm = O.params(tdump=1000)
    gam = 0.01 + 0.002 * np.sin(2 * np.pi * x / L)
    gg = O.grad_gamma_1d(gam, form=0)
    for kw in (dict(), dict(gamma_field=gam, grad_gamma=gg, grad_gamma_sign=-1)):
        st = O.State1D(L)
        st.height[...] = h
        st.feq[...] = O.equilibrium_1d(h, np.zeros(L), 0.0)
        O.time_loop_1d(st, prm, n, **kw)
        h64, u64 = O.time_loop_1d_f32(h, prm, n, dtype=np.float64, **kw)
        assert np.array_equal(h64, st.height) and np.array_equal(u64, st.vel)
