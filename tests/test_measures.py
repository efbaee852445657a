"""Measurements of src/measures.jl (wetted!, fluid_dry!, surfacearea!,
∇f_simple!, t0): the oracle against the reference's known answers
(tests/golden/reference_kat_measures.json), the GPU against both, and the
fused time-loop series against the oracle measured step by step."""
import json
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = json.load(open(os.path.join(HERE, "golden", "reference_kat_measures.json")))["cases"]


def _arr(v, shape):
    return np.asarray(v, dtype=float).reshape(shape, order="F")


def _check_case(c, meas, grad, t0):
    """meas(h, theta_field, theta) -> (dict, mask); grad(f) -> (gx, gy)."""
    op = c["op"]
    if op == "wetted":
        for st in c["steps"]:
            m, _ = meas(_arr(st["h"], c["shape"]), None, 0.0)
            assert m["count"] == st["area"] and m["max"] == st["maxh"], c["name"]
    elif op == "fluid_dry":
        for st in c["steps"]:
            _, mask = meas(_arr(st["h"], c["shape"]), None, 0.0)
            assert np.array_equal(mask.ravel(order="F").astype(float), np.asarray(st["mask"])), c["name"]
    elif op == "grad_simple":
        gx, gy = grad(_arr(c["f"], c["shape"]))
        assert np.array_equal(gx, _arr(c["gx"], c["shape"])), c["name"]
        assert np.array_equal(gy, _arr(c["gy"], c["shape"])), c["name"]
    elif op == "surfacearea":
        tf = None if c["theta_field"] is None else _arr(c["theta_field"], c["shape"])
        m, _ = meas(_arr(c["h"], c["shape"]), tf, c["theta"] or 0.0)
        assert abs(m["area"] - c["area"]) <= c["atol"], c["name"]
        if "energy" in c:
            want = c["energy"]
        else:
            th = c["theta"] if c["theta"] is not None else tf[0, 0]
            want = c["area"] - c["energy_minus_cos"] * O.cospi(th)
        assert abs(m["energy"] - want) <= c["atol"], c["name"]
    elif op == "t0":
        got = t0()[0]
        assert abs(got - c["expect"]) <= c["rtol"] * c["expect"], (c["name"], got)
    else:
        raise ValueError(op)


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_oracle_measures_reference_kat(case):
    _check_case(case, lambda h, tf, th: O.measure(h, tf, th), O.grad_simple, O.t0)


def test_oracle_surfacearea_matches_its_definition():
    # the reference's multi-step check (test/measures.jl:98-115): area = Σ sqrt(gx² + gy² + 1)
    # over every node (all above the threshold), energy = area - Σ cos πθ
    h = np.arange(1.0, 26.0).reshape((5, 5), order="F")
    th = np.full((5, 5), 1 / 9)
    for t in range(5):
        h[t, t] = t + 1
        th[t, t] = 0.0
        gx, gy = O.grad_simple(h)
        m, _ = O.measure(h, th)
        area = np.sqrt(gx ** 2 + gy ** 2 + 1).sum()
        assert np.isclose(m["area"], area, rtol=1e-14)
        assert np.isclose(m["energy"], area - sum(O.cospi(v) for v in th.ravel()), rtol=1e-14)


def test_host_t0_matches_oracle():
    import swalbe as S
    assert S.t0()[0] == O.t0()[0]
    assert S.t0(hᵦ=0.1, γ=0.02, θ=1 / 9) == O.t0(0.1, 0.02, 1 / 6, 1 / 9)


# ---------------------------------------------------------------- GPU


@pytest.fixture(scope="module")
def S():
    import swalbe
    if swalbe.device_count() == 0:
        pytest.skip("no GPU")
    return swalbe


def _gpu_meas(S, T):
    def meas(h, tf, th):
        m = np.zeros(h.shape, dtype=np.uint8, order="F")
        r = S.measure(h.astype(T), 0.055, θ=th if tf is None else tf.astype(T), mask=m)
        return r, m
    return meas


def _gpu_grad(S, T):
    def grad(f):
        gx, gy = np.zeros(f.shape, dtype=T, order="F"), np.zeros(f.shape, dtype=T, order="F")
        S.grad_f_simple(gx, gy, f.astype(T))
        return gx.astype(float), gy.astype(float)
    return grad


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_gpu_measures_reference_kat(S, case):
    _check_case(case, _gpu_meas(S, np.float64), _gpu_grad(S, np.float64), S.t0)


@pytest.mark.gpu
def test_gpu_measure_of_a_state_matches_oracle(S):
    Lx, Ly = 130, 70
    rng = np.random.default_rng(2)
    h = 0.05 + np.abs(rng.standard_normal((Lx, Ly))) * 0.3
    th = rng.uniform(1 / 18, 1 / 6, (Lx, Ly))
    want, wmask = O.measure(h, th, 0.0)
    sysc = S.SysConst(Lx, Ly, S.Taumucs())
    for T, tol in ((np.float64, 1e-12), (np.float32, 1e-6)):
        st = S.Sys(sysc, "GPU", T=T)
        st.height = h
        mask = np.zeros((Lx, Ly), dtype=np.uint8, order="F")
        got = S.measure(st, 0.055, θ=th, mask=mask)
        assert abs(got["area"] - want["area"]) <= tol * want["area"]
        assert abs(got["energy"] - want["energy"]) <= tol * abs(want["area"])
        if T is np.float64:
            assert got["count"] == want["count"] and got["max"] == want["max"]
            assert np.array_equal(mask, wmask)
        # run to run: the same bits
        assert S.measure(st, 0.055, θ=th) == S.measure(st, 0.055, θ=th)


@pytest.mark.gpu
def test_fused_loop_series_match_oracle_step_by_step(S):
    # time_loop with the measurement series == the oracle's loop measured
    # before every step (FP64, the step bit-exact, the sums to 1e-12)
    Lx, Ly, n = 64, 48, 30
    rng = np.random.default_rng(5)
    h0 = 0.6 + 0.1 * rng.standard_normal((Lx, Ly))
    h0 = np.asfortranarray(np.maximum(h0, 0.06))
    th = np.asfortranarray(rng.uniform(1 / 18, 1 / 6, (Lx, Ly)))
    sysc = S.SysConst(Lx, Ly, S.Taumucs(Tmax=n, tdump=1000))
    st = S.Sys(sysc, "GPU", T=np.float64)
    st.height = h0
    S.equilibrium(st, sysc)
    st.set("theta", th)
    maxh, area, energy = [], [], []
    fluid = np.zeros((n, Lx * Ly), dtype=bool)
    S.time_loop(sysc, st, maxheight=maxh, area_lv=area, red_energy=energy, fluid=fluid,
                measure_thresh=0.6, surface_theta_field=True)
    ost = O.State(Lx, Ly)
    ost.height[...] = h0
    ost.feq[...], _ = O.equilibrium(ost.height, ost.velx, ost.vely, 0.0)
    prm = O.params(tdump=1000)
    for t in range(n):
        m, mask = O.measure(ost.height, th, 0.0, thresh=0.6)
        assert maxh[t] == m["max"], t
        assert abs(area[t] - m["area"]) <= 1e-12 * m["area"], t
        assert abs(energy[t] - m["energy"]) <= 1e-12 * abs(m["area"]), t
        assert np.array_equal(fluid[t], mask.ravel(order="F").astype(bool)), t
        O.time_loop(ost, prm, 1, t0=t)
    assert np.array_equal(st.height, ost.height)


@pytest.mark.gpu
def test_reference_style_callbacks_on_a_state(S):
    # wetted!(area, maxheight, state, t), fluid_dry!, surfacearea! with host arrays
    Lx, Ly = 40, 30
    h = np.full((Lx, Ly), 0.05)
    h[10:20, 5:15] = 1.0
    sysc = S.SysConst(Lx, Ly, S.Taumucs())
    st = S.Sys(sysc, "GPU")
    st.height = h
    area, maxh = np.zeros(3), np.zeros(3)
    S.wetted(area, maxh, st, 2)
    assert area[1] == 100 and maxh[1] == 1.0
    fluid, dummy = np.zeros((2, Lx * Ly), dtype=bool), np.zeros((Lx, Ly), dtype=bool)
    S.fluid_dry(fluid, dummy, st, 1)
    assert np.array_equal(dummy, h > 0.055) and fluid[0].sum() == 100
    alv, ren = np.zeros(1), np.zeros(1)
    dx, dy = np.zeros((Lx, Ly)), np.zeros((Lx, Ly))
    S.surfacearea(alv, ren, h, 1 / 9, dx, dy, None, None, 1)
    want, _ = O.measure(h, None, 1 / 9)
    assert abs(alv[0] - want["area"]) <= 1e-12 * want["area"]
    gx, gy = O.grad_simple(h)
    assert np.array_equal(dx, gx) and np.array_equal(dy, gy)
