"""GPU parity of the 1D (D1Q3) path: the reference's 1D known answers through
the HIP library, and multi-step FP64 runs equal to the oracle to the bit for
every term of the step (noise, θ field, γ field and ∇γ, walls, τ ≠ 1, body
force), on both fused loops (one launch per step, whole film in LDS), the
per-op loop and the ensemble launch; FP32 within a stated tolerance."""
import sys

import numpy as np
import pytest

import oracle as O
from kat1d import load_cases, run_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S():
    import swalbe
    if swalbe.device_count() == 0:
        pytest.skip("no GPU")
    return swalbe


def _sys(S, L, nsteps, **kw):
    P = dict(tau=1.0, delta=1.0, kbt=0.0, gamma=0.01, n=9, m=3, hmin=0.1, hcrit=0.05,
             theta=1 / 9, g=0.0, tdump=50)
    P.update(kw)
    mu = (1 / 3) * (P["tau"] - 0.5)
    tm = S.Taumucs(Tmax=nsteps, tdump=P["tdump"], τ=P["tau"], μ=mu, δ=P["delta"], kbt=P["kbt"],
                   γ=P["gamma"], n=P["n"], m=P["m"], hmin=P["hmin"], hcrit=P["hcrit"], θ=P["theta"],
                   g=P["g"])
    return S.SysConst_1D(L=L, param=tm), O.params(mu=mu, **P)


def _film(L, seed=0):
    rng = np.random.default_rng(seed)
    x = np.arange(L)
    return 1.0 + 0.2 * np.sin(2 * np.pi * 3 * x / L) + 0.01 * rng.standard_normal(L)


# ---------------------------------------------------------------- known answers


class HIP1D:
    """The kat1d backend on the HIP library (FP64 states)."""
    S = None

    @classmethod
    def _state(cls, L):
        sysc = cls.S.SysConst_1D(L=L, param=cls.S.Taumucs(Tmax=1))
        return sysc, cls.S.Sys(sysc, T=np.float64)

    @classmethod
    def equilibrium(cls, h, u, g):
        sysc, st = cls._state(len(h))
        st.height, st.vel = h, u
        cls.S.equilibrium(st, sysc, g=g)
        return st.feq

    @classmethod
    def bgk_and_stream(cls, feq, ftemp, F, tau):
        sysc, st = cls._state(feq.shape[0])
        st.feq, st.ftemp, st.F = feq, ftemp, F
        cls.S.BGKandStream(st, sysc, τ=tau)
        assert np.array_equal(st.fout, st.ftemp)  # fout .= ftemp
        return st.fout

    @classmethod
    def moments(cls, f):
        sysc, st = cls._state(f.shape[0])
        st.fout = f
        cls.S.moments(st)
        return st.height, st.vel

    @classmethod
    def filmpressure(cls, h, gamma, theta, n, m, hmin, hcrit, form):
        sysc, st = cls._state(len(h))
        st.height = h
        cls.S.one_d.filmpressure(st, sysc, θ=theta, γ=gamma, n=n, m=m, hmin=hmin, hcrit=hcrit,
                                 raw=form == "raw")
        return st.pressure

    @classmethod
    def grad(cls, f, a):
        out = np.zeros(len(f))
        cls.S.one_d.grad_f(out, f, None, a)
        return out

    @classmethod
    def laplacian(cls, f):
        out = np.zeros(len(f))
        cls.S.laplacian_f(out, f, np.zeros((len(f), 2)))
        return out

    @classmethod
    def slippage(cls, h, u, delta, mu):
        sysc, st = cls._state(len(h))
        sysc.param.δ, sysc.param.μ = delta, mu
        st.height, st.vel = h, u
        cls.S.slippage(st, sysc)
        return st.slip

    @classmethod
    def hgradp(cls, p, h):
        sysc, st = cls._state(len(p))
        st.pressure, st.height = p, h
        cls.S.h_grad_p(st)
        return st.get("h∇p")

    @classmethod
    def thermal(cls, h, kbt, mu, delta, seed, step):
        sysc, st = cls._state(len(h))
        sysc.param.kbt, sysc.param.μ, sysc.param.δ = kbt, mu, delta
        st.height = h
        cls.S.thermal(st, sysc, seed=seed, step=step)
        return st.kbt

    @classmethod
    def inclination(cls, F, h, alpha, fac):
        from swalbe._lib import check, lib
        sysc, st = cls._state(len(h))
        st.F, st.height = F, h
        check(lib().swalbe1d_inclination(st.handle, float(alpha), float(fac)), "inclination")
        return st.F

    @classmethod
    def run_flat(cls, L, Tmax, tdump):
        sysc = cls.S.SysConst_1D(L=L, param=cls.S.Taumucs(Tmax=Tmax, tdump=tdump))
        return cls.S.run_flat(sysc, verbos=False)


CASES = load_cases()


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_reference_kat_1d_on_gpu(S, case):
    HIP1D.S = S
    run_case(HIP1D, case)


def test_inclination_fac_from_t(S):
    # inclination!(α, state; t, tstart, tsmooth): fac = 0.5 + 0.5 tanh((t - tstart)/tsmooth)
    sysc = S.SysConst_1D(L=30, param=S.Taumucs(Tmax=1))
    st = S.Sys(sysc)
    st.F = 0.0
    S.inclination(0.1, st, t=1, tstart=0, tsmooth=1)
    assert np.all(st.F == 0.1 * (0.5 + 0.5 * np.tanh(1.0)))


# ---------------------------------------------------------------- multi-step parity

CONFIGS = {
    "plain": dict(),
    "thermal": dict(thermal=True, prm=dict(kbt=1e-6)),
    "theta_field": dict(theta_field=True),
    "body": dict(body=(1e-5, 1.0)),
    "walls": dict(walls=True),
    "tau": dict(prm=dict(tau=0.8)),
    "gamma_minus": dict(gamma_field=True, grad_gamma=-1),
    "gamma_plus": dict(gamma_field=True, grad_gamma=1),
    "gravity_n32": dict(prm=dict(g=0.05, n=3, m=2)),
}


def _setup(S, L, nsteps, cfg, T=np.float64, variant=0):
    sysc, prm = _sys(S, L, nsteps, **cfg.get("prm", {}))
    h = _film(L)
    walls = None
    if cfg.get("walls"):
        obs = np.zeros(L)
        obs[:6] = 1
        obs[L // 2: L // 2 + 6] = 1
        sysc = S.SysConstWithBound_1D(L=L, param=sysc.param, obs=obs)
        S.obslist(sysc)
        walls = sysc
    st = S.Sys(sysc, T=T, kind="gamma" if cfg.get("gamma_field") else "simple")
    st.height = h
    st.vel = 0.0
    S.equilibrium(st, sysc)
    st.ftemp = st.feq
    ost = O.State1D(L)
    ost.height[...] = h
    ost.feq[...] = O.equilibrium_1d(h, np.zeros(L), prm.g)
    ost.ftemp[...] = ost.feq
    okw = dict(thermal=cfg.get("thermal", False), seed=11, dh=True, wetted=1.05, mass=True)
    hkw = dict(thermal=cfg.get("thermal", False), seed=11, variant=variant)
    args = ()
    if cfg.get("theta_field"):
        th = np.where(np.arange(L) % 40 < 20, 1 / 9, 1 / 4)
        okw["theta_field"] = th
        args = (th,)
    if cfg.get("body"):
        okw["body"] = cfg["body"]
    if walls is not None:
        okw["borders"] = (walls.border[0], walls.border[1])
        okw["interior"] = walls.interior
    if cfg.get("gamma_field"):
        gam = 0.01 + 0.002 * np.sin(2 * np.pi * np.arange(L) / L)  # γ(x) changes h by ~1 %
        gg = O.grad_gamma_1d(gam, form=0)
        okw.update(gamma_field=gam, grad_gamma=gg, grad_gamma_sign=cfg["grad_gamma"])
        st.set("gamma", gam)
        S.grad_gamma(st)
        if T is np.float32:
            # the FP32 reference takes the FP64 ∇γ rounded once to FP32
            st.set("gradgamma", gg)
        hkw.update(γ=gam, grad_gamma_sign=cfg["grad_gamma"])
    return sysc, prm, st, ost, okw, hkw, args


def _run_hip(S, sysc, st, hkw, args, body):
    if body is not None:
        return S.time_loop(sysc, st, S.inclination, body[0], **hkw)
    dh = []
    S.time_loop(sysc, st, dh, *args, **hkw) if not args else S.time_loop(sysc, st, *args, **hkw)
    return dh


@pytest.mark.parametrize("variant", [1, 2], ids=["per_step", "resident"])
@pytest.mark.parametrize("name", list(CONFIGS))
def test_fp64_time_loop_bitwise_vs_oracle(S, name, variant):
    cfg = CONFIGS[name]
    L, n = 300, 200
    print(f"1D parity {name} variant {variant} L={L} steps={n}", flush=True)
    sysc, prm, st, ost, okw, hkw, args = _setup(S, L, n, cfg, variant=variant)
    out = O.time_loop_1d(ost, prm, n, **okw)
    dh = _run_hip(S, sysc, st, hkw, args, cfg.get("body"))
    for f_h, f_o in (("height", "height"), ("vel", "vel"), ("fout", "fout"), ("ftemp", "ftemp"),
                     ("pressure", "pressure"), ("F", "F"), ("slip", "slip"), ("hgradp", "hgp"),
                     ("feq", "feq")):
        a, b = st.get(f_h), getattr(ost, f_o)
        assert np.array_equal(a, b), (name, f_h, np.abs(a - b).max())
    if cfg.get("thermal"):
        assert np.array_equal(st.kbt, ost.kbt)
    if isinstance(dh, list) and dh:
        assert np.array_equal(np.array(dh), out["dh"])


@pytest.mark.parametrize("threads", [512, 1024])
@pytest.mark.parametrize("L,name", [(70, "thermal"), (3001, "thermal"), (1500, "walls"),
                                    (2049, "gamma_minus"), (1100, "theta_field")],
                         ids=lambda v: str(v))
def test_per_step_workgroup_sizes_bitwise_vs_oracle(S, L, name, threads):
    """The one-launch-per-step kernel with 512 or 1024 nodes per workgroup
    (a film shorter than one workgroup, ragged lengths, walls, ∇γ, θ field):
    the oracle's bits, Δh series included."""
    cfg = CONFIGS[name]
    n = 120
    print(f"1D per-step {name} L={L} threads {threads}", flush=True)
    sysc, prm, st, ost, okw, hkw, args = _setup(S, L, n, cfg, variant=1)
    hkw["threads"] = threads
    out = O.time_loop_1d(ost, prm, n, **okw)
    dh = _run_hip(S, sysc, st, hkw, args, cfg.get("body"))
    for f_h, f_o in (("height", "height"), ("vel", "vel"), ("fout", "fout"),
                     ("pressure", "pressure")):
        assert np.array_equal(st.get(f_h), getattr(ost, f_o)), (name, f_h)
    if isinstance(dh, list) and dh:
        assert np.array_equal(np.array(dh), out["dh"])


@pytest.mark.parametrize("threads", [256, 512])
@pytest.mark.parametrize("L,name", [(1030, "plain"), (1300, "thermal"), (1500, "walls"),
                                    (1100, "gamma_plus"), (1200, "theta_field")],
                         ids=lambda v: str(v))
def test_resident_nodes_per_thread_bitwise_vs_oracle(S, L, name, threads):
    """The resident kernel with several nodes per thread: the first two held
    with their per-node constants in registers, the rest read each step;
    films with and without walls (the latter compute the next step's
    pressure in the stream phase).  The oracle's bits, Δh series included."""
    cfg = CONFIGS[name]
    n = 120
    print(f"1D resident {name} L={L} threads {threads}", flush=True)
    sysc, prm, st, ost, okw, hkw, args = _setup(S, L, n, cfg, variant=2)
    hkw["threads"] = threads
    out = O.time_loop_1d(ost, prm, n, **okw)
    dh = _run_hip(S, sysc, st, hkw, args, cfg.get("body"))
    for f_h, f_o in (("height", "height"), ("vel", "vel"), ("fout", "fout"),
                     ("pressure", "pressure")):
        assert np.array_equal(st.get(f_h), getattr(ost, f_o)), (name, f_h)
    if isinstance(dh, list) and dh:
        assert np.array_equal(np.array(dh), out["dh"])


@pytest.mark.parametrize("threads", [256, 512])
@pytest.mark.parametrize("mask", ["fractional", "negative_zero"])
def test_resident_masks_of_any_value_bitwise_vs_oracle(S, mask, threads):
    """Wall masks other than the 0 and 1 that obslist! makes (a 0.5 and a
    0.25, or -0 whose products keep their sign) in the resident kernel, with
    the masks of each thread's first two nodes held in registers and the
    rest read from global memory: the oracle's bits."""
    L, n = 1500, 120
    print(f"1D resident masks {mask} L={L} threads {threads}", flush=True)
    sysc, prm, st, ost, okw, hkw, args = _setup(S, L, n, CONFIGS["walls"], variant=2)
    b1, b2 = sysc.border[0], sysc.border[1]
    if mask == "fractional":
        b1[L // 4] = 0.5
        b2[L // 4 + 1] = 0.25
    else:
        b1[np.flatnonzero(b1 == 0)[:7]] = -0.0
    st.set("border1", b1)
    st.set("border2", b2)
    okw["borders"] = (b1.copy(), b2.copy())
    hkw["threads"] = threads
    out = O.time_loop_1d(ost, prm, n, **okw)
    dh = _run_hip(S, sysc, st, hkw, args, None)
    for f_h, f_o in (("height", "height"), ("vel", "vel"), ("fout", "fout"),
                     ("pressure", "pressure")):
        assert np.array_equal(st.get(f_h), getattr(ost, f_o)), (mask, f_h)
    assert np.array_equal(np.array(dh), out["dh"])


@pytest.mark.parametrize("L", [3, 255, 257, 1000])
def test_fp64_lengths_and_both_loops_agree(S, L):
    print(f"1D lengths L={L}", flush=True)
    n = 60
    res = {}
    for variant in (1, 2):
        sysc, prm, st, ost, okw, hkw, args = _setup(S, L, n, CONFIGS["thermal"], variant=variant)
        S.time_loop(sysc, st, **hkw)
        res[variant] = (st.height, st.vel, st.fout)
    O.time_loop_1d(ost, prm, n, **okw)
    for v in (1, 2):
        assert np.array_equal(res[v][0], ost.height), v
        assert np.array_equal(res[v][1], ost.vel), v
        assert np.array_equal(res[v][2], ost.fout), v


def test_series_match_the_oracle(S):
    L, n = 400, 150
    sysc, prm, st, ost, okw, hkw, args = _setup(S, L, n, CONFIGS["walls"], variant=2)
    out = O.time_loop_1d(ost, prm, n, **okw)
    for variant in (1, 2):
        sysc, prm, st, _, _, hkw, _ = _setup(S, L, n, CONFIGS["walls"], variant=variant)
        dh, mass, area = [], [], []
        S.time_loop(sysc, st, dh, mass=mass, **hkw)
        assert np.array_equal(np.array(dh), out["dh"]), variant
        # Σ(h - interior h) in a different summation order than the oracle's
        assert np.allclose(mass, out["mass"][:len(mass)], rtol=1e-13), variant
        assert len(mass) == n // prm.tdump
        sysc, prm, st, _, _, hkw, _ = _setup(S, L, n, CONFIGS["walls"], variant=variant)
        S.time_loop(sysc, st, S.one_d.wetted, area, **hkw)
        want = O.time_loop_1d(_setup(S, L, n, CONFIGS["walls"])[3], prm, n,
                              borders=okw["borders"], interior=okw["interior"], wetted=0.055)
        assert np.array_equal(np.array(area), want["wetted"]), variant


def test_snapshots_match_the_oracle(S):
    L, n, dump = 256, 120, 40
    fluid = np.zeros((n // dump, L))
    sysc, prm, st, ost, okw, hkw, args = _setup(S, L, n, CONFIGS["plain"], variant=2)
    S.time_loop(sysc, st, snapshots=fluid, dumping=dump, **hkw)
    for k in range(n // dump):
        O.time_loop_1d(ost, prm, dump, t0=k * dump)
        assert np.array_equal(fluid[k], ost.height), k
    for variant in (1,):
        fl2 = np.zeros_like(fluid)
        sysc, prm, st, _, _, hkw, _ = _setup(S, L, n, CONFIGS["plain"], variant=variant)
        S.time_loop(sysc, st, snapshots=fl2, dumping=dump, **hkw)
        assert np.array_equal(fl2, fluid)


def test_per_op_loop_bitwise_vs_oracle(S):
    # a measurement callback forces the op-by-op loop (src/simulate.jl:159-178)
    L, n = 200, 40
    sysc, prm, st, ost, okw, hkw, args = _setup(S, L, n, CONFIGS["gamma_minus"])
    seen = []
    S.time_loop(sysc, st, lambda meas, s: meas.append(1), seen, γ=hkw["γ"], grad_gamma_sign=-1)
    assert len(seen) == n
    O.time_loop_1d(ost, prm, n, **okw)
    assert np.array_equal(st.height, ost.height)
    assert np.array_equal(st.fout, ost.fout)


def test_continuing_a_loop_equals_one_long_loop(S):
    L = 300
    sysc, prm, st, ost, okw, hkw, args = _setup(S, L, 80, CONFIGS["thermal"], variant=2)
    S.time_loop(sysc, st, **hkw)
    hkw["variant"] = 1
    S.time_loop(sysc, st, **hkw)  # t0 carries on: steps 81..160, other loop
    O.time_loop_1d(ost, prm, 160, **{k: v for k, v in okw.items() if k in ("thermal", "seed")})
    assert np.array_equal(st.height, ost.height)


def test_ensemble_equals_single_films_and_oracle(S):
    specs = [(300, "plain"), (128, "thermal"), (500, "gamma_minus"), (77, "walls"),
             (500, "tau"), (256, "theta_field")]
    n = 100
    films, systems, singles, oracles = [], [], [], []
    gam_for = {}
    for L, name in specs:
        cfg = CONFIGS[name]
        sysc, prm, st, ost, okw, hkw, args = _setup(S, L, n, cfg, variant=2)
        films.append((st, hkw, args))
        systems.append(sysc)
        O.time_loop_1d(ost, prm, n, **okw)
        oracles.append(ost)
    print(f"1D ensemble of {len(films)} films", flush=True)
    S.ensemble_time_loop(systems, [f[0] for f in films], n,
                         thermal=[f[1]["thermal"] for f in films], seed=11,
                         γ=[f[1].get("γ") for f in films],
                         grad_gamma_sign=[f[1].get("grad_gamma_sign", 0) for f in films],
                         θ=[f[2][0] if f[2] else None for f in films])
    for (st, _, _), ost, (L, name) in zip(films, oracles, specs):
        assert np.array_equal(st.height, ost.height), name
        assert np.array_equal(st.fout, ost.fout), name
        assert np.array_equal(st.pressure, ost.pressure), name


def test_ensemble_two_nodes_per_thread_equals_oracle(S):
    # 512 threads per film and films longer than 512 nodes: every thread
    # holds two nodes (the scan benchmark's shape), walled and open films mixed
    specs = [(1100, "plain"), (700, "thermal"), (1024, "gamma_plus"), (900, "walls"),
             (600, "theta_field")]
    n = 100
    films, systems, oracles = [], [], []
    for L, name in specs:
        sysc, prm, st, ost, okw, hkw, args = _setup(S, L, n, CONFIGS[name], variant=2)
        films.append((st, hkw, args))
        systems.append(sysc)
        O.time_loop_1d(ost, prm, n, **okw)
        oracles.append(ost)
    print(f"1D ensemble of {len(films)} films, 512 threads each", flush=True)
    S.ensemble_time_loop(systems, [f[0] for f in films], n,
                         thermal=[f[1]["thermal"] for f in films], seed=11,
                         γ=[f[1].get("γ") for f in films],
                         grad_gamma_sign=[f[1].get("grad_gamma_sign", 0) for f in films],
                         θ=[f[2][0] if f[2] else None for f in films], threads=512)
    for (st, _, _), ost, (L, name) in zip(films, oracles, specs):
        assert np.array_equal(st.height, ost.height), name
        assert np.array_equal(st.fout, ost.fout), name
        assert np.array_equal(st.pressure, ost.pressure), name


@pytest.mark.parametrize("name", ["thermal", "walls", "tau"])
def test_resident_at_its_maximum_length_bitwise_vs_oracle(S, name):
    """A film of exactly swalbe1d_resident_max_length nodes (every byte of
    LDS a workgroup may take; τ ≠ 1 holds f_t too, so its limit is lower)
    runs resident and gives the oracle's bits."""
    cfg = CONFIGS[name]
    L = S.resident_max_length(np.float64, tau_is_one=name != "tau")
    n = 60
    print(f"1D resident at the maximum length {name} L={L}", flush=True)
    sysc, prm, st, ost, okw, hkw, args = _setup(S, L, n, cfg, variant=2)
    out = O.time_loop_1d(ost, prm, n, **okw)
    dh = _run_hip(S, sysc, st, hkw, args, None)
    for f_h, f_o in (("height", "height"), ("vel", "vel"), ("fout", "fout"),
                     ("pressure", "pressure")):
        assert np.array_equal(st.get(f_h), getattr(ost, f_o)), (name, L, f_h)
    assert np.array_equal(np.array(dh), out["dh"])


def test_fp32_resident_at_its_maximum_length(S):
    """The FP32 limit (half the bytes per node, so twice the FP64 length)
    runs resident: rel-L2(h) against the FP64 oracle <= 1e-6 after 60 steps."""
    L = S.resident_max_length(np.float32)
    assert L > S.resident_max_length(np.float64)
    n = 60
    print(f"1D FP32 resident at the maximum length L={L}", flush=True)
    sysc, prm, st, ost, okw, hkw, args = _setup(S, L, n, CONFIGS["plain"], T=np.float32,
                                                variant=2)
    S.time_loop(sysc, st, **hkw)
    O.time_loop_1d(ost, prm, n, **okw)
    h = st.get("height", np.float64)
    rel = np.linalg.norm(h - ost.height) / np.linalg.norm(ost.height)
    print(f"1D FP32 L={L}: rel-L2 h {rel:.3e}", flush=True)
    assert np.isfinite(h).all() and rel <= 1e-6


def test_large_film_runs_per_step(S):
    L = 1 << 20
    n = 8
    assert S.resident_max_length(np.float64) < L
    sysc, prm, st, ost, okw, hkw, args = _setup(S, L, n, CONFIGS["thermal"])
    S.time_loop(sysc, st, **hkw)  # variant 0 picks one launch per step
    O.time_loop_1d(ost, prm, n, thermal=True, seed=11)
    assert np.array_equal(st.height, ost.height)


# FP32 against FP64 after 1000 steps: relative L2 of h.  The Marangoni case
# is set by FP32 arithmetic itself: the float32 restatement of the same
# expression order (oracle.time_loop_1d_f32) lands 2.8e-5 from FP64 too.
FP32_TOL = {"plain": 1e-5, "thermal": 1e-5, "walls": 1e-5, "gamma_minus": 5e-5}


@pytest.mark.parametrize("name", list(FP32_TOL))
def test_fp32_close_to_oracle(S, name):
    L, n = 512, 1000
    cfg = CONFIGS[name]
    sysc, prm, st, ost, okw, hkw, args = _setup(S, L, n, cfg, T=np.float32)
    S.time_loop(sysc, st, *args, **hkw)
    O.time_loop_1d(ost, prm, n, **okw)
    h = st.get("height", np.float64)
    rel = np.linalg.norm(h - ost.height) / np.linalg.norm(ost.height)
    assert rel <= FP32_TOL[name], (name, rel)


@pytest.mark.parametrize("name", ["plain", "gamma_minus"])
def test_fp32_matches_the_float32_restatement(S, name):
    # every operation rounded to float32 in the kernel's order on the CPU:
    # the GPU's FP32 film agrees to within a few float32 ulps of h
    L, n = 512, 1000
    cfg = CONFIGS[name]
    sysc, prm, st, ost, okw, hkw, args = _setup(S, L, n, cfg, T=np.float32)
    st.vel = 0.0
    S.time_loop(sysc, st, *args, **hkw)
    kw = {}
    if cfg.get("gamma_field"):
        kw = dict(gamma_field=okw["gamma_field"], grad_gamma=okw["grad_gamma"],
                  grad_gamma_sign=okw["grad_gamma_sign"])
    h32, _ = O.time_loop_1d_f32(ost.height, prm, n, **kw)
    h = st.get("height", np.float64)
    rel = np.linalg.norm(h - h32) / np.linalg.norm(h32)
    print(f"{name}: GPU FP32 vs float32 restatement rel-L2 {rel:.3e}, "
          f"bitwise {np.array_equal(h, h32)}", flush=True)
    assert rel <= 1e-6, (name, rel)


# ---------------------------------------------------------------- drivers and errors


def test_two_droplets_reference_doctest(S):
    # src/initialvalues.jl:262-268: maximum ≈ rad (1 - cospi θ), atol 0.01
    rad, th = 45, 1 / 4
    sysc = S.SysConst_1D(L=200, param=S.Taumucs())
    h = S.two_droplets(sysc, r1=rad, r2=rad, θ1=th, θ2=th)
    assert abs(h.max() - rad * (1 - np.cos(np.pi * th))) <= 0.01


def test_run_gamma_snapshots_and_mass(S):
    L = 512
    sysc = S.SysConst_1D(L=L, param=S.Taumucs(Tmax=2000, γ=0.01, θ=1 / 9))
    gamma = np.where(np.arange(L) < L // 2, 0.01, 0.012)
    fluid = S.run_gamma(sysc, gamma, r1=60, r2=60, verbos=False, dump=500)
    assert fluid.shape == (4, L)
    m0 = S.two_droplets(sysc, r1=60, r2=60, θ1=1 / 9, θ2=1 / 9).sum()
    assert np.allclose(fluid.sum(axis=1), m0, rtol=1e-10)
    assert np.all(np.isfinite(fluid))


def test_rayleigh_taylor_1d_grows(S):
    # test/simulate.jl:38-40 in 1D: L = 512, g = -0.002, k = 4, ϵ = 0.01
    sysc = S.SysConst_1D(L=512, param=S.Taumucs(Tmax=1000, tdump=500, g=-0.002))
    h, diff = S.run_rayleightaylor(sysc, k=4, eps=0.01, verbos=False)
    assert len(diff) == 1000 and diff[-1] > diff[0]


def test_dropletrelax_1d_area_series(S):
    sysc = S.SysConst_1D(L=300, param=S.Taumucs(Tmax=500))
    h, area = S.run_dropletrelax(sysc, radius=40, θ0=1 / 6, verbos=False)
    assert len(area) == 500 and area[0] > 0


def test_errors_are_reported(S):
    sysc = S.SysConst_1D(L=64, param=S.Taumucs(Tmax=3))
    st = S.Sys(sysc)
    with pytest.raises(ValueError):  # wall masks of another length
        S.one_d.BGKandStream(st, S.SysConstWithBound_1D(L=32, param=sysc.param))
    with pytest.raises(S.DomainError):
        S.one_d.filmpressure(st, sysc, n=5, m=2, raw=True)
    big = S.SysConst_1D(L=S.resident_max_length(np.float64) + 1, param=S.Taumucs(Tmax=1))
    with pytest.raises(S.SwalbeError):
        S.time_loop(big, S.Sys(big), variant=2)
    with pytest.raises(S.SwalbeError):
        S.time_loop(sysc, st, grad_gamma_sign=-1)  # no ∇γ field


_MANY_FILMS = r"""
import sys
import numpy as np
sys.path[:0] = [sys.argv[1]]
import swalbe as S
nf, L, n = int(sys.argv[2]), 128, 200
print(f"[many films] creating {nf} films", flush=True)
systems, states = [], []
for k in range(nf):
    sy = S.SysConst_1D(L=L, param=S.Taumucs(Tmax=n, tdump=10 ** 9, γ=0.01))
    st = S.Sys(sy)
    i = np.arange(1, L + 1)
    st.height = 1.0 + 0.01 * np.sin(2 * np.pi * (1 + k % 7) * i / L + 0.001 * k)
    S.equilibrium(st, sy)
    systems.append(sy)
    states.append(st)
print(f"[many films] ensemble of {nf} films, {n} steps", flush=True)
S.ensemble_time_loop(systems, states, n, materialize=False)
for k in (0, 1, nf // 2, nf - 1):
    sy = S.SysConst_1D(L=L, param=S.Taumucs(Tmax=n, tdump=10 ** 9, γ=0.01))
    with S.Sys(sy) as one:
        i = np.arange(1, L + 1)
        one.height = 1.0 + 0.01 * np.sin(2 * np.pi * (1 + k % 7) * i / L + 0.001 * k)
        S.equilibrium(one, sy)
        S.time_loop(sy, one, variant=2)
        assert np.array_equal(one.get("height"), states[k].get("height")), k
for st in states:
    st.close()
print("[many films] OK", flush=True)
"""


def test_ensemble_of_many_live_films(S):
    """6000 live films in one ensemble, each equal to the same film alone.
    The films share the library's few streams per device and take one
    allocation each.  Runs in a child process: the parent test session is
    not at stake if the process ends abnormally."""
    import os
    import subprocess
    import sys
    pkg = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "swalbe.jl_amd")
    r = subprocess.run([sys.executable, "-c", _MANY_FILMS, pkg, "6000"], capture_output=True,
                       text=True, timeout=240)
    print(r.stdout, r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    assert "[many films] OK" in r.stdout
