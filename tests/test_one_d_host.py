"""CPU checks of the 1D host side: initial conditions against the reference's
tests, the obstacle masks against the oracle, and the 1D dispatch of the
package's generic names (no GPU calls)."""
import numpy as np

import oracle as O


def test_singledroplet_1d_reference_kat():
    # test/initialvalues.jl:25-33: rad 50, θ = 1/3, center 100 on ones(200)
    import swalbe as S
    rad, th, center = 50, 1 / 3, 100
    h = S.singledroplet(np.ones(200), rad, th, center)
    assert h.shape == (200,)
    assert h.max() == rad * (1 - S.cospi(th))
    assert int(np.argmax(h)) + 1 == center  # 1-based like the reference
    # the 1D method clamps h <= hcrit to hcrit (src/initialvalues.jl:236-240)
    assert h.min() == 0.05


def test_two_droplets_reference_kat():
    # test/initialvalues.jl:36-44
    import swalbe as S
    sysc = S.SysConst_1D(L=200, param=S.Taumucs())
    h = S.two_droplets(sysc, r1=45, r2=45, θ1=1 / 4, θ2=1 / 4)
    assert h.shape == (200,)
    assert abs(h.max() - 45 * (1 - S.cospi(1 / 4))) <= 0.01
    # Unicode keywords as the reference spells them
    h2 = S.two_droplets(sysc, **{"r₁": 45, "r₂": 45, "θ₁": 1 / 4, "θ₂": 1 / 4})
    assert np.array_equal(h, h2)


def test_obslist_matches_oracle_and_reference_wall():
    import swalbe as S
    rng = np.random.default_rng(5)
    obs = (rng.random(64) < 0.4).astype(float)
    interior, border = S.obslist1D(obs, verbose=False)
    oi, ob = O.obslist_1d(obs)
    assert np.array_equal(interior, oi)
    assert np.array_equal(border[0], ob[0]) and np.array_equal(border[1], ob[1])
    # test/collide.jl:131-137: walls at the two ends of a periodic line
    obst = np.zeros(30)
    obst[0] = obst[-1] = 1
    sysc = S.SysConstWithBound_1D(L=30, param=S.Taumucs(τ=0.75), obs=obst)
    S.obslist(sysc)
    assert not sysc.interior.any()
    assert sysc.border[0][29] == 1 and sysc.border[0].sum() == 1  # obsright
    assert sysc.border[1][0] == 1 and sysc.border[1].sum() == 1  # obsleft


def test_randinterface_1d_uses_the_counter_stream():
    import swalbe as S
    h = np.ones(100)
    S.randinterface(h, 2.0, 0.1, seed=4)
    z0, _ = S.normals(100, 4)
    assert np.array_equal(h, 2.0 * (1.0 + 0.1 * z0))


def test_one_d_names_are_exported():
    import swalbe as S
    for name in ("SysConst_1D", "SysConstWithBound_1D", "HipState1D", "obslist1D", "obslist",
                 "two_droplets", "run_gamma", "ensemble_time_loop", "grad_gamma",
                 "resident_max_length"):
        assert hasattr(S, name), name
    assert set(S.one_d.__all__) <= set(dir(S.one_d))
