"""The 2D run_* drivers (src/simulate.jl:236-484) on the GPU.

Two kinds of check:
* the reference's own physics tests of the drivers (test/simulate.jl:3-169):
  a flat film stays flat, a random interface flattens, a hanging film grows
  (Rayleigh-Taylor), a relaxing or patterned droplet keeps its volume and
  reaches the equilibrium base radius within 10 %, a sliding droplet moves
  along x only and stays slow;
* bit-exact parity with the FP64 oracle over a few hundred steps of each
  driver.  The drivers run the reference's statements, in-place writes into
  state.height included (node-by-node assignment, randinterface! and
  singledroplet writing into the state's field), and the oracle starts from
  the same heights built on plain host arrays.
"""
import math

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


def _oracle_run(h0, prm, nsteps, **kw):
    Lx, Ly = h0.shape
    st = O.State(Lx, Ly)
    st.height[...] = h0
    st.feq[...], _ = O.equilibrium(st.height, st.velx, st.vely, prm.g)
    ser = O.time_loop(st, prm, nsteps, **kw)
    return st, ser


def _cap_volume(rad, th):
    return math.pi / 3 * rad ** 3 * (2 + math.cos(math.pi * th)) * (1 - math.cos(math.pi * th)) ** 2


def _equilibrium_base_radius(rad, th0, th1):
    """Base radius of the cap of contact angle th1 with the volume of the
    cap of sphere radius rad and angle th0 (test/simulate.jl:48-50)."""
    R1 = ((rad ** 3 * (2 + math.cos(math.pi * th0)) * (1 - math.cos(math.pi * th0)) ** 2)
          / ((2 + math.cos(math.pi * th1)) * (1 - math.cos(math.pi * th1)) ** 2)) ** (1 / 3)
    return math.sin(math.pi * th1) * R1


def _droplet_shape(h):
    """test/simulate.jl:51-58: base radius from the wet part of row x = 75
    (1-based), height from the maximum, volume of the cap they span."""
    h = np.asarray(h)
    drop = np.nonzero(h[74, :] > 0.055)[0]
    droprad = len(drop) / 2
    droph = h.max()
    vnum = 1 / 6 * math.pi * droph * (3 * droprad ** 2 + droph ** 2)
    return droprad, vnum


# ---------------------------------------------------------------- the reference's physics tests


def test_flat_interface(S):
    # test/simulate.jl:4-8
    sysc = S.SysConst(25, 25, S.Taumucs(Tmax=200, tdump=100))
    h = S.run_flat(sysc, "GPU", verbos=False)
    assert np.all(np.asarray(h) == 1.0)
    assert np.asarray(h).sum() == 25 * 25
    S.run_flat(sysc, "GPU", verbos=True)


def test_random_interface_flattens(S):
    # test/simulate.jl:17-22: after 10^4 steps max - min < the initial ϵ
    sysc = S.SysConst(25, 25, S.Taumucs(Tmax=10000, tdump=5000))
    h = S.run_random(sysc, "GPU", eps=0.1, verbos=False)
    assert h.max() - h.min() < 0.1
    S.run_random(S.SysConst(25, 25, S.Taumucs(Tmax=100, tdump=500)), "GPU", eps=0.1, verbos=True)


def test_rayleigh_taylor_grows(S):
    # test/simulate.jl:33-35
    sysc = S.SysConst(100, 100, S.Taumucs(Tmax=1000, tdump=500, g=-0.002))
    _, diff = S.run_rayleightaylor(sysc, "GPU", kx=4, ky=5, eps=0.01, verbos=True)
    assert len(diff) == 1000
    assert diff[0] < diff[-1]


def test_relaxing_droplet(S):
    # test/simulate.jl:43-61: volume and base radius within 10 % after 10^4 steps
    rads = 35
    sysc = S.SysConst(150, 150, S.Taumucs(Tmax=10000, δ=3.0))
    h, A = S.run_dropletrelax(sysc, "GPU", radius=rads, verbos=False)
    vol = _cap_volume(rads, 1 / 6)
    r1 = _equilibrium_base_radius(rads, 1 / 6, 1 / 9)
    droprad, vnum = _droplet_shape(h)
    print(f"[drivers] relaxing droplet: volume {vnum:.1f} (cap {vol:.1f}), base radius "
          f"{droprad} (equilibrium {r1:.2f}), wetted {A[0]} -> {A[-1]}", flush=True)
    assert abs(vol - vnum) <= vol / 100 * 10
    assert abs(r1 - droprad) <= r1 / 100 * 10
    assert A[0] < A[-1]
    S.run_dropletrelax(S.SysConst(150, 150, S.Taumucs(Tmax=100, δ=3.0, tdump=50)), "GPU",
                       radius=rads, verbos=True)


def test_patterned_droplet(S):
    # test/simulate.jl:97-113
    sysc = S.SysConst(150, 150, S.Taumucs(Tmax=10000, δ=3.0))
    h = S.run_dropletpatterned(sysc, "GPU", radius=35, θs=np.full((150, 150), 1 / 9),
                               verbos=False)
    vol = _cap_volume(35, 1 / 6)
    r1 = _equilibrium_base_radius(35, 1 / 6, 1 / 9)
    droprad, vnum = _droplet_shape(h)
    print(f"[drivers] patterned droplet: volume {vnum:.1f} (cap {vol:.1f}), base radius "
          f"{droprad} (equilibrium {r1:.2f})", flush=True)
    assert abs(vol - vnum) <= vol / 100 * 10
    assert abs(r1 - droprad) <= r1 / 100 * 10


def test_sliding_droplet(S):
    # test/simulate.jl:147-156: the maximum moves in x (not at x = 75), stays
    # at y = 75 (1-based), and the velocities stay below 0.1
    sysc = S.SysConst(150, 150, S.Taumucs(Tmax=5000, δ=2.0))
    h, u, v = S.run_dropletforced(sysc, "GPU", radius=35, fx=1e-4, verbos=False)
    h = np.asarray(h)
    ix, iy = np.unravel_index(np.argmax(h), h.shape)
    print(f"[drivers] sliding droplet: maximum at ({ix + 1}, {iy + 1}) (1-based), "
          f"max u {np.asarray(u).max():.3e}, max v {np.asarray(v).max():.3e}", flush=True)
    assert ix + 1 != 75
    assert iy + 1 == 75
    assert np.all(np.asarray(u) < 0.1)
    assert np.all(np.asarray(v) < 0.1)


# ---------------------------------------------------------------- bit-exact against the oracle


def test_rayleigh_taylor_driver_matches_oracle(S):
    """The node-by-node writes of the sine undulation (src/simulate.jl:350-353)
    reach the device: 300 steps equal the oracle from the same heights."""
    Lx = Ly = 100
    sysc = S.SysConst(Lx, Ly, S.Taumucs(Tmax=300, tdump=100, g=-0.002))
    h, diff = S.run_rayleightaylor(sysc, "GPU", kx=4, ky=5, eps=0.01, verbos=False)
    h0 = np.empty((Lx, Ly), order="F")
    for i in range(1, Lx + 1):
        for j in range(1, Ly + 1):
            h0[i - 1, j - 1] = 1.0 * (1 + 0.01 * math.sin(2 * math.pi * 4 * i / (Lx - 1))
                                      * math.sin(2 * math.pi * 5 * j / (Ly - 1)))
    ost, ser = _oracle_run(h0, O.params(g=-0.002, tdump=100), 300, dh=True)
    assert np.array_equal(np.asarray(h), ost.height)
    assert np.array_equal(np.array(diff), ser["dh"])


def test_random_driver_matches_oracle(S):
    """randinterface! writes into state.height (src/simulate.jl:289)."""
    sysc = S.SysConst(25, 25, S.Taumucs(Tmax=300, tdump=100))
    h = S.run_random(sysc, "GPU", eps=0.1, verbos=False, seed=3)
    h0 = np.ones((25, 25), order="F")
    S.randinterface(h0, 1.0, 0.1, seed=3)
    ost, _ = _oracle_run(h0, O.params(theta=1 / 9, tdump=100), 300)
    assert np.array_equal(np.asarray(h), ost.height)


def test_dropletrelax_driver_matches_oracle(S):
    """singledroplet writes into state.height (src/simulate.jl:394) and the
    wetted! callback series (src/measures.jl:13-17)."""
    sysc = S.SysConst(150, 150, S.Taumucs(Tmax=300, δ=3.0, tdump=100))
    h, A = S.run_dropletrelax(sysc, "GPU", radius=35, verbos=False)
    h0 = S.singledroplet(np.ones((150, 150), order="F"), 35, 1 / 6, (75, 75))
    ost, ser = _oracle_run(h0, O.params(delta=3.0, tdump=100), 300, wetted=0.055)
    assert np.array_equal(np.asarray(h), ost.height)
    assert np.array_equal(np.array(A), ser["wetted"])


def test_dropletpatterned_driver_matches_oracle(S):
    """A two-valued contact-angle field whose boundary crosses the droplet
    (src/simulate.jl:433-436), on the reference test's geometry (150^2,
    radius 35: singledroplet's rule keeps tiny positive heights at the rim of
    some other radii, which blow the film up in the oracle as well)."""
    Lx = Ly = 150
    th = np.full((Lx, Ly), 1 / 9, order="F")
    th[:, : Ly // 2] = 1 / 8
    sysc = S.SysConst(Lx, Ly, S.Taumucs(Tmax=300, δ=3.0, tdump=100))
    h = S.run_dropletpatterned(sysc, "GPU", radius=35, θs=th, verbos=False)
    h0 = S.singledroplet(np.ones((Lx, Ly), order="F"), 35, 1 / 6, (75, 75))
    ost, _ = _oracle_run(h0, O.params(delta=3.0, tdump=100), 300, theta_field=th)
    assert np.isfinite(ost.height).all()
    assert np.array_equal(np.asarray(h), ost.height)


def test_dropletforced_driver_matches_oracle(S):
    """inclination! each step (src/simulate.jl:475-481, fac = 1 at t = 1000)."""
    sysc = S.SysConst(150, 150, S.Taumucs(Tmax=300, δ=2.0, tdump=100))
    h, u, v = S.run_dropletforced(sysc, "GPU", radius=35, fx=1e-4, fy=-2e-5, verbos=False)
    h0 = S.singledroplet(np.ones((150, 150), order="F"), 35, 1 / 6, (75, 75))
    ost, _ = _oracle_run(h0, O.params(delta=2.0, tdump=100), 300, body=(1e-4, -2e-5, 1.0))
    assert np.array_equal(np.asarray(h), ost.height)
    assert np.array_equal(np.asarray(u), ost.velx)
    assert np.array_equal(np.asarray(v), ost.vely)


# ---------------------------------------------------------------- the state's fields write through


def test_state_fields_write_through(S):
    """The reference's in-place statements on a state, against the same
    statements on host arrays: whole-field fills, node writes, slices,
    in-place arithmetic, field-to-field assignment and np.copyto, each seen
    by the next op; and a field read after an op sees the op's result."""
    Lx, Ly = 37, 29
    sysc = S.SysConst(Lx, Ly, S.Taumucs(Tmax=20, tdump=10))
    for T in (np.float64, np.float32):
        st = S.Sys(sysc, "GPU", T=T)
        ref = np.ones((Lx, Ly), order="F")
        st.height[...] = 1.0                       # state.height .= 1.0
        for i in range(1, Lx + 1):                 # state.height[i, j] = …
            for j in range(1, Ly + 1, 3):
                v = 1.0 + 0.01 * math.sin(0.3 * i) * math.cos(0.2 * j)
                st.height[i - 1, j - 1] = v
                ref[i - 1, j - 1] = v
        st.height[3:9, 2] = 1.02                   # a slice
        ref[3:9, 2] = 1.02
        st.height *= 1.001                         # in-place arithmetic
        ref *= 1.001
        h = st.height                              # a view kept across an op
        tol = 0 if T is np.float64 else 1e-7
        assert np.allclose(np.asarray(h), ref, rtol=tol, atol=0)
        S.equilibrium(st, sysc)                    # the op sees every write
        feq_ref, _ = O.equilibrium(ref, np.zeros_like(ref), np.zeros_like(ref), 0.0)
        assert np.allclose(st.get("feq", np.float64), feq_ref, rtol=tol, atol=tol)
        st.ftemp[...] = st.feq                     # state.ftemp .= state.feq
        assert np.array_equal(st.get("ftemp"), st.get("feq"))
        np.copyto(st.Fx, 0.25)
        st.Fy[:, :] = np.full((Lx, Ly), -0.5)
        assert np.all(st.get("Fx") == 0.25) and np.all(st.get("Fy") == -0.5)
        st.Fx += st.Fy                             # field-to-field in place
        assert np.all(st.get("Fx") == -0.25)
        S.time_loop(sysc, st)
        assert np.array_equal(np.asarray(h), st.get("height"))  # the kept view is live
        st.close()
