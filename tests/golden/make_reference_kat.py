"""Writes reference_kat.json: the known-answer tests of Swalbe.jl's 2D hot path,
transcribed from the reference's own test files and doctests.

Each case records its source (file:line in the reference), its inputs, the
expected output exactly as the reference test states it, and the comparison
the reference uses ("eq" for ==, "approx" for isapprox with the given atol,
or Julia's default rtol = sqrt(eps) when atol is absent).  Matrices are
written as Julia displays them (row i, column j), so numpy's [i][j] indexing
matches Julia's [i, j].  Where the reference writes the expected value as a
formula (circshift of a plane, a closed form in f), the formula is evaluated
here with numpy (np.roll is Julia's circshift: out[i] = in[i - s]).

Run:  python tests/golden/make_reference_kat.py
"""
import json
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
R = lambda a: np.asarray(a, dtype=float).tolist()  # noqa: E731

F25 = np.arange(1.0, 26.0).reshape(5, 5, order="F")  # reshape(collect(1.0:25), 5, 5)

SOLX = [[-1.5] * 5, [1.0] * 5, [1.0] * 5, [1.0] * 5, [-1.5] * 5]
SOLY = [[-7.5, 5.0, 5.0, 5.0, -7.5]] * 5
LAP = [[-30.0, -5.0, -5.0, -5.0, 20.0],
       [-25.0, 0.0, 0.0, 0.0, 25.0],
       [-25.0, 0.0, 0.0, 0.0, 25.0],
       [-25.0, 0.0, 0.0, 0.0, 25.0],
       [-20.0, 5.0, 5.0, 5.0, 30.0]]
DOCP = [[30.0, 5.0, 5.0, 5.0, -20.0],
        [25.0, 0.0, 0.0, 0.0, -25.0],
        [25.0, 0.0, 0.0, 0.0, -25.0],
        [25.0, 0.0, 0.0, 0.0, -25.0],
        [20.0, -5.0, -5.0, -5.0, -30.0]]

C = [(0, 0), (1, 0), (0, 1), (-1, 0), (0, -1), (1, 1), (-1, 1), (-1, -1), (1, -1)]


def roll(a, s):
    return np.roll(np.roll(a, s[0], axis=0), s[1], axis=1)


cases = []

# ---------------------------------------------------------------- stencils
cases.append(dict(name="grad_3arg", op="grad", source="test/differences.jl:6-32; src/differences.jl:112-144",
                  f=R(F25), a=None, expect_x=SOLX, expect_y=SOLY, cmp="eq"))
cases.append(dict(name="grad_4arg", op="grad", source="test/differences.jl:33-53",
                  f=R(F25), a=0.1, expect_x=R(0.1 * np.array(SOLX)), expect_y=R(0.1 * np.array(SOLY)),
                  cmp="eq"))
cases.append(dict(name="laplacian", op="laplacian", source="test/differences.jl:104-124; src/differences.jl:26-48",
                  f=R(F25), gamma=-1.0, expect=LAP, cmp="approx", atol=1e-10))

# ---------------------------------------------------------------- pressure
SOLP = np.array(LAP)
cases.append(dict(name="filmpressure_no_angle", op="filmpressure", source="test/pressure.jl:3-33",
                  h=R(F25), gamma=1.0, theta=0.0, n=3, m=2, hmin=0.1, hcrit=0.1,
                  forms=["state", "raw"], expect=R(SOLP), cmp="approx", atol=1e-10))
cases.append(dict(name="filmpressure_angle_half", op="filmpressure", source="test/pressure.jl:35-43",
                  h=R(F25), gamma=1.0, theta=0.5, n=3, m=2, hmin=0.1, hcrit=0.0,
                  forms=["state", "raw"],
                  expect=R(-1 * (-SOLP + 20 * ((0.1 / F25) ** 3 - (0.1 / F25) ** 2))),
                  cmp="approx", atol=1e-10))
cases.append(dict(name="filmpressure_flat", op="filmpressure", source="test/pressure.jl:44-53",
                  h=R(np.ones((5, 5))), gamma=1.0, theta=0.5, n=3, m=2, hmin=0.1, hcrit=0.0,
                  forms=["state", "raw"], expect=R(np.full((5, 5), -2 * (0.1 ** 2 - 0.1))),
                  cmp="approx", atol=1e-10))
cases.append(dict(name="filmpressure_doctest", op="filmpressure", source="src/pressure.jl:39-63",
                  h=R(F25), gamma=0.01, theta=0.0, n=3, m=2, hmin=0.1, hcrit=0.05,
                  forms=["raw"], expect=R(np.array(DOCP) / -100.0), scale=-100.0, expect_scaled=DOCP,
                  cmp="approx", atol=1e-12))
cases.append(dict(name="power_broad", op="power_broad", source="test/pressure.jl:139-161; src/pressure.jl:348-358",
                  items=[[2.0, 3, 8.0], [2.0, 6, 64.0], [5.0, 3, 125.0], [5.0, 6, 15625.0],
                         [3.0, 3, 27.0], [4.0, 3, 64.0], [3.0, 6, 729.0], [4.0, 6, 4096.0],
                         [6.0, 2, 36.0]], cmp="eq"))

# ---------------------------------------------------------------- forces
cases.append(dict(name="hgradp_5x5", op="hgradp", source="test/forcing.jl:104-139",
                  p=R(F25), h=R(np.ones((5, 5))), expect_x=SOLX, expect_y=SOLY, cmp="eq"))
P100 = np.arange(1.0, 101.0).reshape(10, 10, order="F")
cases.append(dict(name="hgradp_doctest_10x10", op="hgradp_rows", source="src/forcing.jl:144-158",
                  p=R(P100), h=R(np.ones((10, 10))), rows={"0": -4.0, "1": 1.0}, cmp="eq"))
for ux, uy, delta, ex, ey, src in [
        (0.0, 0.0, 1.0, 0.0, 0.0, "test/forcing.jl:13-22"),
        (0.1, 0.0, 1.0, 0.1 / 11, 0.0, "test/forcing.jl:24-33"),
        (0.0, 0.1, 1.0, 0.0, 0.1 / 11, "test/forcing.jl:35-46"),
        (-0.1, 0.1, 1.0, -0.1 / 11, 0.1 / 11, "test/forcing.jl:48-58"),
        (-0.1, 0.1, 0.0, -0.1 / 2, 0.1 / 2, "test/forcing.jl:60-68")]:
    cases.append(dict(name=f"slippage_{ux}_{uy}_{delta}", op="slippage", source=src,
                      h=1.0, ux=ux, uy=uy, delta=delta, mu=1 / 6, shape=[5, 5],
                      expect_x=ex, expect_y=ey, cmp="approx", atol=1e-10,
                      exact_zero=True))
cases.append(dict(name="thermal_stats", op="thermal_stats", source="test/forcing.jl:141-164; src/forcing.jl:270-287",
                  shape=[50, 50], kbts=[0.01, 0.1], mu=1 / 6, delta=1.0,
                  mean_atol=1e-2, var_rel=0.1))
cases.append(dict(name="inclination", op="inclination", source="test/forcing.jl:166-184",
                  alpha=[0.1, 0.1], h=1.0, shape=[5, 5],
                  steps=[[0, 0.05], [1, 0.1 * (0.5 + 0.5 * math.tanh(1.0))]], cmp="eq"))

# ---------------------------------------------------------------- equilibrium
eq = []
eq.append(dict(h=0.0, ux=0.0, uy=0.0, g=0.0, exact={"all": 0.0}, approx={}))
eq.append(dict(h=1.0, ux=0.0, uy=0.0, g=0.0, exact={"0": 1.0, "1-8": 0.0}, approx={}))
eq.append(dict(h=1.0, ux=0.0, uy=0.0, g=0.1, exact={"0": 1.0 - 1 / 12},
               approx={"1-4": 1 / 9 * 1.5 * 0.1, "5-8": 1 / 36 * 1.5 * 0.1}))
eq.append(dict(h=1.0, ux=0.1, uy=-0.1, g=0.0, exact={"0": 1.0 - 2 / 3 * 0.02},
               approx={"1": 1 / 9 * (3 * 0.1 + 4.5 * 0.01 - 3 / 2 * 0.02),
                       "2": 1 / 9 * (3 * -0.1 + 4.5 * 0.01 - 3 / 2 * 0.02),
                       "3": 1 / 9 * (3 * -0.1 + 4.5 * 0.01 - 3 / 2 * 0.02),
                       "4": 1 / 9 * (3 * 0.1 + 4.5 * 0.01 - 3 / 2 * 0.02),
                       "5": 1 / 36 * (-3 / 2 * 0.02),
                       "6": 1 / 36 * (3 * -0.2 + 4.5 * 0.2 ** 2 - 3 / 2 * 0.02),
                       "7": 1 / 36 * (-3 / 2 * 0.02),
                       "8": 1 / 36 * (3 * 0.2 + 4.5 * 0.2 ** 2 - 3 / 2 * 0.02)}))
eq.append(dict(h=1.0, ux=0.1, uy=-0.1, g=0.1, exact={"0": 1.0 - 1 / 12 - 2 / 3 * 0.02},
               approx={"1": 1 / 9 * (1.5 * 0.1 + 3 * 0.1 + 4.5 * 0.01 - 3 / 2 * 0.02),
                       "2": 1 / 9 * (1.5 * 0.1 + 3 * -0.1 + 4.5 * 0.01 - 3 / 2 * 0.02),
                       "3": 1 / 9 * (1.5 * 0.1 + 3 * -0.1 + 4.5 * 0.01 - 3 / 2 * 0.02),
                       "4": 1 / 9 * (1.5 * 0.1 + 3 * 0.1 + 4.5 * 0.01 - 3 / 2 * 0.02),
                       "5": 1 / 36 * (1.5 * 0.1 + -3 / 2 * 0.02),
                       "6": 1 / 36 * (1.5 * 0.1 + 3 * -0.2 + 4.5 * 0.2 ** 2 - 3 / 2 * 0.02),
                       "7": 1 / 36 * (1.5 * 0.1 + -3 / 2 * 0.02),
                       "8": 1 / 36 * (1.5 * 0.1 + 3 * 0.2 + 4.5 * 0.2 ** 2 - 3 / 2 * 0.02)}))
# doctest: feq[:,:,1] displays 0.91; second call ≈ 1 - 2/3*0.01
eq.append(dict(h=1.0, ux=0.1, uy=0.0, g=0.1, exact={}, approx={"0": 0.91}, atol=5e-7))  # compact display, 6 digits
eq.append(dict(h=1.0, ux=0.1, uy=0.0, g=0.0, exact={}, approx={"0": 1 - 2 / 3 * 0.01}))
cases.append(dict(name="equilibrium", op="equilibrium",
                  source="test/equilibrium.jl:1-85; src/equilibrium.jl:35-54",
                  shape=[5, 5], items=eq))

# ---------------------------------------------------------------- collide + stream
omega = 1.0 - 1.0 / 0.75
onebytau = 1.0 / 0.75
feq_a = np.ones((5, 5, 9))
feq_a[0, 0, :] = 2.0
p0 = feq_a[:, :, 0]
coll = []
# τ = 1, no force (test/collide.jl:27-43)
coll.append(dict(tau=1.0, Fx=0.0, Fy=0.0, feq=R(feq_a), ftemp=1.0,
                 expect=[R(roll(p0, C[k])) for k in range(9)], cmp=["eq"] * 9))
# τ = 1, forces (0.1, -0.1), feq all ones (test/collide.jl:44-71)
ones = np.ones((5, 5))
add = [0.0, 1 / 30, -1 / 30, -1 / 30, 1 / 30, 0.0, -1 / 24 * 0.2, 0.0, 1 / 24 * 0.2]
exp = []
for k in range(9):
    if k in (0, 5, 7):
        exp.append(R(roll(ones, C[k])))
    elif k in (1, 4):
        exp.append(R(roll(ones + 1 / 30, C[k])))
    elif k in (2, 3):
        exp.append(R(roll(ones - 1 / 30, C[k])))
    elif k == 6:
        exp.append(R(roll(ones - 1 / 24 * 0.2, C[k])))
    else:
        exp.append(R(roll(ones + 1 / 24 * 0.2, C[k])))
coll.append(dict(tau=1.0, Fx=0.1, Fy=-0.1, feq=R(np.ones((5, 5, 9))), ftemp=1.0, expect=exp,
                 cmp=["eq"] * 9))
# τ = 0.75, no forces (test/collide.jl:72-97); plane 2 uses ≈
base = omega * 1.0 + onebytau * p0
coll.append(dict(tau=0.75, Fx=0.0, Fy=0.0, feq=R(feq_a), ftemp=1.0,
                 expect=[R(base)] + [R(roll(base, C[k])) for k in range(1, 9)],
                 cmp=["eq", "approx"] + ["eq"] * 7))
# τ = 0.75, forces (test/collide.jl:99-123)
exp = [R(base)]
for k in range(1, 9):
    if k in (1, 4):
        v = omega * 1.0 + onebytau * p0 + 1 / 30
    elif k in (2, 3):
        v = omega * 1.0 + onebytau * p0 - 1 / 30
    elif k in (5, 7):
        v = omega * 1.0 + onebytau * p0
    elif k == 6:
        v = omega * 1.0 + onebytau * p0 - 1 / 24 * 0.2
    else:
        v = omega * 1.0 + onebytau * p0 + 1 / 24 * 0.2
    exp.append(R(roll(v, C[k])))
coll.append(dict(tau=0.75, Fx=0.1, Fy=-0.1, feq=R(feq_a), ftemp=1.0, expect=exp, cmp=["eq"] * 9))
# doctest src/collide.jl:33-60: feq ones with [1,1,:]=2, τ=1: plane 6 has the 2 at [2,2]
dt = np.ones((5, 5))
dt[1, 1] = 2.0
coll.append(dict(tau=1.0, Fx=0.0, Fy=0.0, feq=R(feq_a), ftemp=0.0, expect_plane={"5": R(dt)},
                 cmp=["eq"]))
cases.append(dict(name="bgk_and_stream", op="bgk_and_stream", source="test/collide.jl:1-123; src/collide.jl:33-60",
                  items=coll))

# ---------------------------------------------------------------- moments
mom = [
    dict(set={"0": 1.0}, h=1.0, h_cmp="eq"),
    dict(set={"1": 0.1}, h=1.1, h_cmp="eq", ux=0.1 / 1.1, ux_cmp="eq", uy=0.0, uy_cmp="eq"),
    dict(set={"2": 0.2}, h=1.3, h_cmp="eq", ux=0.1 / 1.3, ux_cmp="eq", uy=0.2 / 1.3, uy_cmp="eq"),
    dict(set={"3": -0.2}, h=1.1, h_cmp="eq", ux=0.3 / 1.1, ux_cmp="atol", uy=0.2 / 1.1, uy_cmp="atol"),
    dict(set={"4": -0.1}, h=1.0, h_cmp="eq", ux=0.3 / 1.0, ux_cmp="atol", uy=0.3 / 1.0, uy_cmp="atol"),
    dict(set={"5": 0.1}, h=1.1, h_cmp="eq", ux=0.4 / 1.1, ux_cmp="atol", uy=0.4 / 1.1, uy_cmp="atol"),
]
cases.append(dict(name="moments_sequence", op="moments", source="test/moments.jl:1-135; src/moments.jl:18-34",
                  shape=[5, 5], cumulative=True, items=mom, atol=1e-10))

# ---------------------------------------------------------------- drivers
cases.append(dict(name="run_flat", op="run_flat", source="test/simulate.jl:3-15",
                  Lx=25, Ly=25, Tmax=200, tdump=100, expect_h=1.0, expect_sum=625.0, cmp="eq"))
cases.append(dict(name="singledroplet", op="singledroplet", source="test/initialvalues.jl:16-34; src/initialvalues.jl:203-224",
                  shape=[100, 100], radius=50, theta=1 / 3, center=[50, 50]))

doc = {"reference": "Zitzeronion/Swalbe.jl v1.0.6", "note": __doc__.strip().splitlines()[0],
       "cases": cases}
with open(os.path.join(HERE, "reference_kat.json"), "w") as fh:
    json.dump(doc, fh, indent=1)
print("wrote", len(cases), "cases")
