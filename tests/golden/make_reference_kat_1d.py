"""Writes reference_kat_1d.json: the known-answer tests of Swalbe.jl's 1D
(D1Q3) path, transcribed from the reference's own test files.

Same conventions as make_reference_kat.py: each case names its source
(file:line in the reference), its inputs, the expected output as the
reference test states it, and the comparison ("eq" for ==, "approx" with the
reference's atol, or Julia's default rtol when atol is absent).  Vectors are
0-based here (Julia's x[1] is x[0]); np.roll is Julia's circshift.

Run:  python tests/golden/make_reference_kat_1d.py
"""
import json
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
R = lambda a: np.asarray(a, dtype=float).tolist()  # noqa: E731

cases = []

# ---------------------------------------------------------------- equilibrium
# test/equilibrium.jl:87-153 (L = 30)
for name, h, u, g, exp0, exp1, exp2, cmps in (
        ("nothing", 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, ("eq", "eq", "eq")),
        ("density", 1.0, 0.0, 0.0, 1.0, 0.0, 0.0, ("eq", "eq", "eq")),
        ("density_gravity", 1.0, 0.0, 0.1, 1.0 - 0.05, 0.025, 0.025, ("eq", "eq", "eq")),
        ("density_velocity", 1.0, 0.1, 0.0, 1.0 - 0.01, 0.5 * 0.1 + 0.5 * 0.01,
         -0.5 * 0.1 + 0.5 * 0.01, ("eq", "approx", "approx")),
        ("density_gravity_velocity", 1.0, 0.1, 0.1, 1.0 - 0.5 * 0.1 - 0.01,
         0.025 + 0.5 * 0.1 + 0.5 * 0.01, 0.025 - 0.5 * 0.1 + 0.5 * 0.01, ("eq", "approx", "approx"))):
    cases.append(dict(name=f"equilibrium_{name}", op="equilibrium", source="test/equilibrium.jl:87-153",
                      L=30, h=h, u=u, g=g, expect=[exp0, exp1, exp2], cmp=list(cmps)))

# ---------------------------------------------------------------- BGK and stream
# test/collide.jl:125-226: feq = 1 with feq[1, :] = 2, ftemp = 1
feq = np.ones((30, 3))
feq[0, :] = 2.0
tau_cases = []
for tau, F in ((1.0, 0.0), (0.75, 0.0), (1.0, 0.1), (0.75, 0.1)):
    onebytau = 1.0 / tau
    omega = 1.0 - 1.0 / tau
    if tau == 1.0:
        base = feq[:, 0].copy()
    else:
        base = omega * 1.0 + onebytau * feq[:, 0]
    e0 = base
    e1 = np.roll(base + 1 / 20 if F else base, 1)
    e2 = np.roll(base - 1 / 20 if F else base, -1)
    cases.append(dict(name=f"bgk_tau{tau}_F{F}", op="bgk_and_stream", source="test/collide.jl:141-216",
                      feq=R(feq), ftemp=1.0, F=F, tau=tau, expect=[R(e0), R(e1), R(e2)], cmp="eq"))

# ---------------------------------------------------------------- moments
# test/moments.jl:13-83: population columns set in sequence
f = np.zeros((30, 3))
f[:, 0] = 1.0
cases.append(dict(name="moments_rest", op="moments", source="test/moments.jl:13-33",
                  f=R(f), expect_h=1.0, expect_u=0.0, cmp="eq"))
f[:, 1] = 0.1
cases.append(dict(name="moments_f1", op="moments", source="test/moments.jl:34-64",
                  f=R(f), expect_h=1.1, expect_u=0.1 / 1.1, cmp="eq"))
f[:, 2] = 0.2
cases.append(dict(name="moments_f2", op="moments", source="test/moments.jl:65-89",
                  f=R(f), expect_h=1.3, expect_u=-0.1 / 1.3, cmp="approx", atol=1e-10))

# ---------------------------------------------------------------- pressure
# test/pressure.jl:56-103, h = 1:30, n = 3, m = 2, γ = 1, hmin = 0.1
f30 = np.arange(1.0, 31.0)
sol = np.zeros(30)
sol[0], sol[-1] = 30.0, -30.0
for form in ("raw", "state"):
    cases.append(dict(name=f"filmpressure_{form}_no_angle", op="filmpressure", form=form,
                      source="test/pressure.jl:76-83", h=R(f30), gamma=1.0, theta=0.0, n=3, m=2,
                      hmin=0.1, hcrit=0.1, expect=R(-sol), cmp="eq"))
    cases.append(dict(name=f"filmpressure_{form}_angle", op="filmpressure", form=form,
                      source="test/pressure.jl:84-91", h=R(f30), gamma=1.0, theta=0.5, n=3, m=2,
                      hmin=0.1, hcrit=0.0,
                      expect=R(-1 * (sol + 20 * ((0.1 / f30) ** 3 - (0.1 / f30) ** 2))),
                      cmp="approx", atol=1e-10))
    cases.append(dict(name=f"filmpressure_{form}_flat", op="filmpressure", form=form,
                      source="test/pressure.jl:93-102", h=1.0, L=30, gamma=1.0, theta=0.5, n=3, m=2,
                      hmin=0.1, hcrit=0.0, expect=-2 * (0.1 ** 2 - 0.1), cmp="approx", atol=1e-10))

# ---------------------------------------------------------------- differences
# test/differences.jl:79-100 and 126-139, f = 1:25
f25 = np.arange(1.0, 26.0)
g = np.ones(25)
g[0] = g[-1] = -11.5
cases.append(dict(name="grad_3arg", op="grad", source="test/differences.jl:90-99", f=R(f25), a=None,
                  expect=R(g), cmp="eq"))
cases.append(dict(name="grad_4arg", op="grad", source="test/differences.jl:81-89", f=R(f25), a=1.0,
                  expect=R(g), cmp="eq"))
lap = np.zeros(25)
lap[0], lap[-1] = 25.0, -25.0
cases.append(dict(name="laplacian", op="laplacian", source="test/differences.jl:126-139", f=R(f25),
                  expect=R(lap), cmp="eq"))

# ---------------------------------------------------------------- forcing
# test/forcing.jl:71-102 (h = 1, μ = 1/6)
for u, delta, exp, cmp in ((0.0, 1.0, 0.0, "eq"), (0.1, 1.0, 0.1 / 11, "approx"),
                           (-0.1, 0.0, -0.1 / 2, "approx")):
    cases.append(dict(name=f"slippage_{u}_{delta}", op="slippage", source="test/forcing.jl:71-102",
                      L=30, h=1.0, u=u, delta=delta, mu=1 / 6, expect=exp, cmp=cmp, atol=1e-10))
# test/forcing.jl:104-139: p = 1:30, h = 1
hgp = np.ones(30)
hgp[0] = hgp[-1] = -14.0
cases.append(dict(name="hgradp", op="hgradp", source="test/forcing.jl:104-139", p=R(f30), h=1.0,
                  expect=R(hgp), cmp="eq"))
# test/forcing.jl:141-164: L = 100000, h = 1
cases.append(dict(name="thermal_stats", op="thermal_stats", source="test/forcing.jl:141-164",
                  L=100000, kbts=[0.01, 0.1], mu=1 / 6, delta=1.0))
# test/forcing.jl:166-184: α = 0.1, tstart = 0, tsmooth = 1
cases.append(dict(name="inclination", op="inclination", source="test/forcing.jl:166-184", L=30,
                  alpha=0.1, facs={"0": 0.5 + 0.5 * math.tanh(0.0), "1": 0.5 + 0.5 * math.tanh(1.0)},
                  expect={"0": 0.05, "1": 0.1 * (0.5 + 0.5 * math.tanh(1.0))}, cmp="eq"))

# ---------------------------------------------------------------- drivers
# test/simulate.jl:9-13: run_flat(SysConst_1D(L = 25, Tmax = 200, tdump = 100))
cases.append(dict(name="run_flat", op="run_flat", source="test/simulate.jl:9-13", L=25, Tmax=200,
                  tdump=100))

with open(os.path.join(HERE, "reference_kat_1d.json"), "w") as fh:
    json.dump({"reference": "Zitzeronion/Swalbe.jl v1.0.6", "cases": cases}, fh, indent=1)
print(f"{len(cases)} cases")
