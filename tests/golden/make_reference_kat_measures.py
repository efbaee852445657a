"""Writes reference_kat_measures.json: the known answers of Swalbe.jl's
measurement functions (src/measures.jl), transcribed from test/measures.jl.

Arrays are written column-major flattened (Julia's vec) with their shape;
indices in the notes are the reference's 1-based ones.

Run:  python tests/golden/make_reference_kat_measures.py
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
F = lambda a: np.asarray(a, dtype=float).flatten(order="F").tolist()  # noqa: E731
cases = []

# test/measures.jl:2-9: height[2:3, 2:3] = 1 on 5x5 zeros -> area 4, max 1
h = np.zeros((5, 5))
h[1:3, 1:3] = 1.0
cases.append(dict(name="wetted_block", op="wetted", source="test/measures.jl:2-9", shape=[5, 5],
                  steps=[dict(h=F(h), area=4.0, maxh=1.0)]))
# test/measures.jl:11-23: height[t, t] = t accumulated, t = 1..5 -> area t, max t
h = np.zeros((5, 5))
steps = []
for t in range(1, 6):
    h[t - 1, t - 1] = t
    steps.append(dict(h=F(h), area=float(t), maxh=float(t)))
cases.append(dict(name="wetted_diagonal", op="wetted", source="test/measures.jl:11-23",
                  shape=[5, 5], steps=steps))

# test/measures.jl:26-34: fluid_dry! of the block: vec positions 7, 8 (1-based) wet
h = np.zeros((5, 5))
h[1:3, 1:3] = 1.0
mask = np.zeros(25)
mask[[6, 7, 11, 12]] = 1.0  # vec(dummy) of the 2x2 block at rows/cols 2:3
cases.append(dict(name="fluid_dry_block", op="fluid_dry", source="test/measures.jl:26-34",
                  shape=[5, 5], steps=[dict(h=F(h), mask=mask.tolist())],
                  checked="fluid[1, 1:6] == false and fluid[1, 7:8] == true (the test's assertions); "
                          "the rest of the mask follows from dummy[height .> hthresh] .= true"))
# test/measures.jl:36-52: the sol2 rows, one per step
sol2 = np.array([
    [1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0],
    [1, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0],
    [1, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0],
    [1, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0],
    [1, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 1]], dtype=float)
h = np.zeros((5, 5))
steps = []
for t in range(1, 6):
    h[t - 1, t - 1] = t
    steps.append(dict(h=F(h), mask=sol2[t - 1].tolist()))
cases.append(dict(name="fluid_dry_diagonal", op="fluid_dry", source="test/measures.jl:36-52",
                  shape=[5, 5], steps=steps))

# test/measures.jl:55-97: height = reshape(1:25, 5, 5); the ∇f_simple! gradients
Xgrad = np.array([[-1.5] * 5, [1.0] * 5, [1.0] * 5, [1.0] * 5, [-1.5] * 5])
Ygrad = np.array([[-7.5, 5.0, 5.0, 5.0, -7.5]] * 5)
hs = np.arange(1.0, 26.0).reshape((5, 5), order="F")
solsurf = float(np.sqrt(Xgrad ** 2 + Ygrad ** 2 + 1).sum())
cases.append(dict(name="grad_simple", op="grad_simple", source="test/measures.jl:67-78",
                  shape=[5, 5], f=F(hs), gx=F(Xgrad), gy=F(Ygrad), cmp="eq"))
cases.append(dict(name="surfacearea_half", op="surfacearea", source="test/measures.jl:55-87",
                  shape=[5, 5], h=F(hs), theta_field=F(np.full((5, 5), 0.5)), theta=None,
                  area=solsurf, energy=solsurf - 25 * 0.0, atol=1e-10,
                  note="cospi(0.5) == 0 exactly"))
cases.append(dict(name="surfacearea_ninth_field", op="surfacearea", source="test/measures.jl:88-97",
                  shape=[5, 5], h=F(hs), theta_field=F(np.full((5, 5), 1 / 9)), theta=None,
                  area=solsurf, energy_minus_cos=25, atol=1e-10,
                  note="energy = area - sum(cospi.(θ)): 25 cospi(1/9)"))
cases.append(dict(name="surfacearea_ninth_scalar", op="surfacearea", source="test/measures.jl:88-97",
                  shape=[5, 5], h=F(hs), theta_field=None, theta=1 / 9,
                  area=solsurf, energy_minus_cos=25, atol=1e-10))

# test/measures.jl:117-120
cases.append(dict(name="t0_default", op="t0", source="test/measures.jl:117-120",
                  expect=177428.32340802, rtol=1e-12, note="Julia ≈ default rtol √eps"))

with open(os.path.join(HERE, "reference_kat_measures.json"), "w") as fh:
    json.dump({"reference": "Zitzeronion/Swalbe.jl v1.0.6", "cases": cases}, fh, indent=1)
print(f"{len(cases)} cases")
