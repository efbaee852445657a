"""Runs the reference's 1D known-answer tests (tests/golden/reference_kat_1d.json)
against a backend: the CPU oracle or the HIP library.  A backend takes and
returns numpy vectors (L,) and distributions (L, 3)."""
import json
import os

import numpy as np

from kat import _check

HERE = os.path.dirname(os.path.abspath(__file__))


def load_cases():
    with open(os.path.join(HERE, "golden", "reference_kat_1d.json")) as fh:
        return json.load(fh)["cases"]


def _vec(v, L):
    return np.broadcast_to(np.asarray(v, dtype=float), (L,)).copy()


def run_case(be, c):
    op, name = c["op"], c["name"]
    if op == "equilibrium":
        L = c["L"]
        feq = be.equilibrium(_vec(c["h"], L), _vec(c["u"], L), c["g"])
        for k in range(3):
            _check(feq[:, k], c["expect"][k], c["cmp"][k], None, f"{name} f{k}")
    elif op == "bgk_and_stream":
        feq = np.asarray(c["feq"], dtype=float)
        L = feq.shape[0]
        fout = be.bgk_and_stream(feq, np.full((L, 3), c["ftemp"]), _vec(c["F"], L), c["tau"])
        for k in range(3):
            _check(fout[:, k], c["expect"][k], c["cmp"], None, f"{name} f{k}")
    elif op == "moments":
        h, u = be.moments(np.asarray(c["f"], dtype=float))
        _check(h, c["expect_h"], c["cmp"], c.get("atol"), f"{name} h")
        _check(u, c["expect_u"], c["cmp"], c.get("atol"), f"{name} u")
    elif op == "filmpressure":
        L = c.get("L") or len(c["h"])
        h = _vec(c["h"], L)
        p = be.filmpressure(h, c["gamma"], c["theta"], c["n"], c["m"], c["hmin"], c["hcrit"],
                            c["form"])
        _check(p, c["expect"], c["cmp"], c.get("atol"), name)
    elif op == "grad":
        f = np.asarray(c["f"], dtype=float)
        a = None if c["a"] is None else _vec(c["a"], len(f))
        _check(be.grad(f, a), c["expect"], c["cmp"], None, name)
    elif op == "laplacian":
        _check(be.laplacian(np.asarray(c["f"], dtype=float)), c["expect"], c["cmp"], None, name)
    elif op == "slippage":
        L = c["L"]
        s = be.slippage(_vec(c["h"], L), _vec(c["u"], L), c["delta"], c["mu"])
        _check(s, c["expect"], c["cmp"], c.get("atol"), name)
    elif op == "hgradp":
        p = np.asarray(c["p"], dtype=float)
        _check(be.hgradp(p, _vec(c["h"], len(p))), c["expect"], c["cmp"], None, name)
    elif op == "thermal_stats":
        L = c["L"]
        for kb in c["kbts"]:
            z = be.thermal(np.ones(L), kb, c["mu"], c["delta"], 11, 1)
            var = 2 * kb / 11
            assert abs(z.mean()) <= 1e-2, (name, kb, z.mean())
            assert abs(z.var(ddof=1) - var) <= var / 10, (name, kb, z.var())
    elif op == "inclination":
        L = c["L"]
        for t, fac in c["facs"].items():
            F = be.inclination(np.zeros(L), np.ones(L), c["alpha"], fac)
            _check(F, c["expect"][t], c["cmp"], None, f"{name} t={t}")
    elif op == "run_flat":
        h = be.run_flat(c["L"], c["Tmax"], c["tdump"])
        assert np.all(h == 1.0), name
        assert h.sum() == c["L"], name
    else:
        raise ValueError(op)
