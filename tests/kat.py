"""Runs the reference's known-answer tests (tests/golden/reference_kat.json)
against a backend: the CPU oracle or the HIP library.

A backend exposes numpy-in / numpy-out versions of the 2D ops (arrays of
shape (Lx, Ly) or (Lx, Ly, 9), Julia index order).  Comparisons follow the
reference test: "eq" is bit-exact ==, "approx" uses the reference's atol, or
Julia's default isapprox (rtol = sqrt(eps)) when the reference gives none.
"""
import json
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
RTOL_DEFAULT = math.sqrt(np.finfo(float).eps)


def load_cases():
    with open(os.path.join(HERE, "golden", "reference_kat.json")) as fh:
        return json.load(fh)["cases"]


def _approx(a, b, atol=None):
    a = np.asarray(a, dtype=float)
    b = np.asarray(b, dtype=float)
    if atol is None:
        # Julia isapprox: |a-b| <= rtol * max(|a|, |b|)
        return bool(np.all(np.abs(a - b) <= RTOL_DEFAULT * np.maximum(np.abs(a), np.abs(b))))
    return bool(np.all(np.abs(a - b) <= atol))


def _check(got, want, cmp, atol=None, what=""):
    got = np.asarray(got, dtype=float)
    want = np.broadcast_to(np.asarray(want, dtype=float), got.shape)
    if cmp == "eq":
        assert np.array_equal(got, want), f"{what}: not bit-exact\n got={got}\nwant={want}"
    else:
        assert _approx(got, want, atol), f"{what}: not approx (atol={atol})\n got={got}\nwant={want}"


def _planes(spec):
    if spec == "all":
        return list(range(9))
    if "-" in spec:
        a, b = spec.split("-")
        return list(range(int(a), int(b) + 1))
    return [int(spec)]


def run_case(B, case):
    op = case["op"]
    if op == "grad":
        f = np.array(case["f"])
        a = None if case["a"] is None else np.full(f.shape, case["a"])
        gx, gy = B.grad(f, a)
        _check(gx, case["expect_x"], case["cmp"], what="∂x")
        _check(gy, case["expect_y"], case["cmp"], what="∂y")
    elif op == "laplacian":
        out = B.laplacian(np.array(case["f"]), case["gamma"])
        _check(out, case["expect"], case["cmp"], case.get("atol"), "∇²f")
    elif op == "filmpressure":
        h = np.array(case["h"])
        for form in case["forms"]:
            p = B.filmpressure(h, case["gamma"], case["theta"], case["n"], case["m"], case["hmin"],
                               case["hcrit"], raw=(form == "raw"))
            if "scale" in case:
                _check(case["scale"] * p, case["expect_scaled"], case["cmp"], case.get("atol"),
                       f"p ({form})")
            else:
                _check(p, case["expect"], case["cmp"], case.get("atol"), f"p ({form})")
    elif op == "power_broad":
        for x, n, want in case["items"]:
            got = B.power_broad(x, n)
            assert got == want, (x, n, got, want)
    elif op == "hgradp":
        gx, gy = B.hgradp(np.array(case["p"]), np.array(case["h"]))
        _check(gx, case["expect_x"], "eq", what="h∂xp")
        _check(gy, case["expect_y"], "eq", what="h∂yp")
    elif op == "hgradp_rows":
        gx, _ = B.hgradp(np.array(case["p"]), np.array(case["h"]))
        for row, want in case["rows"].items():
            _check(gx[int(row), :], want, "eq", what=f"h∂xp row {row}")
    elif op == "slippage":
        shape = tuple(case["shape"])
        sx, sy = B.slippage(np.full(shape, case["h"]), np.full(shape, case["ux"]),
                            np.full(shape, case["uy"]), case["delta"], case["mu"])
        for got, want, lab in ((sx, case["expect_x"], "slipx"), (sy, case["expect_y"], "slipy")):
            if want == 0.0:
                _check(got, 0.0, "eq", what=lab)
            else:
                _check(got, want, "approx", case["atol"], lab)
    elif op == "thermal_stats":
        shape = tuple(case["shape"])
        h = np.ones(shape)
        for i, kbt in enumerate(case["kbts"]):
            kx, ky = B.thermal(h, kbt, case["mu"], case["delta"], seed=1234 + i, step=1)
            var = 2 * kbt / 11
            for arr in (kx, ky):
                assert abs(arr.mean()) <= case["mean_atol"]
                assert abs(arr.var(ddof=1) - var) <= var * case["var_rel"]
    elif op == "inclination":
        shape = tuple(case["shape"])
        ax, ay = case["alpha"]
        for t, want in case["steps"]:
            fac = 0.5 + 0.5 * math.tanh((t - 0) / 1)
            Fx, Fy = B.inclination(np.zeros(shape), np.zeros(shape), np.full(shape, case["h"]),
                                   ax, ay, fac)
            _check(Fx, want, "eq", what=f"Fx t={t}")
            _check(Fy, want, "eq", what=f"Fy t={t}")
    elif op == "equilibrium":
        shape = tuple(case["shape"])
        for it in case["items"]:
            feq, _ = B.equilibrium(np.full(shape, it["h"]), np.full(shape, it["ux"]),
                                   np.full(shape, it["uy"]), it["g"])
            for spec, want in it["exact"].items():
                for k in _planes(spec):
                    _check(feq[:, :, k], want, "eq", what=f"feq[{k}] {it}")
            for spec, want in it["approx"].items():
                for k in _planes(spec):
                    _check(feq[:, :, k], want, "approx", it.get("atol"), f"feq[{k}] {it}")
    elif op == "bgk_and_stream":
        for it in case["items"]:
            feq = np.array(it["feq"])
            ftemp = np.full(feq.shape, it["ftemp"])
            Fx = np.full(feq.shape[:2], it["Fx"])
            Fy = np.full(feq.shape[:2], it["Fy"])
            fout, ft = B.bgk_and_stream(feq, ftemp, Fx, Fy, it["tau"])
            assert np.array_equal(fout, ft), "fout .= ftemp must leave both equal"
            if "expect" in it:
                for k in range(9):
                    _check(fout[:, :, k], it["expect"][k], it["cmp"][k], None,
                           f"fout[{k}] τ={it['tau']} F=({it['Fx']},{it['Fy']})")
            else:
                for k, want in it["expect_plane"].items():
                    _check(fout[:, :, int(k)], want, "eq", what=f"fout[{k}]")
    elif op == "moments":
        shape = tuple(case["shape"])
        f = np.zeros(shape + (9,))
        for it in case["items"]:
            for k, v in it["set"].items():
                f[:, :, int(k)] = v
            h, ux, uy = B.moments(f)
            _check(h, it["h"], "eq", what=f"h {it['set']}")
            for arr, key in ((ux, "ux"), (uy, "uy")):
                if key in it:
                    cmp = "eq" if it[key + "_cmp"] == "eq" else "approx"
                    _check(arr, it[key], cmp, case["atol"] if cmp == "approx" else None,
                           f"{key} {it['set']}")
    elif op == "run_flat":
        h = B.run_flat(case["Lx"], case["Ly"], case["Tmax"], case["tdump"])
        assert np.all(h == case["expect_h"])
        assert float(np.sum(h)) == case["expect_sum"]
    elif op == "singledroplet":
        h = B.singledroplet(np.ones(tuple(case["shape"])), case["radius"], case["theta"],
                            tuple(case["center"]))
        want = case["radius"] * (1 - B.cospi(case["theta"]))
        assert h.max() == want
        i, j = np.unravel_index(np.argmax(h), h.shape)
        # Julia's 1-based CartesianIndex(center) is 0-based (center - 1)
        assert (i + 1, j + 1) == tuple(case["center"])
    else:
        raise KeyError(op)
