"""Throughput of the fused thin-film LBM step on one MI355X.

Workload (the headline configuration): thin-film rupture with thermal noise,
D2Q9 on a 4096 x 4096 periodic film, FP32 storage and compute, τ = 1,
θ = 1/9, n = 9, m = 3, hmin = 0.1, hcrit = 0.05, γ = 0.01, δ = 1, kbt = 1e-6
(Philox noise keyed by seed, step, node), h0 = 1 + 0.001 N(0,1) from a seed.

A "step" is one fused time step of the whole film (one kernel launch).
W untimed warmup steps, then K timed steps bracketed by device
synchronisation; the step kernels' time is measured with HIP events on the
stream they are launched on.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "swalbe.jl_amd"))

HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: (Lx, Ly, dtype, thermal)
    "rupture4096_f32": (4096, 4096, "f32", True),
    "rupture4096_f64": (4096, 4096, "f64", True),
    "flat4096_f32_nonoise": (4096, 4096, "f32", False),
    "droplet1024_f32": (1024, 1024, "f32", False),
}


def bytes_per_lup(esz: int, tau_is_one: bool) -> int:
    """Algorithmic HBM bytes per lattice update of the fused step: reads h, ux,
    uy (and the 9 populations f_t when τ ≠ 1), writes the 9 streamed
    populations and h, ux, uy, p."""
    reads = 3 + (0 if tau_is_one else 9)
    writes = 9 + 4
    return (reads + writes) * esz


def cpu_baseline(seconds_target: float = 12.0):
    """The oracle's unfused FP64 step (reference-equivalent sweeps, 1 thread)
    timed on a 4096 x 128 strip of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    Lx, Ly = 4096, 128
    st = O.State(Lx, Ly)
    st.height[...] = O.init_normal(Lx, Ly, 1234, 1.0, 0.001)
    st.feq[...], _ = O.equilibrium(st.height, st.velx, st.vely, 0.0)
    prm = O.params(gamma=0.01, kbt=1e-6, tdump=10 ** 9)
    t = time.perf_counter()
    O.time_loop(st, prm, 1, thermal=True, seed=7)
    one = time.perf_counter() - t
    n = max(2, int(seconds_target / max(one, 1e-3)))
    t = time.perf_counter()
    O.time_loop(st, prm, n, thermal=True, seed=7, t0=1)
    dt = time.perf_counter() - t
    return {"value": Lx * Ly * n / dt / 1e9, "unit": "GLUPS", "cores": 1, "kind": "port",
            "sample": f"{Lx}x{Ly} strip of the same workload, FP64, {n} steps, {dt:.1f} s, "
                      "unfused reference-order sweeps (oracle)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--workload", default="rupture4096_f32", choices=sorted(WORKLOADS))
    ap.add_argument("--tile-y", type=int, default=0)
    ap.add_argument("--variant", type=int, default=0)
    ap.add_argument("--nontemporal", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(the films are independent or a single slab; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import ctypes as C
    import numpy as np
    import swalbe as S
    from swalbe import _lib

    Lx, Ly, dt, thermal = WORKLOADS[a.workload]
    T = np.float32 if dt == "f32" else np.float64
    param = S.Taumucs(Tmax=a.steps, tdump=10 ** 9, kbt=1e-6 if thermal else 0.0, γ=0.01, δ=1.0,
                      n=9, m=3, hmin=0.1, hcrit=0.05, θ=1 / 9)
    sysc = S.SysConst(Lx, Ly, param)
    st = S.Sys(sysc, "GPU", T=T, kind="thermal" if thermal else "simple")
    _lib.check(_lib.lib().swalbe_init_random(st.handle, 1, 1234, 1.0, 0.001), "init")
    S.equilibrium(st, sysc)
    prm = param.cparams()
    o = _lib.StepOpts()
    o.thermal, o.seed, o.tile_y, o.variant = int(thermal), 7, a.tile_y, a.variant
    o.nontemporal = a.nontemporal
    ms = C.c_double()
    L = _lib.lib()
    # warmup
    o.t0 = 0
    _lib.check(L.swalbe_time_loop_timed(st.handle, C.byref(prm), a.warmup, C.byref(o), C.byref(ms)))
    st.synchronize()
    # timed
    o.t0 = a.warmup
    w0 = time.perf_counter()
    _lib.check(L.swalbe_time_loop_timed(st.handle, C.byref(prm), a.steps, C.byref(o), C.byref(ms)))
    st.synchronize()
    wall = time.perf_counter() - w0
    t_s = ms.value / 1e3
    N = Lx * Ly
    glups = N * a.steps / t_s / 1e9
    esz = 4 if dt == "f32" else 8
    bpl = bytes_per_lup(esz, True)
    achieved = bpl * N * a.steps / t_s / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", f"traffic_step_{a.workload}.json")
    if os.path.exists(tf):
        with open(tf) as fh:
            traffic = json.load(fh).get("bytes_per_launch")
    h = st.get("height", np.float64)
    ok = bool(np.isfinite(h).all())
    line = {
        "metric": "GLUPS (lattice updates/s) of the fused FP32 D2Q9 thin-film step at 4096^2",
        "value": round(glups, 3),
        "unit": "GLUPS",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms.value / a.steps, 5),
        "wall_ms_per_step": round(wall * 1e3 / a.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dt,
        "data": "synthetic (seeded Philox film h0 = 1 + 0.001 N(0,1))",
        "config": {"workload": f"{a.workload}: D2Q9 thin-film rupture with thermal noise, "
                               f"{Lx}x{Ly} periodic, kbt=1e-6, theta=1/9, n=9, m=3, gamma=0.01, tau=1",
                   "Lx": Lx, "Ly": Ly, "parallelism": "single GPU", "tile_y": a.tile_y or 16,
                   "variant": a.variant, "nontemporal": a.nontemporal,
                   "plane_pad": int(os.environ.get("SWALBE_PLANE_PAD", "0")), "finite": ok},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "bytes_per_lup": bpl},
    }
    if not a.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
