"""Throughput of the fused thin-film LBM step on one MI355X.

Workload (the headline configuration): thin-film rupture with thermal noise,
D2Q9 on a 4096 x 4096 periodic film, FP32 storage and compute, τ = 1,
θ = 1/9, n = 9, m = 3, hmin = 0.1, hcrit = 0.05, γ = 0.01, δ = 1, kbt = 1e-6
(Philox noise keyed by seed, step, node), h0 = 1 + 0.001 N(0,1) from a seed.

A "step" is one fused time step of the whole film (one kernel launch).
The live ceilings and about a second of untimed steps run first (the film is
then re-initialised), then W untimed warmup steps, then K timed steps
bracketed by device synchronisation; the step kernels' time is measured with
HIP events on the stream they are launched on.  The same line carries the
second half of the headline metric: continuous-batched serving of Poisson
solve requests (solves/s, p50/p99 latency, the speed-up over one-at-a-time
solves and its mix-bound ceiling, a 5 % FP64 sample), and the 256-film batch
(one launch per step for all films).  Prints one JSON line.

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
    "flat4096_f64_nonoise": (4096, 4096, "f64", False),
    "droplet1024_f32": (1024, 1024, "f32", False),
    "droplet1024_f64": (1024, 1024, "f64", False),
}
# workloads that start from the sessile drop of config 2 (cap sphere radius
# 150 at the centre, θ = 1/9, precursor 0.1) instead of a perturbed flat film
DROPLET = ("droplet1024_f32", "droplet1024_f64")
# many-film workloads (FP32 batched kernel), the 1D coalescence scan and the copy calibration
BATCH_WORKLOADS = ("coalesce256_batch", "serve_poisson", "copy4096_f32", "pattern4096_f32",
                   "coalesce1d_scan", "single1d_f64", "long1d_f64")


def bytes_per_lup(esz: int, tau_is_one: bool) -> int:
    """Algorithmic HBM bytes per lattice update of the fused step: reads h, ux,
    uy (and the 9 populations f_t when τ ≠ 1), writes the 9 streamed
    populations and h, ux, uy, p."""
    reads = 3 + (0 if tau_is_one else 9)
    writes = 9 + 4
    return (reads + writes) * esz


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for ln in fh:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


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
                      f"unfused reference-order sweeps (oracle), {cpu_model()}"}


def cpu_baseline_fused(seconds_target: float = 8.0):
    """The same step as three fused OpenMP sweeps on all host threads
    (oracle/swalbe_cpu_fused.c, bit-identical to the oracle), FP64, on a
    4096 x 512 strip of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    Lx, Ly = 4096, 512
    st = O.State(Lx, Ly)
    st.height[...] = O.init_normal(Lx, Ly, 1234, 1.0, 0.001)
    st.feq[...], _ = O.equilibrium(st.height, st.velx, st.vely, 0.0)
    prm = O.params(gamma=0.01, kbt=1e-6, tdump=10 ** 9)
    O.fused_time_loop(st, prm, 2, thermal=True, seed=7)  # threads and pages warm up
    t = time.perf_counter()
    O.fused_time_loop(st, prm, 4, thermal=True, seed=7, t0=2)
    one = (time.perf_counter() - t) / 4
    n = max(4, int(seconds_target / max(one, 1e-4)))
    t = time.perf_counter()
    O.fused_time_loop(st, prm, n, thermal=True, seed=7, t0=6)
    dt = time.perf_counter() - t
    return {"value": Lx * Ly * n / dt / 1e9, "unit": "GLUPS", "cores": O.fused_threads(),
            "kind": "port", "sample": f"{Lx}x{Ly} strip of the same workload, FP64, {n} steps, "
                                      f"{dt:.1f} s, three fused OpenMP sweeps per step, {cpu_model()}"}


def batch_run(steps, warmup, graph_steps=100, autotune=True, final_outputs_only=False,
              nontemporal=-1):
    """256 independent droplet-coalescence films (Lx in {128..1024} step 64,
    Ly = Lx/2, two tangent caps, per-film θ, γ, radius) stepped together by
    the batched FP32 kernel; `steps` timed batch steps after `warmup`, run as
    captured graphs.  Each film's tile height (8 or 16 rows) is the one the
    autotuner measured fastest for its shape (the tuning runs at admission,
    before the timed steps, and is reported).  nontemporal >= 0 sets the
    films' store mode (otherwise the batch's default, 2)."""
    import numpy as np
    import swalbe as S
    from swalbe import serve
    reqs = serve.coalescence_sweep(256, seed=1, nsteps=warmup + steps)
    cap = sum(S.film_bytes(r.Lx, r.Ly) for r in reqs)
    b = S.Batch(cap, 256, graph_steps=graph_steps, autotune=autotune,
                final_outputs_only=final_outputs_only)
    if nontemporal >= 0:
        b.set_nontemporal(nontemporal)
    t0 = time.perf_counter()
    slots = [b.add(r.param, r.Lx, r.Ly, r.init, r.nsteps) for r in reqs]
    b.sync()
    admit_s = time.perf_counter() - t0
    rows = [b.tile_y(sl) for sl in slots]
    shapes = {}
    for r, ty in zip(reqs, rows):
        shapes[f"{r.Lx}x{r.Ly}"] = ty
    st = b.stats()
    b.step(warmup)
    b.sync()
    w0 = time.perf_counter()
    ms = b.step_timed(steps)
    wall = time.perf_counter() - w0
    N = st["nodes"]
    glups = N * steps / (ms / 1e3) / 1e9
    ok = all(bool(np.isfinite(b.get(s, "height")).all()) for s in (0, 127, 255))
    b.close()
    return {"glups": glups, "ms": ms, "wall": wall, "nodes": N, "films": st["films"],
            "cap": cap, "finite": ok, "rows": rows, "shapes": shapes, "admit_s": admit_s,
            "nontemporal": nontemporal if nontemporal >= 0 else 2}


def bench_batch(a):
    """Config 4 on its own line (coalesce256_batch): see batch_run()."""
    r = batch_run(a.steps, a.warmup, a.graph_steps, not a.no_autotune,
                  nontemporal=a.nontemporal)
    bpl = bytes_per_lup(4, True)
    achieved = bpl * r["glups"]
    line = {
        "metric": "GLUPS of 256 batched FP32 droplet-coalescence films (one launch per step)",
        "value": round(r["glups"], 3), "unit": "GLUPS", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(r["ms"] / a.steps, 5),
        "wall_ms_per_step": round(r["wall"] * 1e3 / a.steps, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded two-cap films)",
        "config": {"workload": BATCH_DESC, "films": r["films"], "nodes": r["nodes"],
                   "graph_steps": a.graph_steps, "arena_bytes": r["cap"],
                   "nontemporal": r["nontemporal"],
                   "parallelism": "single GPU", "finite": r["finite"]},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None, "bytes_per_lup": bpl},
        "autotune": {"on": not a.no_autotune, "shapes": len(r["shapes"]),
                     "films_rows8": r["rows"].count(8), "films_rows16": r["rows"].count(16),
                     "rows_by_shape": r["shapes"], "admission_s": round(r["admit_s"], 3)},
    }
    print(json.dumps(line), flush=True)


BATCH_DESC = ("coalesce256_batch: 256 films, Lx uniform in {128..1024} step 64, Ly = Lx/2, "
              "two tangent caps R ~ U(0.15,0.3) Ly, theta/pi ~ U(1/18,1/6), "
              "gamma ~ U(0.005,0.02), tau=1")


# FP64 operations per node and step of the 1D resident kernel in the scan
# configuration (γ field, ∇γ force, n = 9, m = 3, τ = 1), counted from the
# source: pressure 19 + 1 division, force 12 + 1 division, collide and
# stream 33, moments 3 + 1 division (a division counted as one operation)
FLOP_PER_LUP_1D = 70
FP64_PEAK_TFLOPS = 78.6


def _scan_film(S, k, L, steps):
    """Film k of the 1D coalescence scan (analyze/droplet_coalescence.jl):
    two droplets r = 500, θ = 1/9 on L = 1024, γ₀ = 1e-4 with a tanh step of
    20 % and width sl, slip δ, hmin; every other film has walls at the ends."""
    import numpy as np
    sls = [1, 2, 5, 10, 20, 50, 100, 200]
    sl = sls[k % len(sls)]
    delta = (11.0, 15.0)[(k // len(sls)) % 2]
    hmin = (0.1, 0.14)[(k // (2 * len(sls))) % 2]
    g0 = 1e-4
    l = np.arange(1.0, L + 1)
    smooth = np.abs(1 - (0.5 + 0.5 * np.tanh((l - L // 2) / sl)))
    gamma = g0 * smooth + (1 - smooth) * (g0 - g0 * 20 / 100)
    param = S.Taumucs(Tmax=steps, tdump=10 ** 9, γ=g0, δ=delta, n=9, m=3, hmin=hmin, θ=1 / 9)
    if k % 2:
        obs = np.zeros(L)
        obs[:5] = 1
        obs[-5:] = 1
        sysc = S.SysConstWithBound_1D(L=L, param=param, obs=obs)
        S.obslist(sysc)
    else:
        sysc = S.SysConst_1D(L=L, param=param)
    st = S.Sys(sysc, T=np.float64, kind="gamma")
    st.height = S.two_droplets(sysc, r1=500, r2=500, θ1=1 / 9, θ2=1 / 9)
    S.equilibrium(st, sysc)
    st.set("gamma", gamma)
    S.grad_gamma(st)
    gg = st.get("gradgamma")
    gg[-5:] = 0.0
    gg[:4] = 0.0  # state.∇γ[L-4:L] .= 0; state.∇γ[1:4] .= 0
    st.set("gradgamma", gg)
    return sysc, st, gamma


def cpu_baseline_1d(L=1024, seconds_target=10.0):
    """The oracle's FP64 1D step (reference-order sweeps, 1 thread) on one scan film."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as O
    import swalbe as S
    sysc = S.SysConst_1D(L=L, param=S.Taumucs())
    h = S.two_droplets(sysc, r1=500, r2=500, θ1=1 / 9, θ2=1 / 9)
    st = O.State1D(L)
    st.height[...] = h
    st.feq[...] = O.equilibrium_1d(h, np.zeros(L), 0.0)
    l = np.arange(1.0, L + 1)
    smooth = np.abs(1 - (0.5 + 0.5 * np.tanh((l - L // 2) / 10)))
    gamma = 1e-4 * smooth + (1 - smooth) * 0.8e-4
    gg = O.grad_gamma_1d(gamma, form=0)
    prm = O.params(gamma=1e-4, delta=11.0, tdump=10 ** 9)
    kw = dict(gamma_field=gamma, grad_gamma=gg, grad_gamma_sign=1)
    t = time.perf_counter()
    O.time_loop_1d(st, prm, 200, **kw)
    one = (time.perf_counter() - t) / 200
    n = max(200, int(seconds_target / max(one, 1e-7)))
    t = time.perf_counter()
    O.time_loop_1d(st, prm, n, **kw)
    dt = time.perf_counter() - t
    return {"value": L * n / dt / 1e9, "unit": "GLUPS", "cores": 1, "kind": "port",
            "sample": f"one L={L} film of the scan, FP64, {n} steps, {dt:.1f} s, "
                      "unfused reference-order sweeps (oracle)"}


def bench_1d(a):
    """1D coalescence scan: a.films FP64 films of L = 1024 (scan parameters per
    film) stepped together, one workgroup per film, whole films in LDS, K
    steps in one launch (single1d_f64: one film, both loops)."""
    import ctypes as C
    import numpy as np
    import swalbe as S
    from swalbe import _lib
    L = 1024
    nf = 1 if a.workload == "single1d_f64" else a.films
    films = [_scan_film(S, k, L, a.steps) for k in range(nf)]
    systems = [f[0] for f in films]
    states = [f[1] for f in films]
    gammas = [f[2] for f in films]

    def run(nsteps):
        for sy in systems:
            sy.param.Tmax = nsteps
        w0 = time.perf_counter()
        ms = S.ensemble_time_loop(systems, states, nsteps, γ=gammas, grad_gamma_sign=1,
                                  materialize=False, threads=a.threads)
        return ms / 1e3, time.perf_counter() - w0

    run(a.warmup)
    dt, wall = run(a.steps)
    lups = nf * L * a.steps
    glups = lups / dt / 1e9
    line = {
        "metric": "GLUPS of the 1D D1Q3 coalescence scan (films x L x steps / s)",
        "value": round(glups, 3), "unit": "GLUPS", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(dt * 1e3 / a.steps, 5),
        "wall_GLUPS": round(lups / wall / 1e9, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (two droplets r=500, tanh gamma steps of the reference scan)",
        "config": {"workload": f"{a.workload}: {nf} films x L={L}, FP64, gamma field + grad "
                               "gamma force, half the films walled, tau=1, resident kernel",
                   "films": nf, "L": L, "threads_per_film": a.threads or "auto",
                   "parallelism": "single GPU, one workgroup per film"},
        "roofline": {"bound": "fp64-valu", "achieved": round(glups * FLOP_PER_LUP_1D / 1e3, 3),
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(glups * FLOP_PER_LUP_1D / 1e3 / FP64_PEAK_TFLOPS, 4),
                     "traffic": None, "flop_per_lup": FLOP_PER_LUP_1D},
    }
    h = states[0].height
    line["config"]["finite"] = bool(np.isfinite(h).all())
    if a.workload == "single1d_f64":
        # the same film with one launch per step
        st2 = _scan_film(S, 0, L, a.steps)
        o = _lib.StepOpts1D()
        o.variant, o.grad_gamma, o.use_gamma_field = 1, 1, 1
        st2[1].set("gamma", st2[2])
        prm = st2[0].param.cparams()
        ms = C.c_double()
        _lib.check(_lib.lib().swalbe1d_time_loop_timed(st2[1].handle, C.byref(prm), a.steps,
                                                       C.byref(o), C.byref(ms)), "1d loop")
        line["per_step_loop_GLUPS"] = round(L * a.steps / (ms.value / 1e3) / 1e9, 4)
    if not a.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline_1d(L)
    print(json.dumps(line), flush=True)


def bench_long1d(a):
    """One long 1D film (L = 2^24, FP64, periodic, h = 1 + 0.01 N(0,1)), one
    launch per step (k1_step): 64 algorithmic bytes per node update (reads
    h, u; writes f0..f2, h, u, p)."""
    import ctypes as C
    import numpy as np
    import swalbe as S
    from swalbe import _lib
    L = 1 << 24
    sysc = S.SysConst_1D(L=L, param=S.Taumucs(Tmax=a.steps, tdump=10 ** 9))
    st = S.Sys(sysc, T=np.float64)
    z0, _ = S.normals(L, 1234)
    st.height = 1.0 + 0.01 * z0
    S.equilibrium(st, sysc)
    prm = sysc.param.cparams()
    o = _lib.StepOpts1D()
    o.variant, o.materialize, o.threads = 1, 0, a.threads
    ms = C.c_double()
    o.t0 = 0
    _lib.check(_lib.lib().swalbe1d_time_loop_timed(st.handle, C.byref(prm), a.warmup, C.byref(o),
                                                   C.byref(ms)), "1d loop")
    o.t0 = a.warmup
    _lib.check(_lib.lib().swalbe1d_time_loop_timed(st.handle, C.byref(prm), a.steps, C.byref(o),
                                                   C.byref(ms)), "1d loop")
    t = ms.value / 1e3
    glups = L * a.steps / t / 1e9
    achieved = 64 * L * a.steps / t / 1e9
    line = {"metric": "GLUPS of the 1D D1Q3 step on one long film (one launch per step)",
            "value": round(glups, 3), "unit": "GLUPS", "n_gpus": 1, "steps": a.steps,
            "warmup": a.warmup, "ms_per_step": round(ms.value / a.steps, 5),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic (seeded Philox film h0 = 1 + 0.01 N(0,1))",
            "config": {"workload": f"long1d_f64: L = {L}, FP64, periodic, tau = 1, k1_step",
                       "L": L, "finite": bool(np.isfinite(st.height[:4096]).all())},
            "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                         "traffic": None, "bytes_per_lup": 64}}
    print(json.dumps(line), flush=True)


def bench_copy(a):
    """Device-to-device copy of the nine distribution planes of a 4096^2 FP32
    state (feq -> ftemp, 604 MB each way): the achievable HBM bandwidth of a
    plain copy, and the known-size kernel that calibrates the byte counters."""
    import swalbe as S
    sysc = S.SysConst(4096, 4096, S.Taumucs(Tmax=1))
    import numpy as np
    st = S.Sys(sysc, "GPU", T=np.float32)
    for _ in range(max(1, a.warmup)):
        st.copy("ftemp", "feq")
    st.synchronize()
    t = time.perf_counter()
    for _ in range(a.steps):
        st.copy("ftemp", "feq")
    st.synchronize()
    dt = time.perf_counter() - t
    nbytes = 9 * 4096 * 4096 * 4
    gbs = 2 * nbytes * a.steps / dt / 1e9
    print(json.dumps({"metric": "device-to-device copy bandwidth (read + write)",
                      "value": round(gbs, 1), "unit": "GB/s", "n_gpus": 1, "steps": a.steps,
                      "warmup": a.warmup, "ms_per_step": round(dt * 1e3 / a.steps, 5),
                      "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                      "dtype": "f32", "data": "synthetic",
                      "config": {"workload": "copy4096_f32: 9 planes of 4096^2 FP32",
                                 "bytes_per_copy": nbytes}}), flush=True)


def probe_lib():
    """libswalbe_probe.so: the access-pattern probe (a bench tool, built
    beside the library by swalbe.jl_amd/build.py)."""
    import ctypes as C
    L = C.CDLL(os.path.join(ROOT, "swalbe.jl_amd", "lib", "libswalbe_probe.so"))
    L.swalbe_pattern_probe.restype = C.c_int
    L.swalbe_pattern_probe.argtypes = [C.c_int32] * 6 + [C.POINTER(C.c_double)]
    return L


def pattern_ms(Lx, Ly, tile_y, iters):
    """Mean device ms per launch of the probe reading 3 / writing 13 planes."""
    import ctypes as C
    ms = C.c_double()
    rc = probe_lib().swalbe_pattern_probe(Lx, Ly, 3, 13, tile_y, iters, C.byref(ms))
    if rc != 0:
        raise RuntimeError(f"swalbe_pattern_probe failed with status {rc}")
    return ms.value


def ceilings(st, tile_y):
    """What the part allows for this step, measured live beside it: a plain
    device-to-device copy of nine 4096^2 FP32 planes (GB/s, read + write) and
    the step's own access pattern with no arithmetic (read 3 / write 13 planes
    over the same 64 x tile_y tiles, every store streaming, GLUPS)."""
    for _ in range(3):
        st.copy("ftemp", "feq")
    st.synchronize()
    reps = 20
    t = time.perf_counter()
    for _ in range(reps):
        st.copy("ftemp", "feq")
    st.synchronize()
    copy_gbs = 2 * 9 * 4096 * 4096 * 4 * reps / (time.perf_counter() - t) / 1e9
    ms = pattern_ms(4096, 4096, tile_y, 50)
    return {"copy_GBs": round(copy_gbs, 1),
            "pattern_GLUPS": round(4096 * 4096 / (ms / 1e3) / 1e9, 3),
            "pattern": f"read 3 / write 13 planes over 64x{tile_y} tiles, no arithmetic, "
                       "streaming stores"}


def bench_pattern(a):
    """The fused step's data movement alone at 4096^2 FP32: read 3 planes (h,
    ux, uy), write 13 (9 f, h, ux, uy, p) over the step's 64 x 16 tiles, no
    arithmetic, every store streaming: the ceiling the step's access pattern
    allows."""
    L = 4096
    ty = a.tile_y or 16
    ms = pattern_ms(L, L, ty, a.steps)
    t = ms / 1e3
    gbs = 64 * L * L / t / 1e9
    print(json.dumps({"metric": "data-movement ceiling of the fused step's access pattern",
                      "value": round(L * L / t / 1e9, 3), "unit": "GLUPS", "n_gpus": 1,
                      "steps": a.steps, "warmup": 3, "ms_per_step": round(ms, 5),
                      "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                      "dtype": "f32", "data": "synthetic",
                      "config": {"workload": "pattern4096_f32: read 3 / write 13 planes of "
                                             f"4096^2 FP32 over 64x{ty} tiles, no arithmetic, "
                                             "streaming stores"},
                      "achieved_GBs": round(gbs, 1)}), flush=True)


def serving(requests=48, rate=200.0, tick=500, graph_steps=100, alone=True,
            low_rate=2.0, low_requests=16, sat_glups=None):
    """Config 5, continuous-batching serving of coalescence solve requests:
    Poisson arrivals at `rate`/s, Lx in [128, 2048] (multiples of 64),
    Ly = Lx/2 or Lx, T in {10^3, 10^4, 5 10^4}, admission by HBM footprint.

    Reports solves/s and p50/p99 latency (arrival -> height on the host);
    the same requests solved one at a time with the fused single-film kernel
    (each solve timed on its own: state, initial height, steps, download);
    the speed-up over that and its mix-bound ceiling; latencies at an arrival
    rate below capacity; the op-by-op (reference step structure) estimate;
    and the rel-L2 of a 5 % sample against FP64 solves.

    The ceiling: every solve alone takes t_alone(r); a server cannot step the
    mix faster than the saturated rate, the largest of sat_glups (the fused
    kernel on a film that fills the device, final outputs only), the batched
    kernel on the 256 films of the batch benchmark (final outputs only) and
    the batched rate measured while serving, so no server beats
    sum(t_alone) / (sum(LUP) / saturated rate) x the one-at-a-time rate."""
    import numpy as np
    import swalbe as S
    from swalbe import serve
    reqs = serve.poisson_requests(requests, rate=rate, seed=5)
    free, total = S.device_memory()
    cap = int(min(free * 0.8, 200 << 30))
    be = serve.BatchBackend(cap, 4096, graph_steps=graph_steps)
    # warm up the kernels and the graph path on a tiny film
    w = serve.coalescence_request(-1, 128, 64, 2 * graph_steps + 3, np.random.default_rng(0), 1)
    serve.Server(be, tick=tick).serve([w], realtime=False)
    srv = serve.Server(be, tick=tick)
    done = srv.serve(reqs)
    st = serve.latency_stats(done)
    lups = sum(r.nodes * r.nsteps for r in reqs)
    finite = sum(bool(np.isfinite(r.result["height"]).all()) for r in done)
    out = {"workload": f"serve_poisson: {requests} requests, lambda={rate}/s, Lx in [128,2048] "
                       "step 64, Ly in {Lx/2, Lx}, T in {1e3,1e4,5e4}, two tangent caps, "
                       "seeded theta/gamma/radius (radius capped at gamma*H <= "
                       f"{serve.MAX_GAMMA_H}, the scheme's stability limit), FP32, tau=1",
           "finite_results": f"{finite}/{len(done)}",
           "solves_per_s": round(st["solves_per_s"], 4), "p50_s": round(st["p50_s"], 4),
           "p99_s": round(st["p99_s"], 4), "makespan_s": round(st["makespan_s"], 3),
           "batched_glups": round(st["glups"], 3), "lattice_updates": lups,
           "max_resident": srv.max_resident, "peak_arena_bytes": srv.peak_bytes,
           "arena_bytes": cap, "tick": tick, "graph_steps": graph_steps}
    if low_requests:
        # latency below capacity: arrivals slower than the server's solve rate
        low = serve.poisson_requests(low_requests, rate=low_rate, seed=6)
        sl = serve.latency_stats(serve.Server(be, tick=tick).serve(low))
        out["below_capacity"] = {"rate_per_s": low_rate, "requests": low_requests,
                                 "solves_per_s": round(sl["solves_per_s"], 4),
                                 "p50_s": round(sl["p50_s"], 4), "p99_s": round(sl["p99_s"], 4)}
    be.close()
    if alone:
        # the same requests one at a time, each with the fused single-film kernel
        t_alone = []
        for r in reqs:
            t = time.perf_counter()
            serve.solve_alone(r, "fused")
            t_alone.append(time.perf_counter() - t)
        alone_s = sum(t_alone)
        out["one_at_a_time_fused"] = {"solves_per_s": round(len(reqs) / alone_s, 4),
                                      "seconds": round(alone_s, 3),
                                      "glups": round(lups / alone_s / 1e9, 3)}
        out["speedup_vs_one_at_a_time_fused"] = round(st["solves_per_s"] * alone_s / len(reqs), 3)
        if sat_glups:
            sat_batch = batch_run(300, 20, graph_steps, final_outputs_only=True)["glups"]
            sat = max(sat_glups, sat_batch, st["glups"])
            ceil = alone_s / (lups / (sat * 1e9))
            out["speedup_ceiling_mix_bound"] = round(ceil, 3)
            out["ceiling_note"] = (f"sum of the solo times / the mix's lattice updates at "
                                   f"{sat:.1f} GLUPS, the largest of: the fused kernel on a "
                                   f"film that fills the device ({sat_glups:.1f}), the batched "
                                   f"kernel on the 256 batch-benchmark films ({sat_batch:.1f}), "
                                   f"both with final outputs only, and the batched rate while "
                                   f"serving ({st['glups']:.1f})")
        # plain GPU solves: the reference's op-by-op step (one kernel per op,
        # driven from the host), timed for 20 steps per distinct film shape and
        # scaled by each request's step count
        per_step = {}
        for r in reqs:
            if (r.Lx, r.Ly) not in per_step:
                per_step[(r.Lx, r.Ly)] = serve.perop_step_seconds(r)
        plain = sum(per_step[(r.Lx, r.Ly)] * r.nsteps for r in reqs)
        out["one_at_a_time_plain"] = {"solves_per_s": round(len(reqs) / plain, 4),
                                      "seconds_estimated": round(plain, 3),
                                      "method": "20 timed op-by-op steps per film shape x steps"}
        out["speedup_vs_one_at_a_time_plain"] = round(st["solves_per_s"] * plain / len(reqs), 3)
    pick = np.random.default_rng(17).choice(len(done), size=max(1, round(0.05 * len(done))),
                                            replace=False)
    sample = [done[i] for i in sorted(pick)]
    errs, dmass, dmax = [], [], []
    for r in sample:
        ref = serve.solve_alone(r, "fused", T=np.float64)["height"]
        h = r.result["height"]
        errs.append(float(np.linalg.norm(h - ref) / np.linalg.norm(ref)))
        dmass.append(float(abs(h.sum() - ref.sum()) / ref.sum()))
        dmax.append(float(abs(h.max() - ref.max()) / ref.max()))
    out["fp64_sample"] = {"requests": [f"{r.Lx}x{r.Ly} T={r.nsteps}" for r in sample],
                          "relL2_h": errs, "mass_rel_diff": dmass, "max_h_rel_diff": dmax,
                          "note": "served results are the bits of the same FP32 solve alone"}
    return out


def bench_serve(a):
    """Config 5 on its own line (serve_poisson): see serving()."""
    sv = serving(a.requests, a.rate, a.tick, a.graph_steps, alone=not a.no_alone)
    out = {"metric": "solves/s of continuous-batched FP32 coalescence solves (Poisson arrivals)",
           "value": sv["solves_per_s"], "unit": "solves/s", "n_gpus": 1,
           "steps": a.requests, "warmup": 1, "ms_per_step": None,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (seeded Poisson arrivals, seeded two-cap films)",
           "config": {"workload": sv.pop("workload"), "parallelism": "single GPU"}}
    out.update(sv)
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--workload", default="rupture4096_f32",
                    choices=sorted(WORKLOADS) + list(BATCH_WORKLOADS))
    ap.add_argument("--graph-steps", type=int, default=100,
                    help="batch workloads: steps per captured HIP graph (0 = plain launches)")
    ap.add_argument("--requests", type=int, default=48, help="serve_poisson: number of requests")
    ap.add_argument("--rate", type=float, default=200.0, help="serve_poisson: arrivals per second")
    ap.add_argument("--tick", type=int, default=500, help="serve_poisson: steps per scheduler tick")
    ap.add_argument("--no-alone", action="store_true",
                    help="serve_poisson: skip the one-at-a-time baseline")
    ap.add_argument("--films", type=int, default=2048, help="coalesce1d_scan: number of films")
    ap.add_argument("--threads", type=int, default=0,
                    help="1D resident kernel: threads per film (0 = by L, 256, 512, 1024)")
    ap.add_argument("--tile-y", type=int, default=0)
    ap.add_argument("--variant", type=int, default=0)
    ap.add_argument("--nontemporal", type=int, default=-1,
                    help="fused-step stores: 0 cached, 1 all streaming, 2 f and p streaming; "
                         "default the faster measured for the pass: FP32 films 1 (0 with final "
                         "outputs only) above 32 MB of inputs, else 0 (1); FP64 films and "
                         "batches 2")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-serving", action="store_true",
                    help="default workload: skip the config-5 serving measurement")
    ap.add_argument("--no-batch", action="store_true",
                    help="default workload: skip the config-4 batch measurement")
    ap.add_argument("--prewarm-s", type=float, default=1.0,
                    help="untimed seconds of steps before the warmup steps")
    ap.add_argument("--no-autotune", action="store_true",
                    help="coalesce256_batch: every film in 16-row tiles")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (independent "
              "films would shard by film; one large film by y-slabs with a halo exchange "
              "per step)", file=sys.stderr)
        sys.exit(2)

    if a.workload == "coalesce256_batch":
        return bench_batch(a)
    if a.workload == "serve_poisson":
        return bench_serve(a)
    if a.workload == "copy4096_f32":
        return bench_copy(a)
    if a.workload == "pattern4096_f32":
        return bench_pattern(a)
    if a.workload in ("coalesce1d_scan", "single1d_f64"):
        return bench_1d(a)
    if a.workload == "long1d_f64":
        return bench_long1d(a)

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

    def init_film():
        if a.workload in DROPLET:
            st.set_reference_height(0.1)
            st.init_height(S.film_init(S.INIT_CAP, precursor=0.1, href=0.1, clamp=True,
                                       cap1=(150.0, 1 / 9, Lx / 2 + 0.5, Ly / 2 + 0.5)))
        else:
            _lib.check(_lib.lib().swalbe_init_random(st.handle, 1, 1234, 1.0, 0.001), "init")
        S.equilibrium(st, sysc)

    init_film()
    prm = param.cparams()
    o = _lib.StepOpts()
    o.thermal, o.seed, o.tile_y, o.variant = int(thermal), 7, a.tile_y, a.variant
    # store modes (the same bits in each): by default the faster measured for
    # the film and the pass (swalbe.store_mode)
    if a.nontemporal >= 0:
        nt_full = nt_final = a.nontemporal
    else:
        nt_full = S.store_mode(Lx, Ly, dt == "f64", False)
        nt_final = S.store_mode(Lx, Ly, dt == "f64", True)
    o.nontemporal = nt_full
    ms = C.c_double()
    L = _lib.lib()
    # the live ceilings, then an untimed device prewarm: the part reaches its
    # working clocks before the warmup steps (the film is re-initialised, so
    # the timed steps start from the same state either way)
    ceil = ceilings(st, a.tile_y or 16) if (Lx, Ly, dt) == (4096, 4096, "f32") else None
    pre_steps, w0 = 0, time.perf_counter()
    while time.perf_counter() - w0 < a.prewarm_s:
        o.t0 = pre_steps
        _lib.check(L.swalbe_time_loop_timed(st.handle, C.byref(prm), 100, C.byref(o), C.byref(ms)))
        pre_steps += 100
    pre_s = time.perf_counter() - w0
    init_film()
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
    # the same steps with the distributions and pressure written on the last
    # step only (τ = 1: no step reads them; identical final state), reported
    # beside the headline figure, which writes every step's outputs
    o.final_outputs_only = 1
    o.nontemporal = nt_final
    o.t0 = a.warmup + a.steps
    ms2 = C.c_double()
    # untimed warmup of this pass: its kernel (its own store mode, or its own
    # kernel in FP32) may not have run yet, and its first launches are slower
    _lib.check(L.swalbe_time_loop_timed(st.handle, C.byref(prm), a.warmup, C.byref(o),
                                        C.byref(ms2)))
    o.t0 += a.warmup
    _lib.check(L.swalbe_time_loop_timed(st.handle, C.byref(prm), a.steps, C.byref(o), C.byref(ms2)))
    st.synchronize()
    o.final_outputs_only = 0
    o.nontemporal = nt_full
    fo_glups = N * a.steps / (ms2.value / 1e3) / 1e9
    h = st.get("height", np.float64)
    ok = bool(np.isfinite(h).all())
    # the same number of steps with the film handed over by the host and the
    # height handed back (PCIe both ways, on the wall clock); never `value`
    st.synchronize()
    w0 = time.perf_counter()
    st.height = h
    o.t0 = 2 * (a.warmup + a.steps)
    _lib.check(L.swalbe_time_loop(st.handle, C.byref(prm), a.steps, C.byref(o)))
    h_back = st.get("height", np.float64)
    pcie_s = time.perf_counter() - w0
    ok = ok and bool(np.isfinite(h_back).all())
    st.close()
    if ceil:
        ceil["frac_of_copy"] = round(achieved / ceil["copy_GBs"], 4)
        ceil["frac_of_pattern"] = round(glups / ceil["pattern_GLUPS"], 4)
    line = {
        "metric": f"GLUPS (lattice updates/s) of the fused {'FP32' if dt == 'f32' else 'FP64'} "
                  f"D2Q9 thin-film step at {Lx}x{Ly}" if (Lx, Ly, dt) != (4096, 4096, "f32")
                  else "GLUPS (lattice updates/s) of the fused FP32 D2Q9 thin-film step at 4096^2",
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
        "data": ("synthetic (spherical cap on a 0.1 precursor)" if a.workload in DROPLET
                 else "synthetic (seeded Philox film h0 = 1 + 0.001 N(0,1))"),
        "config": {"workload": f"{a.workload}: D2Q9 thin film "
                               + ("sessile drop (cap radius 150, precursor 0.1), "
                                  if a.workload in DROPLET else "")
                               + ("rupture with thermal noise, kbt=1e-6, " if thermal else "without noise, ")
                               + f"{Lx}x{Ly} periodic, theta=1/9, n=9, m=3, gamma=0.01, tau=1",
                   "Lx": Lx, "Ly": Ly, "parallelism": "single GPU",
                   "tile_y": a.tile_y or (16 if dt == "f32" or Lx * Ly * 24 > (32 << 20) else 8),
                   "variant": a.variant, "nontemporal": nt_full,
                   "plane_pad": int(os.environ.get("SWALBE_PLANE_PAD", "0")), "finite": ok},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "bytes_per_lup": bpl},
        "prewarm": {"seconds": round(pre_s, 3), "steps": pre_steps,
                    "note": "untimed steps before the warmup (the film is re-initialised "
                            "after them); the ceilings run before them too"},
        "final_outputs_only": {"GLUPS": round(fo_glups, 3),
                               "bytes_per_lup": 7 * esz, "nontemporal": nt_final,
                               "note": "f and p written on the last step only (opts)"},
        "ceilings": ceil,
        "host_handover": {"GLUPS": round(N * a.steps / pcie_s / 1e9, 3),
                          "seconds": round(pcie_s, 4),
                          "note": "height uploaded, the same steps, height downloaded (PCIe "
                                  "both ways, wall clock)"},
    }
    if not a.no_batch and (Lx, Ly, dt) == (4096, 4096, "f32"):
        # config 4: 256 films in one launch per step, every step's outputs
        # written (its own line: --workload coalesce256_batch)
        r = batch_run(300, 20)
        line["batch"] = {"workload": BATCH_DESC, "GLUPS": round(r["glups"], 3),
                         "ms_per_step": round(r["ms"] / 300, 5), "steps": 300, "warmup": 20,
                         "films": r["films"], "nodes": r["nodes"], "graph_steps": 100,
                         "finite": r["finite"], "films_rows16": r["rows"].count(16),
                         "admission_s": round(r["admit_s"], 3)}
    if not a.no_serving and (Lx, Ly, dt) == (4096, 4096, "f32"):
        # config 5: the second half of the headline metric (batched solves/s
        # and p50/p99 latency)
        line["serving"] = serving(a.requests, a.rate, a.tick, a.graph_steps,
                                  alone=not a.no_alone, sat_glups=fo_glups)
    if not a.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline()
        line["cpu_baseline_fused"] = cpu_baseline_fused()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
