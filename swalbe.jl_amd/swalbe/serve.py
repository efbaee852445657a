"""Continuous-batching solve server for independent thin-film runs.

Requests are parameter-sweep solves (a film size, parameters, an initial
height spec and a step budget).  They arrive over time; the server admits
them in arrival order while the batch arena has room for their footprint
and a film slot is free, advances every resident film together (one launch
per time step, the batched FP32 kernel), and hands back each film's height
as soon as its budget is spent, freeing its memory for the next arrivals.
Work of films admitted at different times overlaps on the device.

The reference runs such sweeps as a serial host loop over single-film runs
(scripts/Rivulet_stability.jl:141-190); `solve_alone` is that one-at-a-time
path on the GPU, used as the baseline and for the alone-vs-in-mix check.

The scheduler talks to the device through a small backend interface
(`BatchBackend` below), so its logic is testable without a GPU.
"""
from __future__ import annotations

import math
import time
from collections import deque
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional, Sequence

import numpy as np

from ._lib import ENOMEM, FilmInit, SwalbeError

# single-film kernel variant and tile height whose arithmetic the batched
# kernel shares (LDS tiles of 64 x 16, collide at the destination)
BATCH_VARIANT, BATCH_TILE_Y = 2, 16

MASK64 = (1 << 64) - 1


def splitmix64(x: int) -> int:
    """SplitMix64 finaliser (Steele, Lea, Flood 2014): a per-domain seed
    from (batch seed, index)."""
    x = (x + 0x9E3779B97F4A7C15) & MASK64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & MASK64
    return x ^ (x >> 31)


def domain_seed(batch_seed: int, k: int) -> int:
    return splitmix64((int(batch_seed) * 0x100000001B3 + int(k)) & MASK64)


@dataclass
class Request:
    rid: int
    Lx: int
    Ly: int
    param: object          # Taumucs
    init: FilmInit
    nsteps: int
    arrival: float = 0.0   # seconds after the start of serving
    fields: Sequence[str] = ("height",)
    # filled in by the server
    slot: int = -1
    t_admit: float = math.nan
    t_done: float = math.nan
    result: Dict[str, np.ndarray] = field(default_factory=dict)

    @property
    def nodes(self) -> int:
        return self.Lx * self.Ly

    @property
    def latency(self) -> float:
        return self.t_done - self.arrival


def two_caps_film(Lx: int, Ly: int, R: float, theta: float, seed: int, precursor: float = 0.1,
                  href: Optional[float] = None) -> FilmInit:
    """Two tangent spherical caps of sphere radius R and contact angle πθ on
    the x mid-line: base radius b = R sin(πθ), centres 2b apart (the 2D
    analogue of two_droplets, src/initialvalues.jl:277-323).  Heights below
    the precursor are raised to it: the reference's rule (only negative
    heights) leaves a neck of h ~ 1e-3 between tangent caps, whose disjoining
    pressure drives velocities of O(100) and blows the film up in ~10 steps."""
    from .batch import INIT_TWO_CAPS, film_init
    b = R * math.sin(math.pi * theta)
    cx, cy = Lx / 2 + 0.5, Ly / 2 + 0.5
    return film_init(INIT_TWO_CAPS, seed=seed, precursor=precursor,
                     href=precursor if href is None else href,
                     cap1=(R, theta, cx - b, cy), cap2=(R, theta, cx + b, cy), clamp=True)


# The reference's scheme (τ = 1, this film-pressure stencil) stays finite
# for caps with γ·H up to about 0.43 and blows up from about 0.48, where H =
# R (1 - cos πθ) is the cap height (checked with the FP64 oracle on two
# tangent caps: γ = 0.02: H = 21.4 finite, 24.1 not; γ = 0.01: H = 40.2
# finite, 53.6 not; γ = 0.005: H = 80.4 finite).  Serving requests on films
# up to 2048 nodes wide draw caps beyond that, which return NaN in FP64 as in
# FP32; their sphere radius is capped at γ·H = MAX_GAMMA_H.
MAX_GAMMA_H = 0.4


def coalescence_request(rid: int, Lx: int, Ly: int, nsteps: int, rng: np.random.Generator,
                        seed: int, arrival: float = 0.0,
                        max_gamma_h: Optional[float] = None) -> Request:
    """One droplet-coalescence solve: R ~ U(0.15, 0.3) Ly, θ/π ~ U(1/18, 1/6),
    γ ~ U(0.005, 0.02); n = 9, m = 3, hmin = 0.1, hcrit = 0.05, δ = 1, τ = 1.
    max_gamma_h caps R so that γ R (1 - cos πθ) stays below it (the draws
    are the same either way, so capped and uncapped sequences pair up)."""
    from .core import Taumucs
    R = rng.uniform(0.15, 0.3) * Ly
    theta = rng.uniform(1 / 18, 1 / 6)
    gamma = rng.uniform(0.005, 0.02)
    if max_gamma_h is not None:
        R = min(R, max_gamma_h / (gamma * (1.0 - math.cos(math.pi * theta))))
    param = Taumucs(Tmax=nsteps, tdump=10 ** 9, γ=gamma, θ=theta, n=9, m=3, hmin=0.1,
                    hcrit=0.05, δ=1.0)
    return Request(rid, Lx, Ly, param, two_caps_film(Lx, Ly, R, theta, seed), nsteps, arrival)


def coalescence_sweep(n: int = 256, seed: int = 1, nsteps: int = 10_000) -> List[Request]:
    """n independent coalescence domains, Lx uniform in {128, 192, ..., 1024},
    Ly = Lx / 2, per-domain seed from (seed, k)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        Lx = int(rng.choice(np.arange(128, 1025, 64)))
        out.append(coalescence_request(k, Lx, Lx // 2, nsteps, rng, domain_seed(seed, k)))
    return out


def poisson_requests(n: int, rate: float = 200.0, seed: int = 5,
                     lx_range=(128, 2048), steps=(1_000, 10_000, 50_000),
                     max_gamma_h: Optional[float] = MAX_GAMMA_H) -> List[Request]:
    """n solve requests with Poisson arrivals at `rate` per second: Lx a
    multiple of 64 in lx_range, Ly = Lx/2 or Lx, T drawn from `steps`,
    seeded θ/γ/radius, the radius capped at the scheme's stability limit
    (MAX_GAMMA_H)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for k in range(n):
        t += rng.exponential(1.0 / rate)
        Lx = int(rng.choice(np.arange(lx_range[0], lx_range[1] + 1, 64)))
        Ly = Lx // 2 if rng.random() < 0.5 else Lx
        T = int(rng.choice(steps))
        out.append(coalescence_request(k, Lx, Ly, T, rng, domain_seed(seed, k), arrival=t,
                                       max_gamma_h=max_gamma_h))
    return out


class BatchBackend:
    """The device side of the server: a swalbe Batch."""

    def __init__(self, capacity_bytes: int, max_films: int = 4096, graph_steps: int = 0,
                 final_outputs_only: bool = True):
        from .batch import Batch, film_bytes
        # requests return heights: a film's distributions and pressure are
        # written on its last step only
        self.batch = Batch(capacity_bytes, max_films, graph_steps, final_outputs_only)
        self.capacity = int(capacity_bytes)
        self.max_films = int(max_films)
        self._film_bytes = film_bytes

    def footprint(self, r: Request) -> int:
        return self._film_bytes(r.Lx, r.Ly, r.param.τ)

    def bytes_free(self) -> int:
        return self.batch.stats()["bytes_free"]

    def add(self, r: Request) -> int:
        return self.batch.add(r.param, r.Lx, r.Ly, r.init, r.nsteps)

    def step(self, n: int) -> None:
        self.batch.step(n)

    def wait(self, steps_behind: int) -> None:
        self.batch.wait(steps_behind)

    def get(self, slot: int, name: str) -> np.ndarray:
        return self.batch.get(slot, name)

    def remove(self, slot: int) -> None:
        self.batch.remove(slot)

    def close(self) -> None:
        self.batch.close()


class Server:
    """Continuous batching over a backend.

    Each tick: admit arrived requests in arrival order while their footprint
    fits and a slot is free (strict FIFO, so a large request is never
    starved), enqueue `k` steps for all resident films, where k is the tick
    length cut to the smallest remaining budget, then hand back the films
    that finished in an earlier tick.  The host runs at most one tick ahead
    of the device, so new arrivals wait at most about two ticks."""

    def __init__(self, backend, tick: int = 500, clock: Callable[[], float] = time.perf_counter,
                 sleep: Callable[[float], None] = time.sleep):
        self.be = backend
        self.tick = int(tick)
        self.clock = clock
        self.sleep = sleep
        self.queue: deque = deque()
        self.resident: Dict[int, Request] = {}   # slot -> request
        self.remaining: Dict[int, int] = {}      # slot -> steps not yet enqueued
        self.done: List[Request] = []
        self.steps_enqueued = 0
        self.finished_at: Dict[int, int] = {}    # slot -> steps_enqueued when its budget ran out
        self.peak_bytes = 0
        self.bytes_used = 0
        self.max_resident = 0
        self._t0 = 0.0

    def now(self) -> float:
        return self.clock() - self._t0

    def admit(self) -> int:
        n = 0
        while self.queue:
            r = self.queue[0]
            need = self.be.footprint(r)
            if len(self.resident) >= self.be.max_films or need > self.be.bytes_free():
                break
            try:
                slot = self.be.add(r)
            except SwalbeError as e:
                if e.status != ENOMEM:
                    raise
                break  # free bytes are fragmented: wait for a retirement
            self.queue.popleft()
            r.slot = slot
            r.t_admit = self.now()
            self.resident[r.slot] = r
            self.remaining[r.slot] = r.nsteps
            self.bytes_used += need
            self.peak_bytes = max(self.peak_bytes, self.bytes_used)
            if r.nsteps == 0:
                self.finished_at[r.slot] = self.steps_enqueued
            n += 1
        self.max_resident = max(self.max_resident, len(self.resident))
        return n

    def _retire(self, upto: int) -> int:
        """Hand back films whose last step was enqueued at or before `upto`."""
        n = 0
        for slot in [s for s, at in self.finished_at.items() if at <= upto]:
            r = self.resident.pop(slot)
            for name in r.fields:
                r.result[name] = self.be.get(slot, name)
            self.be.remove(slot)
            self.bytes_used -= self.be.footprint(r)
            r.t_done = self.now()
            del self.finished_at[slot]
            del self.remaining[slot]
            self.done.append(r)
            n += 1
        return n

    def step_once(self) -> int:
        """One tick; returns the number of steps enqueued."""
        live = [v for v in self.remaining.values() if v > 0]
        if not live:
            return 0
        k = min(self.tick, min(live))
        self.be.step(k)
        prev = self.steps_enqueued
        self.steps_enqueued += k
        for slot in list(self.remaining):
            if self.remaining[slot] > 0:
                self.remaining[slot] -= k
                if self.remaining[slot] <= 0:
                    self.remaining[slot] = 0
                    self.finished_at[slot] = self.steps_enqueued
        # the previous tick is complete once at most this one is pending
        self.be.wait(k)
        self._retire(prev)
        return k

    def serve(self, requests: Sequence[Request], realtime: bool = True) -> List[Request]:
        """Serve `requests` (arrival times in seconds from now) to completion.
        realtime=False admits every request at once (offline batch)."""
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        self._t0 = self.clock()
        if not realtime:
            for r in pending:
                r.arrival = 0.0
        while pending or self.queue or self.resident:
            now = self.now()
            while pending and pending[0].arrival <= now:
                self.queue.append(pending.popleft())
            self.admit()
            if self.step_once() == 0:
                if self.finished_at:
                    self.be.wait(0)
                    self._retire(self.steps_enqueued)
                    continue
                if self.queue and not self.resident:
                    raise RuntimeError(f"request {self.queue[0].rid} ({self.queue[0].Lx}x"
                                       f"{self.queue[0].Ly}) does not fit the empty batch")
                if pending:
                    dt = pending[0].arrival - self.now()
                    if dt > 0:
                        self.sleep(min(dt, 0.01))
        return self.done


class NativeServer:
    """The same scheduler in the library (swalbe_serve_*): for callers that
    drive serving through the C ABI alone."""

    def __init__(self, capacity_bytes: int, max_films: int = 4096, tick: int = 500,
                 graph_steps: int = 0):
        import ctypes as C

        from ._lib import check, lib
        self._C, self._check, self._lib = C, check, lib
        h = C.c_void_p()
        check(lib().swalbe_serve_create(int(capacity_bytes), int(max_films), int(tick),
                                        int(graph_steps), C.byref(h)), "swalbe_serve_create")
        self._h = h
        self.requests: Dict[int, Request] = {}

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib().swalbe_serve_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, r: Request) -> int:
        C = self._C
        rid = C.c_int64()
        prm = r.param.cparams()
        self._check(self._lib().swalbe_serve_submit(self._h, C.byref(prm), r.Lx, r.Ly,
                                                    C.byref(r.init), int(r.nsteps), C.byref(rid)),
                    "swalbe_serve_submit")
        self.requests[rid.value] = r
        return rid.value

    def poll(self) -> List[Request]:
        """One tick; returns the requests finished since the last poll."""
        C = self._C
        nd, nr, nq = C.c_int32(), C.c_int32(), C.c_int32()
        self._check(self._lib().swalbe_serve_poll(self._h, C.byref(nd), C.byref(nr), C.byref(nq)),
                    "swalbe_serve_poll")
        self.resident, self.queued = nr.value, nq.value
        out = []
        for _ in range(nd.value):
            rid, Lx, Ly = C.c_int64(), C.c_int32(), C.c_int32()
            self._check(self._lib().swalbe_serve_peek(self._h, C.byref(rid), C.byref(Lx),
                                                      C.byref(Ly)), "swalbe_serve_peek")
            h = np.empty((Lx.value, Ly.value), order="F")
            self._check(self._lib().swalbe_serve_take(
                self._h, h.ctypes.data_as(C.POINTER(C.c_double)), h.nbytes), "swalbe_serve_take")
            r = self.requests.pop(rid.value)
            r.result["height"] = h
            out.append(r)
        return out

    def serve(self, requests: Sequence[Request], clock=time.perf_counter) -> List[Request]:
        """Submit each request at its arrival time and poll until all are done."""
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        t0 = clock()
        done: List[Request] = []
        self.max_resident = 0
        while pending or self.requests:
            now = clock() - t0
            while pending and pending[0].arrival <= now:
                r = pending.popleft()
                self.submit(r)
            if not self.requests:
                time.sleep(min(max(pending[0].arrival - (clock() - t0), 0.0), 0.01))
                continue
            for r in self.poll():
                r.t_done = clock() - t0
                done.append(r)
            self.max_resident = max(self.max_resident, self.resident)
        return done


def latency_stats(reqs: Sequence[Request]) -> dict:
    lat = np.array([r.latency for r in reqs])
    first = min(r.arrival for r in reqs)
    last = max(r.t_done for r in reqs)
    lups = float(sum(r.nodes * r.nsteps for r in reqs))
    span = last - first
    return {"requests": len(reqs), "makespan_s": span, "solves_per_s": len(reqs) / span,
            "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "glups": lups / span / 1e9}


def perop_step_seconds(r: Request, steps: int = 20, warm: int = 3) -> float:
    """Seconds per step of the op-by-op path on r's film (state set up
    beforehand, device synchronised around the timed steps)."""
    from .core import (BGKandStream, Sys, SysConst, equilibrium, filmpressure, force_sum,
                       h_grad_p, moments, slippage)
    sysc = SysConst(r.Lx, r.Ly, r.param)
    st = Sys(sysc, "GPU", T=np.float32)
    st.set_reference_height(r.init.href, r.param.g)
    st.init_height(r.init)
    equilibrium(st, sysc)

    def one():
        filmpressure(st, sysc)
        h_grad_p(st)
        slippage(st, sysc)
        force_sum(st, False)
        equilibrium(st, sysc)
        BGKandStream(st, sysc)
        moments(st)
    for _ in range(warm):
        one()
    st.synchronize()
    t = time.perf_counter()
    for _ in range(steps):
        one()
    st.synchronize()
    dt = (time.perf_counter() - t) / steps
    st.close()
    return dt


def solve_alone(r: Request, mode: str = "fused", T=np.float32,
                final_outputs_only: bool = True) -> Dict[str, np.ndarray]:
    """The request run by itself on the GPU: a single-film state, the same
    initial height, then `mode` = "fused" (one fused kernel per step) or
    "perop" (the reference's op-by-op step: filmpressure!, h∇p!, slippage!,
    force sum, equilibrium!, BGKandStream!, moments!).  FP32 fused runs give
    the same bits as the film in a batch."""
    import ctypes as C

    from . import _lib
    from .core import (BGKandStream, Sys, SysConst, equilibrium, filmpressure, force_sum,
                       h_grad_p, moments, slippage, store_mode)
    sysc = SysConst(r.Lx, r.Ly, r.param)
    st = Sys(sysc, "GPU", T=T, kind="thermal" if r.init.thermal else "simple")
    st.set_reference_height(r.init.href, r.param.g)
    st.init_height(r.init)
    equilibrium(st, sysc)
    if r.param.τ != 1.0:
        st.copy("ftemp", "feq")
    if mode == "fused":
        o = _lib.StepOpts()
        # the kernel variant batches use, so that the bits agree
        o.thermal, o.seed = int(r.init.thermal), int(r.init.seed)
        o.variant, o.tile_y = BATCH_VARIANT, BATCH_TILE_Y
        o.final_outputs_only = int(bool(final_outputs_only))
        # the store mode measured fastest for the film alone (the same bits in every mode)
        o.nontemporal = store_mode(r.Lx, r.Ly, T is np.float64, bool(final_outputs_only))
        o.body = int(r.init.body)
        o.body_ax, o.body_ay, o.body_fac = r.init.body_ax, r.init.body_ay, r.init.body_fac
        prm = r.param.cparams()
        _lib.check(_lib.lib().swalbe_time_loop(st.handle, C.byref(prm), int(r.nsteps),
                                               C.byref(o)), "time_loop")
    elif mode == "perop":
        if r.init.thermal or r.init.body:
            raise ValueError("perop baseline: plain films only")
        for _ in range(r.nsteps):
            filmpressure(st, sysc)
            h_grad_p(st)
            slippage(st, sysc)
            force_sum(st, False)
            equilibrium(st, sysc)
            BGKandStream(st, sysc)
            moments(st)
    else:
        raise ValueError(mode)
    out = {name: st.get(name, np.float64) for name in r.fields}
    st.close()
    return out


__all__ = ["Request", "Server", "NativeServer", "BatchBackend", "coalescence_sweep", "poisson_requests",
           "perop_step_seconds",
           "coalescence_request", "two_caps_film", "solve_alone", "latency_stats", "domain_seed",
           "splitmix64"]
