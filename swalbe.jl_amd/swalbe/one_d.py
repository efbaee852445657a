"""Swalbe.jl's one-dimensional (D1Q3) API over libswalbe_hip.

SysConst_1D / SysConstWithBound_1D, the 1D state kinds of Sys(sysc; T, kind)
(src/initialize.jl:84-120, 276-341, 587-616), the ops, the time loops of
src/simulate.jl:98-205, the 1D run_* drivers (:247-560), the 1D initial
conditions (src/initialvalues.jl:226-323) and the obstacle masks
(src/obstacle.jl:6-53).  The generic names of the package (Sys,
equilibrium, BGKandStream, moments, filmpressure, h_grad_p, slippage,
thermal, inclination, time_loop, run_*, singledroplet) dispatch here when
given a 1D system or state, as Julia's methods do.

Arrays are numpy vectors of length L; distributions are (L, 3) arrays in
Fortran order (population k = column k, velocities 0, +1, -1).
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Callable, Optional, Sequence

import numpy as np

from . import _lib
from ._lib import FIELDS_1D, check, lib
from .fields import LiveFields

DIST_1D = ("fout", "ftemp", "feq")


# ---------------------------------------------------------------- systems


class SysConst_1D:
    """Lattice length plus Taumucs (src/initialize.jl:96-100)."""

    def __init__(self, L: int = 256, param=None):
        from .core import Taumucs
        self.L = int(L)
        self.param = param if param is not None else Taumucs()


class SysConstWithBound_1D(SysConst_1D):
    """SysConstWithBound_1D (src/initialize.jl:112-120): an obstacle vector
    `obs` and the masks obslist! derives from it (interior, border =
    [obsright, obsleft])."""

    def __init__(self, L: int = 256, param=None, obs=None):
        super().__init__(L, param)
        self.obs = np.zeros(self.L) if obs is None else np.asarray(obs, dtype=float).copy()
        self.interior = np.zeros(self.L)
        self.border = [np.zeros(self.L), np.zeros(self.L)]


def obslist1D(obs, T=np.float64, verbose: bool = True):
    """obslist1D(obs) (src/obstacle.jl:6-35): (interior, [obsright, obsleft])
    with periodic neighbours."""
    if verbose:
        print("WARNING: Always make your obstacles four nodes thick or the algorithm will crash.")
    obs = np.asarray(obs)
    right, left = np.roll(obs, -1), np.roll(obs, 1)  # obs[i+1], obs[i-1]
    interior = ((obs == 1) & (right == 1) & (left == 1)).astype(T)
    obsleft = ((obs == 1) & (left == 1)).astype(T)
    obsright = ((obs == 1) & (right == 1)).astype(T)
    return interior, [obsright, obsleft]


def obslist(sys: SysConstWithBound_1D, T=np.float64, verbose: bool = False) -> None:
    """obslist!(sys) (src/obstacle.jl:41-53): refresh interior and border from obs."""
    interior, border = obslist1D(sys.obs, T=T, verbose=verbose)
    sys.interior[...] = interior
    sys.border[0][...] = border[0]
    sys.border[1][...] = border[1]


# ---------------------------------------------------------------- state


class HipState1D(LiveFields):
    """A 1D film resident on the GPU: the arrays of State_1D plus kbt
    (State_thermal_1D), γ and ∇γ (State_gamma_1D), θ and the wall masks.
    `state.basestate` is the state itself, so reference code written as
    state.basestate.height reads the same.  `state.height` is a live view of
    the device array (reads see the device, in-place writes reach it before
    the next op); assigning an array uploads it, assigning a scalar fills it."""

    _names = dict(FIELDS_1D, **{"h∇p": FIELDS_1D["hgradp"], "γ": FIELDS_1D["gamma"],
                                "∇γ": FIELDS_1D["gradgamma"], "θ": FIELDS_1D["theta"]})
    _kinds = ("simple", "thermal", "gamma", "gamma_bound")
    _is_1d = True

    def __init__(self, L: int, T=np.float64, kind: str = "simple"):
        T = np.dtype(T).type
        if T not in (np.float64, np.float32):
            raise TypeError("T must be Float64 or Float32")
        if kind not in self._kinds:
            raise ValueError(f"kind must be one of {self._kinds}")
        self._init_fields()
        object.__setattr__(self, "L", int(L))
        object.__setattr__(self, "T", T)
        object.__setattr__(self, "kind", kind)
        h = C.c_void_p()
        check(lib().swalbe1d_create(int(L), _lib.F64 if T is np.float64 else _lib.F32,
                                    _lib.THERMAL if kind == "thermal" else _lib.SIMPLE,
                                    C.byref(h)), "swalbe1d_create")
        object.__setattr__(self, "_h", h)
        object.__setattr__(self, "t", 0)
        if kind in ("gamma", "gamma_bound"):
            self.set("gamma", 0.0)
            self.set("gradgamma", 0.0)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().swalbe1d_destroy(self._h)
            object.__setattr__(self, "_h", None)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def basestate(self):
        return self

    @property
    def dtype(self) -> int:
        return _lib.F64 if self.T is np.float64 else _lib.F32

    def _field_shape(self, name):
        return (self.L, 3) if self._names[name] <= FIELDS_1D["feq"] else (self.L,)

    def _download(self, name: str, dtype=None) -> np.ndarray:
        out = np.empty(self._field_shape(name), dtype=np.dtype(dtype or self.T), order="F")
        check(lib().swalbe1d_download(self._h, self._names[name], out.ctypes.data, out.nbytes),
              f"download {name}")
        return out

    def _upload(self, name: str, value) -> None:
        fid = self._names[name]
        if np.ndim(value) == 0:
            check(lib().swalbe1d_fill(self._h, fid, float(value)), f"fill {name}")
            return
        arr = np.asarray(value)
        shape = self._field_shape(name)
        if arr.shape != shape:
            arr = np.broadcast_to(arr, shape)
        arr = np.asfortranarray(arr, dtype=np.float32 if arr.dtype == np.float32 else np.float64)
        check(lib().swalbe1d_upload(self._h, fid, arr.ctypes.data, arr.nbytes), f"upload {name}")

    def synchronize(self):
        check(lib().swalbe1d_synchronize(self.handle), "synchronize")

    def reduce(self, name: str = "height", thresh: float = 0.0, interior: bool = False):
        out = (C.c_double * 4)()
        check(lib().swalbe1d_reduce(self.handle, self._names[name], float(thresh), int(interior),
                                    out), "reduce")
        return {"sum": out[0], "max": out[1], "min": out[2], "count": int(out[3])}


def Sys(sysc: SysConst_1D, T=np.float64, kind: str = "simple") -> HipState1D:
    """Sys(sysc::Consts_1D; T, kind) (src/initialize.jl:587-616): zeros and
    height = 1 on the GPU.  A walled system's masks travel with it."""
    st = HipState1D(sysc.L, T=T, kind=kind)
    if isinstance(sysc, SysConstWithBound_1D):
        _upload_walls(st, sysc)
    return st


def _upload_walls(state: HipState1D, sys: SysConstWithBound_1D) -> None:
    """Uploads the wall masks when they differ from the ones the state holds."""
    key = b"".join(np.ascontiguousarray(a, dtype=np.float64).tobytes()
                   for a in (sys.border[0], sys.border[1], sys.interior))
    if getattr(state, "_walls", None) == key:
        return
    state.set("border1", sys.border[0])
    state.set("border2", sys.border[1])
    state.set("interior", sys.interior)
    object.__setattr__(state, "_walls", key)


# ---------------------------------------------------------------- ops


def equilibrium(state: HipState1D, sys, g=None) -> None:
    """equilibrium!(state, sys; g) (src/equilibrium.jl:169-191)."""
    check(lib().swalbe1d_equilibrium(state.handle, float(sys.param.g if g is None else g)),
          "equilibrium")


def BGKandStream(state: HipState1D, sys, τ=None) -> None:
    """BGKandStream!(state, sys) (src/collide.jl:179-212); with a
    SysConstWithBound_1D the bounce-back method (:214-249)."""
    bound = isinstance(sys, SysConstWithBound_1D)
    if bound:
        _upload_walls(state, sys)
    check(lib().swalbe1d_bgk_and_stream(state.handle, float(sys.param.τ if τ is None else τ),
                                        int(bound)), "BGKandStream")


def moments(state: HipState1D) -> None:
    """moments!(state) (src/moments.jl:54-76)."""
    check(lib().swalbe1d_moments(state.handle), "moments")


def filmpressure(state: HipState1D, sys, θ=None, γ=None, n=None, m=None, hmin=None, hcrit=None,
                 raw: bool = False) -> None:
    """filmpressure!(state, sys; θ, n, m, hmin, hcrit, γ) (src/pressure.jl:231-315);
    θ and γ may be vectors (uploaded as the θ / γ fields)."""
    P = sys.param
    th = P.θ if θ is None else θ
    gm = P.γ if γ is None else γ
    use_t = use_g = 0
    if np.ndim(th) > 0:
        state.set("theta", th)
        use_t, th = 1, 0.0
    if np.ndim(gm) > 0:
        state.set("gamma", gm)
        use_g, gm = 1, 0.0
    check(lib().swalbe1d_filmpressure(state.handle, float(gm), float(th),
                                      int(P.n if n is None else n), int(P.m if m is None else m),
                                      float(P.hmin if hmin is None else hmin),
                                      float(P.hcrit if hcrit is None else hcrit), use_t, use_g,
                                      int(raw)), "filmpressure")


def grad_gamma(state: HipState1D, sys=None) -> None:
    """∇γ!(state) (src/forcing.jl:423-432) or ∇γ!(state, sys) (:434-447)."""
    check(lib().swalbe1d_grad_gamma(state.handle, 0 if sys is None else 1,
                                    1.0 if sys is None else float(sys.param.δ)), "grad_gamma")


def h_grad_p(state: HipState1D) -> None:
    """h∇p!(state) (src/forcing.jl:189-198, 223-232)."""
    check(lib().swalbe1d_hgradp(state.handle), "h_grad_p")


def slippage(state: HipState1D, sys) -> None:
    """slippage!(state, sys) (src/forcing.jl:68-82)."""
    check(lib().swalbe1d_slippage(state.handle, sys.param.δ, sys.param.μ), "slippage")


def thermal(state: HipState1D, sys, seed: int = 0, step: Optional[int] = None) -> None:
    """thermal!(state, sys) (src/forcing.jl:322-335); normals keyed (seed, step, x)."""
    st = state.t if step is None else step
    check(lib().swalbe1d_thermal(state.handle, sys.param.kbt, sys.param.μ, sys.param.δ, int(seed),
                                 int(st)), "thermal")


def _incl_fac(t, tstart, tsmooth):
    return 0.5 + 0.5 * math.tanh((t - tstart) / tsmooth)


def inclination(α, state: HipState1D, t: int = 1000, tstart: int = 0, tsmooth: int = 1) -> None:
    """inclination!(α::Float64, state; t, tstart, tsmooth) (src/forcing.jl:379-389)."""
    check(lib().swalbe1d_inclination(state.handle, float(α), _incl_fac(t, tstart, tsmooth)),
          "inclination")


def force_sum(state: HipState1D, thermal: bool = False, grad_gamma: int = 0) -> None:
    """state.F .= -state.h∇p .- state.slip [.- kbt] [∓ ∇γ] (src/simulate.jl:111, 548)."""
    check(lib().swalbe1d_force_sum(state.handle, int(bool(thermal)), int(grad_gamma)), "force_sum")


def _ptr_f64(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def wetted(area_size, state, hthresh: float = 0.055):
    """wetted!(area_size, state) (src/measures.jl:13-17)."""
    area_size.append(int(state.reduce("height", hthresh)["count"]))


# ---------------------------------------------------------------- time loop


def _options(sys, state, thermal, seed, materialize, variant, grad_gamma_sign):
    o = _lib.StepOpts1D()
    o.thermal, o.seed, o.t0 = int(thermal), int(seed), int(state.t)
    o.materialize, o.variant, o.grad_gamma = int(materialize), int(variant), int(grad_gamma_sign)
    if isinstance(sys, SysConstWithBound_1D):
        _upload_walls(state, sys)
        o.bound = 1
    return o


def time_loop(sys, state: HipState1D, *args, verbose: bool = False, thermal: bool = False,
              seed: int = 0, materialize: bool = True, variant: int = 0, γ=None,
              grad_gamma_sign: int = 0, mass: Optional[list] = None,
              snapshots: Optional[np.ndarray] = None, dumping: int = 0, threads: int = 0):
    """time_loop(sys::SysConst_1D, state[, θ | Δh | f, measure]) and the walled
    loop (src/simulate.jl:98-205), one fused GPU loop.  `threads` (0, 256,
    512, 1024): threads of the resident kernel's film, or nodes per workgroup
    of the one-launch-per-step kernel; 0 picks automatically.

    Known callbacks run fused: wetted (count per step) and inclination (α,
    body force).  Another callable f(measure, state) runs the per-op kernels
    so that f sees the state between the force sum and equilibrium!.
    γ (scalar or vector) is filmpressure!'s γ (run_gamma); grad_gamma_sign
    -1 subtracts the ∇γ field from F, +1 adds it.  `snapshots` (shape
    (Tmax ÷ dumping, L)) receives snapshot!(snapshots, h, t; dumping)."""
    P = sys.param
    prm = P.cparams()
    nsteps = P.Tmax
    o = _options(sys, state, thermal, seed, materialize, variant, grad_gamma_sign)
    o.threads = int(threads)
    if γ is not None:
        if np.ndim(γ) > 0:
            state.set("gamma", γ)
            o.use_gamma_field = 1
        else:
            prm.gamma = float(γ)
    dh = wet = None
    f = measure = None
    if len(args) == 1:
        a = args[0]
        if isinstance(a, list):
            dh = np.zeros(nsteps)
            o.dh_series = _ptr_f64(dh)
        elif np.ndim(a) > 0:
            state.set("theta", a)
            o.use_theta_field = 1
        else:
            prm.theta = float(a)
    elif len(args) == 2:
        f, measure = args
        if f is wetted or getattr(f, "__name__", "") == "wetted":
            wet = np.zeros(nsteps, dtype=np.int64)
            o.wet_series = wet.ctypes.data_as(C.POINTER(C.c_int64))
            o.wet_thresh = 0.055
        elif f is inclination or getattr(f, "__name__", "") == "inclination":
            o.body, o.alpha, o.body_fac = 1, float(measure), _incl_fac(1000, 0, 1)
        else:
            return _time_loop_callback(sys, state, f, measure, verbose, thermal, seed, γ,
                                       grad_gamma_sign)
    ms = None
    if P.tdump > 0:
        ms = np.zeros(nsteps // P.tdump + 2)
        o.mass_series = _ptr_f64(ms)
    snap_buf = None
    if snapshots is not None and dumping > 0:
        nsnap = sum(1 for t in range(state.t + 1, state.t + nsteps + 1) if t % dumping == 0)
        snap_buf = np.zeros((max(1, nsnap), state.L))
        o.snap_every = int(dumping)
        o.snap_series = _ptr_f64(snap_buf)
    check(lib().swalbe1d_time_loop(state.handle, C.byref(prm), nsteps, C.byref(o)), "time_loop")
    t0 = state.t
    state.t = t0 + nsteps
    if snap_buf is not None:
        k = 0
        for t in range(t0 + 1, t0 + nsteps + 1):
            if t % dumping == 0:
                row = t // dumping - 1
                if 0 <= row < snapshots.shape[0]:
                    snapshots[row, :] = snap_buf[k]
                k += 1
    if ms is not None:
        k = 0
        for t in range(t0 + 1, t0 + nsteps + 1):
            if t % P.tdump == 0:
                if verbose:
                    print(f"Time step {t} mass is {round(ms[k], 3)}")
                if mass is not None:
                    mass.append(ms[k])
                k += 1
    if dh is not None:
        args[0].extend(dh.tolist())
        return state
    if wet is not None:
        measure.extend(int(v) for v in wet)
        return state, measure
    if f is not None:
        return state, measure
    return state


def _time_loop_callback(sys, state, f: Callable, measure, verbose, thermal_on, seed, γ,
                        grad_gamma_sign):
    """Per-op loop for an arbitrary callback (src/simulate.jl:159-178)."""
    P = sys.param
    for _ in range(P.Tmax):
        t = state.t + 1
        if P.tdump > 0 and t % P.tdump == 0 and verbose:
            print(f"Time step {t} mass is {round(state.reduce()['sum'], 3)}")
        filmpressure(state, sys, γ=γ)
        h_grad_p(state)
        slippage(state, sys)
        if thermal_on:
            thermal(state, sys, seed=seed, step=t)
        force_sum(state, thermal_on, grad_gamma_sign)
        f(measure, state)
        equilibrium(state, sys)
        BGKandStream(state, sys)
        moments(state)
        state.t = t
    return state, measure


def ensemble_time_loop(systems: Sequence, states: Sequence[HipState1D], nsteps: Optional[int] = None,
                       thermal=False, seed=0, materialize: bool = True, γ=None,
                       grad_gamma_sign=0, θ=None, threads: int = 0) -> float:
    """Steps many 1D films in one launch (one workgroup per film), each with
    its own SysConst_1D.  Per-film options may be given as sequences
    (thermal, seed, γ, θ, grad_gamma_sign); each film ends exactly as its own
    time_loop(..., variant=2) would.  Returns the launch's device time in ms."""
    n = len(states)
    if len(systems) != n:
        raise ValueError("one system per state")
    steps = systems[0].param.Tmax if nsteps is None else int(nsteps)
    per = lambda v, i: v[i] if isinstance(v, (list, tuple)) else v  # noqa: E731
    prms = (_lib.Params * n)()
    opts = (_lib.StepOpts1D * n)()
    for i, (sy, st) in enumerate(zip(systems, states)):
        prms[i] = sy.param.cparams()
        o = _options(sy, st, per(thermal, i), per(seed, i), materialize, 0, per(grad_gamma_sign, i))
        g = per(γ, i)
        if g is not None:
            if np.ndim(g) > 0:
                st.set("gamma", g)
                o.use_gamma_field = 1
            else:
                prms[i].gamma = float(g)
        th = per(θ, i) if not (isinstance(θ, np.ndarray) and θ.ndim == 1) else θ
        if th is not None:
            if np.ndim(th) > 0:
                st.set("theta", th)
                o.use_theta_field = 1
            else:
                prms[i].theta = float(th)
        o.threads = int(threads)
        opts[i] = o
    handles = (C.c_void_p * n)(*[st.handle.value for st in states])
    ms = C.c_double(0.0)
    check(lib().swalbe1d_ensemble_time_loop_timed(handles, n, prms, steps, opts, C.byref(ms)),
          "ensemble_time_loop")
    for st in states:
        st.t += steps
    return ms.value


def resident_max_length(T=np.float64, tau_is_one: bool = True) -> int:
    """Longest film that runs with the whole film in LDS on this GPU."""
    L = C.c_int32(0)
    check(lib().swalbe1d_resident_max_length(_lib.F64 if np.dtype(T).type is np.float64 else _lib.F32,
                                             int(tau_is_one), C.byref(L)), "resident_max_length")
    return L.value


# ---------------------------------------------------------------- raw-vector forms


def grad_f(output, f, dgrad=None, a=None):
    """∇f!(output::Vector, f, dgrad[, a]) (src/differences.jl:208-228) on host vectors."""
    return _raw1(lambda dt, o, x, aa, L: lib().swalbe1d_grad_ptr(dt, o, x, aa, L, None),
                 output, f, a)


def laplacian_f(output, f, dgrad=None):
    """∇²f!(output, f::Vector, dgrad) (src/differences.jl:77-85) on host vectors."""
    return _raw1(lambda dt, o, x, aa, L: lib().swalbe1d_laplacian_ptr(dt, o, x, L, None),
                 output, f, None)


def _raw1(call, output, f, a):
    f = np.asarray(f)
    dt = np.float32 if f.dtype == np.float32 else np.float64
    L = f.shape[0]
    nbytes = L * np.dtype(dt).itemsize
    bufs = []
    try:
        for arr in (output, f, a):
            if arr is None:
                bufs.append(None)
                continue
            p = C.c_void_p()
            check(lib().swalbe_malloc(C.byref(p), nbytes), "malloc")
            bufs.append(p)
            host = np.ascontiguousarray(np.broadcast_to(np.asarray(arr, dtype=dt), (L,)))
            check(lib().swalbe_memcpy_h2d(p, host.ctypes.data, nbytes), "h2d")
        check(call(_lib.F64 if dt is np.float64 else _lib.F32, bufs[0], bufs[1], bufs[2], L),
              "raw 1D op")
        out = np.empty(L, dtype=dt)
        check(lib().swalbe_memcpy_d2h(out.ctypes.data, bufs[0], nbytes), "d2h")
        output[...] = out
    finally:
        for p in bufs:
            if p is not None:
                lib().swalbe_free(p)
    return output


# ---------------------------------------------------------------- initial values


def singledroplet(height: np.ndarray, radius, θ, center, hcrit: float = 0.05):
    """singledroplet(height::Vector, radius, θ, center; hcrit) (src/initialvalues.jl:226-242):
    a cap (cos(asin(r/R)) - cos πθ) R, then every h <= hcrit set to hcrit.
    `center` is 1-based as in the reference."""
    from .core import cospi
    L = height.shape[0]
    i = np.arange(1, L + 1, dtype=float)
    circ = np.sqrt((i - center) ** 2)
    cap = (np.cos(np.arcsin(np.minimum(circ / radius, 1.0))) - cospi(θ)) * radius
    h = np.where(circ <= radius, cap, hcrit)
    height[...] = np.where(h <= hcrit, hcrit, h)
    return height


def two_droplets(sys, r1=230, r2=230, θ1=1 / 9, θ2=1 / 9, center=None, **kw):
    """two_droplets(sys::Consts_1D; r₁, r₂, θ₁, θ₂, center) (src/initialvalues.jl:277-323):
    cap₁ + cap₂ - 0.05 on a 0.05 precursor.  Unicode keywords r₁, θ₁, … work too."""
    from .core import cospi
    r1 = kw.pop("r₁", r1)
    r2 = kw.pop("r₂", r2)
    θ1 = kw.pop("θ₁", θ1)
    θ2 = kw.pop("θ₂", θ2)
    if kw:
        raise TypeError(f"unknown keywords {sorted(kw)}")
    L = sys.L
    center = (L / 3, 2 * L / 3) if center is None else center
    i = np.arange(1, L + 1, dtype=float)

    def cap(c, r, th):
        circ = np.sqrt((i - c) ** 2)
        d = np.where(circ <= r, (np.cos(np.arcsin(np.minimum(circ / r, 1.0))) - cospi(th)) * r, 0.05)
        return np.where(d < 0, 0.05, d)

    h = cap(center[0], r1, θ1) + cap(center[1], r2, θ2) - 0.05
    return np.where(h < 0, 0.05, h)


def randinterface(height: np.ndarray, h0: float, eps: float, seed: int = 0) -> None:
    """randinterface!(height::Vector, h₀, ϵ) (src/initialvalues.jl:23-33):
    h₀ (1 + ϵ N(0,1)), normal x of the counter stream."""
    from .initialvalues import normals
    z0, _ = normals(height.shape[0], seed)
    height[...] = h0 * (1.0 + eps * z0)


# ---------------------------------------------------------------- drivers


def run_flat(sys: SysConst_1D, verbos: bool = True, T=np.float64):
    """run_flat(sys::SysConst_1D) (src/simulate.jl:247-256)."""
    print("Simulating a flat interface without driving forces (nothing should happen) in one dimension")
    state = Sys(sys, T=T)
    state.height[...] = 1.0
    time_loop(sys, state, verbose=verbos)
    return state.height


def run_random(sys: SysConst_1D, h0: float = 1.0, eps: float = 0.01, verbos: bool = True,
               seed: int = 0, T=np.float64):
    """run_random(sys::SysConst_1D; h₀, ϵ) (src/simulate.jl:296-304)."""
    print("Simulating a random undulated interface in one dimension")
    state = Sys(sys, T=T)
    randinterface(state.height, h0, eps, seed=seed)
    equilibrium(state, sys)
    time_loop(sys, state, 1 / 9, verbose=verbos)
    return state.height


def run_rayleightaylor(sys: SysConst_1D, k: int = 15, h0: float = 1.0, eps: float = 0.001,
                       verbos: bool = True, T=np.float64):
    """run_rayleightaylor(sys::SysConst_1D; k, h₀, ϵ) (src/simulate.jl:360-377):
    h = h₀ (1 + ϵ sin(2π k i / (L - 1))), i = 1..L; returns (height, Δh)."""
    print("Simulating the Rayleigh Taylor instability in one dimension")
    state = Sys(sys, T=T)
    for i in range(1, sys.L + 1):
        state.height[i - 1] = h0 * (1 + eps * math.sin(2 * math.pi * k * i / (sys.L - 1)))
    diff = []
    equilibrium(state, sys)
    time_loop(sys, state, diff, verbose=verbos)
    return state.height, diff


def run_dropletrelax(sys: SysConst_1D, radius=20, θ0=1 / 6, center=None, verbos: bool = True,
                     T=np.float64):
    """run_dropletrelax(sys::SysConst_1D; radius, θ₀, center) (src/simulate.jl:401-416)."""
    print("Simulating an out of equilibrium droplet in one dimensions")
    area = []
    state = Sys(sys, T=T)
    singledroplet(state.height, radius, θ0, sys.L // 2 if center is None else center)
    equilibrium(state, sys)
    time_loop(sys, state, wetted, area, verbose=verbos)
    return state.height, area


def run_dropletpatterned(sys: SysConst_1D, radius=20, θ0=1 / 6, center=None, θs=None,
                         verbos: bool = True, T=np.float64):
    """run_dropletpatterned(sys::SysConst_1D; radius, θ₀, center, θₛ) (src/simulate.jl:441-460)."""
    print("Simulating a droplet on a patterned substrate in one dimension")
    θs = np.full(sys.L, 1 / 2) if θs is None else θs
    state = Sys(sys, T=T)
    singledroplet(state.height, radius, θ0, sys.L // 2 if center is None else center)
    equilibrium(state, sys)
    time_loop(sys, state, θs, verbose=verbos)
    return state.height


def run_dropletforced(sys: SysConst_1D, radius=20, θ0=1 / 6, center=None, θs=None, f=0.0,
                      verbos: bool = True, T=np.float64):
    """run_dropletforced(sys::SysConst_1D; radius, θ₀, center, θₛ, f) (src/simulate.jl:486-503):
    inclination!(f, state) each step; returns (height, vel)."""
    print("Simulating a sliding droplet in one dimension")
    state = Sys(sys, T=T)
    singledroplet(state.height, radius, θ0, sys.L // 2 if center is None else center)
    equilibrium(state, sys)
    print("Starting the lattice Boltzmann time loop")
    time_loop(sys, state, inclination, f, verbose=verbos)
    return state.height, state.vel


def run_gamma(sys: SysConst_1D, gamma, r1=115, r2=115, θ0=1 / 9, verbos: bool = True,
              dump: int = 100, fluid: Optional[np.ndarray] = None, T=np.float64):
    """run_gamma(sys, gamma::Vector; r₁, r₂, θ₀, verbos, dump, fluid)
    (src/simulate.jl:518-560): two droplets coalescing under a surface-tension
    field; F = -h∇p - slip - ∇γ with ∇γ!(state) of the γ field; returns the
    height snapshots (Tmax ÷ dump, L)."""
    print("Simulating droplet coalecense with surface tension gardient")
    gamma = np.asarray(gamma, dtype=float)
    fluid = np.zeros((sys.param.Tmax // dump, sys.L)) if fluid is None else fluid
    state = Sys(sys, T=T, kind="gamma")
    state.height = two_droplets(sys, r1=r1, r2=r2, θ1=θ0, θ2=θ0, center=(sys.L / 3, 2 * sys.L / 3))
    equilibrium(state, sys)
    state.set("gamma", gamma)
    grad_gamma(state)
    print("Starting the lattice Boltzmann time loop")
    mass = [] if verbos else None
    time_loop(sys, state, γ=gamma, grad_gamma_sign=-1, snapshots=fluid, dumping=dump, mass=mass)
    if verbos:
        for k, m in enumerate(mass):
            print(f"Time step {(k + 1) * sys.param.tdump} mass is {round(m, 3)}")
    return fluid


__all__ = [
    "SysConst_1D", "SysConstWithBound_1D", "HipState1D", "Sys", "obslist1D", "obslist",
    "equilibrium", "BGKandStream", "moments", "filmpressure", "grad_gamma", "h_grad_p",
    "slippage", "thermal", "inclination", "force_sum", "wetted", "time_loop",
    "ensemble_time_loop", "resident_max_length", "grad_f", "laplacian_f", "singledroplet",
    "two_droplets", "randinterface", "run_flat", "run_random", "run_rayleightaylor",
    "run_dropletrelax", "run_dropletpatterned", "run_dropletforced", "run_gamma",
]
