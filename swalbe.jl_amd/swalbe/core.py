"""Swalbe.jl's 2D API (parameters, state, ops, time loop) over libswalbe_hip.

Names follow the reference without the trailing `!` (Python identifiers
cannot carry it) and with ∇ spelled out: h∇p! -> h_grad_p, ∇f! -> grad_f,
∇²f! -> laplacian_f.  Arrays handed to or returned by the state use the
reference's layout: numpy arrays of shape (Lx, Ly) or (Lx, Ly, 9) in Fortran
(column-major) order, so a[i, j, k] is Julia's a[i+1, j+1, k+1].
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Callable, Optional

import numpy as np

from . import _lib
from ._lib import DIST_FIELDS, FIELDS, check, lib
from .fields import HipField, LiveFields

def _one_d():
    from . import one_d
    return one_d


def _is1d(x) -> bool:
    return getattr(x, "_is_1d", False)


# ---------------------------------------------------------------- parameters


class Taumucs:
    """Run-time constants (src/initialize.jl:43-60).  Keyword names follow the
    reference (τ, cₛ, μ, δ, kbt, γ, n, m, hmin, hcrit, θ, g); ASCII spellings
    (tau, cs, mu, delta, gamma, theta) are accepted too."""

    _ascii = {"tau": "τ", "cs": "cₛ", "mu": "μ", "delta": "δ", "gamma": "γ", "theta": "θ"}

    def __init__(self, Tmax: int = 1000, tdump: Optional[int] = None, **kw):
        for a, u in self._ascii.items():
            if a in kw:
                kw[u] = kw.pop(a)
        self.Tmax = int(Tmax)
        self.tdump = int(tdump) if tdump is not None else self.Tmax // 10
        self.τ = float(kw.pop("τ", 1.0))
        self.cₛ = float(kw.pop("cₛ", 1 / math.sqrt(3.0)))
        mu = kw.pop("μ", None)
        self.μ = float(self.cₛ ** 2 * (self.τ - 0.5) if mu is None else mu)
        self.δ = float(kw.pop("δ", 1.0))
        self.kbt = float(kw.pop("kbt", 0.0))
        self.γ = float(kw.pop("γ", 0.01))
        self.n = int(kw.pop("n", 9))
        self.m = int(kw.pop("m", 3))
        self.hmin = float(kw.pop("hmin", 0.1))
        self.hcrit = float(kw.pop("hcrit", 0.05))
        self.θ = float(kw.pop("θ", 1 / 9))
        self.g = float(kw.pop("g", 0.0))
        if kw:
            raise TypeError(f"Taumucs: unknown keywords {sorted(kw)}")

    def cparams(self) -> _lib.Params:
        return _lib.Params(Tmax=self.Tmax, tdump=self.tdump, tau=self.τ, cs=self.cₛ, mu=self.μ,
                           delta=self.δ, kbt=self.kbt, gamma=self.γ, n=self.n, m=self.m,
                           hmin=self.hmin, hcrit=self.hcrit, theta=self.θ, g=self.g)

    def __repr__(self):
        return (f"Taumucs(Tmax={self.Tmax}, tdump={self.tdump}, τ={self.τ}, μ={self.μ}, δ={self.δ}, "
                f"kbt={self.kbt}, γ={self.γ}, n={self.n}, m={self.m}, hmin={self.hmin}, "
                f"hcrit={self.hcrit}, θ={self.θ}, g={self.g})")


class SysConst:
    """Lattice size plus Taumucs (src/initialize.jl:76-81)."""

    def __init__(self, Lx: int = 256, Ly: int = 256, param: Optional[Taumucs] = None):
        self.Lx, self.Ly = int(Lx), int(Ly)
        self.param = param if param is not None else Taumucs()


# ---------------------------------------------------------------- state


class HipState(LiveFields):
    """A film resident on the GPU: every array of State{T,3}
    (src/initialize.jl:149-168) plus kbtx/kbty and an optional θ field.

    `state.height` is a live view of the device array (a HipField): reads
    see the device's values and in-place writes (`state.height[...] = 1.0`,
    `state.height[i, j] = v`, `state.ftemp[...] = state.fout`,
    `state.Fx += a`) reach the device before the next op, as the reference's
    in-place writes into its state arrays do.  Assigning an array uploads it
    and assigning a scalar fills the field (state.height .= 1.0)."""

    _names = dict(FIELDS, **{"h∇px": FIELDS["hgradpx"], "h∇py": FIELDS["hgradpy"]})

    def __init__(self, Lx: int, Ly: int, T=np.float64, kind: str = "simple"):
        T = np.dtype(T).type
        if T not in (np.float64, np.float32):
            raise TypeError("T must be Float64 or Float32")
        if kind not in ("simple", "thermal"):
            raise ValueError("kind must be 'simple' or 'thermal'")
        self._init_fields()
        object.__setattr__(self, "Lx", int(Lx))
        object.__setattr__(self, "Ly", int(Ly))
        object.__setattr__(self, "T", T)
        object.__setattr__(self, "kind", kind)
        h = C.c_void_p()
        check(lib().swalbe_create(int(Lx), int(Ly), _lib.F64 if T is np.float64 else _lib.F32,
                                  _lib.SIMPLE if kind == "simple" else _lib.THERMAL, C.byref(h)),
              "swalbe_create")
        object.__setattr__(self, "_h", h)
        object.__setattr__(self, "t", 0)  # steps taken by time_loop (noise counter)

    # lifetime
    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().swalbe_destroy(self._h)
            object.__setattr__(self, "_h", None)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def dtype(self) -> int:
        return _lib.F64 if self.T is np.float64 else _lib.F32

    def _field_shape(self, name):
        return (self.Lx, self.Ly, 9) if name in DIST_FIELDS else (self.Lx, self.Ly)

    def _download(self, name: str, dtype=None) -> np.ndarray:
        out = np.empty(self._field_shape(name), dtype=np.dtype(dtype or self.T), order="F")
        check(lib().swalbe_download(self._h, self._names[name], out.ctypes.data, out.nbytes),
              f"download {name}")
        return out

    def _upload(self, name: str, value) -> None:
        fid = self._names[name]
        if np.ndim(value) == 0:
            check(lib().swalbe_fill(self._h, fid, float(value)), f"fill {name}")
            return
        arr = np.asarray(value)
        shape = self._field_shape(name)
        if arr.shape != shape:
            arr = np.broadcast_to(arr, shape)
        dt = np.float64 if arr.dtype != np.float32 else np.float32
        arr = np.asfortranarray(arr, dtype=dt)
        check(lib().swalbe_upload(self._h, fid, arr.ctypes.data, arr.nbytes), f"upload {name}")

    def synchronize(self):
        check(lib().swalbe_synchronize(self.handle), "synchronize")

    def set_reference_height(self, href: float, g: float = 0.0):
        """FP32 storage offset (deviation encoding); no effect for FP64."""
        check(lib().swalbe_set_reference_height(self.handle, float(href), float(g)),
              "set_reference_height")

    def copy(self, dst: str, src: str) -> None:
        """Device-side `dst .= src` between fields of the same kind."""
        check(lib().swalbe_copy(self.handle, self._names[dst], self._names[src]), "copy")

    def init_height(self, init) -> None:
        """Initial height from a FilmInit spec (the same kernel a batch film
        is initialised with)."""
        check(lib().swalbe_init_height(self.handle, C.byref(init)), "init_height")

    def reduce(self, name: str = "height", thresh: float = 0.0):
        out = (C.c_double * 4)()
        check(lib().swalbe_reduce(self.handle, self._names[name], float(thresh), out), "reduce")
        return {"sum": out[0], "max": out[1], "min": out[2], "count": int(out[3])}


def Sys(sysc: SysConst, device: str = "GPU", T=np.float64, kind: str = "simple") -> HipState:
    """Allocates a state like Sys(sysc, device; T, kind) (src/initialize.jl:491-572):
    all fields zero and height = 1, resident on the GPU.  A SysConst_1D gives
    the 1D state of Sys(sysc::Consts_1D; T, kind) (:587-616)."""
    if isinstance(sysc, _one_d().SysConst_1D):
        return _one_d().Sys(sysc, T=T, kind=kind)
    if device not in ("GPU", "HIP"):
        raise ValueError("this package implements the GPU device (device = 'GPU' or 'HIP')")
    return HipState(sysc.Lx, sysc.Ly, T=T, kind=kind)


# ---------------------------------------------------------------- device arrays for raw forms


class DeviceArray:
    """A caller-owned device array for the positional (raw-array) op forms."""

    def __init__(self, shape, dtype=np.float64, data=None):
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        check(lib().swalbe_malloc(C.byref(p), max(self.nbytes, 1)), "swalbe_malloc")
        self.ptr = p
        if data is not None:
            self.upload(data)

    @classmethod
    def from_numpy(cls, a):
        a = np.asarray(a)
        dt = np.float32 if a.dtype == np.float32 else np.float64
        return cls(a.shape, dt, a)

    def upload(self, a):
        a = np.asfortranarray(np.broadcast_to(np.asarray(a, dtype=self.dtype), self.shape))
        check(lib().swalbe_memcpy_h2d(self.ptr, a.ctypes.data, self.nbytes), "memcpy h2d")

    def get(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype, order="F")
        check(lib().swalbe_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes), "memcpy d2h")
        return out

    def __del__(self):
        try:
            if self.ptr and self.ptr.value:
                lib().swalbe_free(self.ptr)
                self.ptr = None
        except Exception:
            pass


def _dt(arrs) -> int:
    for a in arrs:
        if isinstance(a, DeviceArray):
            return _lib.F32 if a.dtype == np.float32 else _lib.F64
        if isinstance(a, np.ndarray) and a.dtype == np.float32:
            return _lib.F32
    return _lib.F64


class _Raw:
    """Stages numpy arguments of a raw-array call in device memory and writes
    the outputs back into the caller's arrays (the reference mutates them)."""

    def __init__(self, dtype: int, shape2d):
        self.dtype = dtype
        self.np_dtype = np.float64 if dtype == _lib.F64 else np.float32
        self.shape2d = tuple(shape2d)
        self.back = []

    def arg(self, a, shape, out=False):
        if a is None:
            return None
        if isinstance(a, DeviceArray):
            return a.ptr
        d = DeviceArray(shape, self.np_dtype, np.broadcast_to(np.asarray(a, dtype=float), shape))
        if out:
            self.back.append((a, d))
        self.keep = getattr(self, "keep", []) + [d]
        return d.ptr

    def finish(self):
        for a, d in self.back:
            a[...] = d.get()


def _shape2(a):
    return tuple(a.shape[:2])


# ---------------------------------------------------------------- ops


def equilibrium(*args, **kw):
    """equilibrium!(state, sys) or equilibrium!(feq, h, ux, uy, vsq, g)
    (src/equilibrium.jl:63-128)."""
    if _is1d(args[0]):
        return _one_d().equilibrium(*args, **kw)
    if isinstance(args[0], HipState):
        state, sys = args[0], args[1]
        check(lib().swalbe_equilibrium(state.handle, float(kw.get("g", sys.param.g))), "equilibrium")
        return None
    feq, h, ux, uy, vsq, g = args
    s2 = _shape2(feq)
    R = _Raw(_dt(args[:5]), s2)
    check(lib().swalbe_equilibrium_ptr(R.dtype, R.arg(feq, s2 + (9,), True), R.arg(h, s2),
                                       R.arg(ux, s2), R.arg(uy, s2), R.arg(vsq, s2, True), float(g),
                                       s2[0], s2[1], None), "equilibrium")
    R.finish()


def BGKandStream(*args, τ=None, tau=None):
    """BGKandStream!(state, sys) or BGKandStream!(fout, feq, ftemp, Fx, Fy, τ)
    (src/collide.jl:70-161)."""
    if _is1d(args[0]):
        return _one_d().BGKandStream(args[0], args[1], τ=τ if τ is not None else tau)
    if isinstance(args[0], HipState):
        state, sys = args[0], args[1]
        t = τ if τ is not None else (tau if tau is not None else sys.param.τ)
        check(lib().swalbe_bgk_and_stream(state.handle, float(t)), "BGKandStream")
        return None
    fout, feq, ftemp, Fx, Fy, t = args
    s2 = _shape2(feq)
    R = _Raw(_dt(args[:5]), s2)
    check(lib().swalbe_bgk_and_stream_ptr(R.dtype, R.arg(fout, s2 + (9,), True), R.arg(feq, s2 + (9,)),
                                          R.arg(ftemp, s2 + (9,), True), R.arg(Fx, s2), R.arg(Fy, s2),
                                          float(t), s2[0], s2[1], None), "BGKandStream")
    R.finish()


def moments(*args):
    """moments!(state) or moments!(height, velx, vely, fout) (src/moments.jl:43-73)."""
    if _is1d(args[0]):
        return _one_d().moments(args[0])
    if isinstance(args[0], HipState):
        check(lib().swalbe_moments(args[0].handle), "moments")
        return None
    h, ux, uy, fout = args
    s2 = _shape2(fout)
    R = _Raw(_dt(args), s2)
    check(lib().swalbe_moments_ptr(R.dtype, R.arg(h, s2, True), R.arg(ux, s2, True), R.arg(uy, s2, True),
                                   R.arg(fout, s2 + (9,)), s2[0], s2[1], None), "moments")
    R.finish()


def filmpressure(*args, θ=None, γ=None, n=None, m=None, hmin=None, hcrit=None, **ascii_kw):
    """filmpressure!(state, sys; θ, γ, n, m, hmin, hcrit) (src/pressure.jl:119-193,
    power_broad form; θ may be a field) or the raw-array method
    filmpressure!(output, f, dgrad, γ, θ, n, m, hmin, hcrit) (src/pressure.jl:72-117,
    fast_93/fast_32 only, DomainError otherwise)."""
    θ = ascii_kw.pop("theta", θ)
    γ = ascii_kw.pop("gamma", γ)
    if ascii_kw:
        raise TypeError(f"unknown keywords {sorted(ascii_kw)}")
    if _is1d(args[0]):
        return _one_d().filmpressure(args[0], args[1], θ=θ, γ=γ, n=n, m=m, hmin=hmin, hcrit=hcrit)
    if isinstance(args[0], HipState):
        state, sys = args[0], args[1]
        P = sys.param
        th = P.θ if θ is None else θ
        use_field = 0
        if np.ndim(th) > 0:
            state.set("theta", th)
            use_field, th = 1, 0.0
        check(lib().swalbe_filmpressure(state.handle, float(P.γ if γ is None else γ), float(th),
                                        int(P.n if n is None else n), int(P.m if m is None else m),
                                        float(P.hmin if hmin is None else hmin),
                                        float(P.hcrit if hcrit is None else hcrit), use_field, 0),
              "filmpressure")
        return None
    output, f, _dgrad, g_, th, n_, m_, hmin_, hcrit_ = args
    s2 = _shape2(f)
    R = _Raw(_dt([output, f]), s2)
    tf = None
    ths = 0.0
    if np.ndim(th) > 0 or isinstance(th, DeviceArray):
        tf = R.arg(th, s2)
    else:
        ths = float(th)
    check(lib().swalbe_filmpressure_ptr(R.dtype, R.arg(output, s2, True), R.arg(f, s2), float(g_), ths, tf,
                                        int(n_), int(m_), float(hmin_), float(hcrit_), 1, s2[0], s2[1],
                                        None), "filmpressure")
    R.finish()


def grad_f(outx, outy, f, *rest):
    """∇f!(outx, outy, f), ∇f!(outx, outy, f, a), ∇f!(outx, outy, f, dgrad, a)
    (src/differences.jl:153-206); on vectors the 1D ∇f!(output, f, dgrad[, a])."""
    if np.ndim(outx) == 1:
        return _one_d().grad_f(outx, outy, f, rest[0] if rest else None)
    a = rest[-1] if rest else None
    s2 = _shape2(f)
    R = _Raw(_dt([outx, outy, f]), s2)
    check(lib().swalbe_grad_ptr(R.dtype, R.arg(outx, s2, True), R.arg(outy, s2, True), R.arg(f, s2),
                                R.arg(a, s2), s2[0], s2[1], None), "grad_f")
    R.finish()


def laplacian_f(output, f, γ):
    """∇²f!(output, f, γ) (src/differences.jl:57-75); on vectors the 1D
    ∇²f!(output, f, dgrad) (:77-85), whose third argument is scratch."""
    if np.ndim(output) == 1:
        return _one_d().laplacian_f(output, f)
    s2 = _shape2(f)
    R = _Raw(_dt([output, f]), s2)
    check(lib().swalbe_laplacian_ptr(R.dtype, R.arg(output, s2, True), R.arg(f, s2), float(γ), s2[0],
                                     s2[1], None), "laplacian_f")
    R.finish()


def h_grad_p(state: HipState):
    """h∇p!(state) (src/forcing.jl:168-221)."""
    if _is1d(state):
        return _one_d().h_grad_p(state)
    check(lib().swalbe_hgradp(state.handle), "h_grad_p")


def h_grad_p_raw(hgpx, hgpy, p, h):
    """The h∇p! stencil on caller arrays: h ⋅ ∇p."""
    s2 = _shape2(p)
    R = _Raw(_dt([hgpx, hgpy, p, h]), s2)
    check(lib().swalbe_hgradp_ptr(R.dtype, R.arg(hgpx, s2, True), R.arg(hgpy, s2, True), R.arg(p, s2),
                                  R.arg(h, s2), s2[0], s2[1], None), "h_grad_p")
    R.finish()


def slippage(*args):
    """slippage!(state, sys) or slippage!(slipx, slipy, h, ux, uy, δ, μ) (src/forcing.jl:42-66)."""
    if _is1d(args[0]):
        return _one_d().slippage(*args)
    if isinstance(args[0], HipState):
        state, sys = args
        check(lib().swalbe_slippage(state.handle, sys.param.δ, sys.param.μ), "slippage")
        return None
    sx, sy, h, ux, uy, d, mu = args
    s2 = _shape2(h) if np.ndim(h) == 2 else _shape2(sx)
    R = _Raw(_dt(args[:5]), s2)
    check(lib().swalbe_slippage_ptr(R.dtype, R.arg(sx, s2, True), R.arg(sy, s2, True), R.arg(h, s2),
                                    R.arg(ux, s2), R.arg(uy, s2), float(d), float(mu), s2[0], s2[1],
                                    None), "slippage")
    R.finish()


def thermal(*args, seed: int = 0, step: Optional[int] = None):
    """thermal!(state, sys) or thermal!(fx, fy, h, kbt, μ, δ) (src/forcing.jl:297-320).
    randn! is replaced by normals keyed (seed, step, node); step defaults to
    the state's step counter."""
    if _is1d(args[0]):
        return _one_d().thermal(args[0], args[1], seed=seed, step=step)
    if isinstance(args[0], HipState):
        state, sys = args
        st = state.t if step is None else step
        check(lib().swalbe_thermal(state.handle, sys.param.kbt, sys.param.μ, sys.param.δ,
                                   int(seed), int(st)), "thermal")
        return None
    fx, fy, h, kbt, mu, d = args
    s2 = _shape2(fx)
    R = _Raw(_dt([fx, fy, h]), s2)
    check(lib().swalbe_thermal_ptr(R.dtype, R.arg(fx, s2, True), R.arg(fy, s2, True), R.arg(h, s2),
                                   float(kbt), float(mu), float(d), int(seed), int(step or 0), s2[0],
                                   s2[1], None), "thermal")
    R.finish()


def _incl_fac(t, tstart, tsmooth):
    return 0.5 + 0.5 * math.tanh((t - tstart) / tsmooth)


def inclination(α, state, t: int = 1000, tstart: int = 0, tsmooth: int = 1):
    """inclination!(α, state; t, tstart, tsmooth) (src/forcing.jl:363-377);
    α a number for a 1D state (:379-389)."""
    if _is1d(state):
        return _one_d().inclination(α, state, t=t, tstart=tstart, tsmooth=tsmooth)
    check(lib().swalbe_inclination(state.handle, float(α[0]), float(α[1]),
                                   _incl_fac(t, tstart, tsmooth)), "inclination")


def force_sum(state: HipState, thermal: bool = False):
    """state.Fx .= -state.h∇px .- state.slipx [.- state.kbtx] (src/simulate.jl:18-19)."""
    if _is1d(state):
        return _one_d().force_sum(state, thermal)
    check(lib().swalbe_force_sum(state.handle, int(bool(thermal))), "force_sum")


# scalar helpers of src/pressure.jl:363-422 (host arithmetic)
def power_broad(arg, n: int):
    t = type(arg)(1)
    for _ in range(int(n)):
        t *= arg
    return t


def fast_93(arg: float) -> float:
    t = arg * arg * arg
    return t * t * t - t


def fast_32(arg: float) -> float:
    return arg * arg * arg - arg * arg


def cospi(x: float) -> float:
    """cos(πx) exactly as the library evaluates κ(θ)."""
    return lib().swalbe_host_cospi(float(x))


# ---------------------------------------------------------------- time loop


def wetted(area_size, *args, hthresh: float = 0.055):
    """wetted!(area_size, state) (src/measures.jl:13-17): push count(h > hthresh);
    wetted!(area_size, maxheight, height, t) (:6-11): area_size[t] and
    maxheight[t] (t 1-based) of a state or a height array, measured on the GPU."""
    if len(args) == 1:
        area_size.append(int(args[0].reduce("height", hthresh)["count"]))
        return None
    maxheight, height, t = args
    m = measure(height, hthresh, surface=False)
    area_size[t - 1] = m["count"]
    maxheight[t - 1] = m["max"]
    return None


def measure(height, thresh: float = 0.055, θ=0.0, surface: bool = True, mask=None):
    """The on-device measurement behind wetted!, fluid_dry! and surfacearea!:
    {count(h > thresh), max h, liquid-vapour area, reduced energy} of a
    state's height or of a height array; θ a scalar or a field; `mask` (a
    uint8 array of the film's shape, Fortran order) receives h > thresh."""
    if _is1d(height):
        raise TypeError("the surface measures are defined for 2D films")
    out = (C.c_double * 4)()
    if isinstance(height, HipState):
        use_field = 0
        th = θ
        if np.ndim(θ) > 0:
            height.set("theta", θ)
            use_field, th = 1, 0.0
        mb = None if mask is None else mask.ctypes.data
        check(lib().swalbe_measure(height.handle, float(thresh), use_field, float(th),
                                   int(bool(surface)), out, mb, 0 if mask is None else mask.nbytes),
              "measure")
    else:
        h = np.asarray(height)
        s2 = _shape2(h)
        R = _Raw(_dt([h]), s2)
        th_f = R.arg(θ, s2) if np.ndim(θ) > 0 else None
        dmask = DeviceArray(s2, dtype=np.uint8) if mask is not None else None
        dout = DeviceArray((4,), dtype=np.float64)
        check(lib().swalbe_measure_ptr(R.dtype, R.arg(h, s2), th_f, float(θ) if th_f is None else 0.0,
                                       s2[0], s2[1], float(thresh),
                                       None if dmask is None else dmask.ptr, dout.ptr, None),
              "measure")
        R.finish()
        vals = dout.get()
        for k in range(4):
            out[k] = vals[k]
        if mask is not None:
            mask[...] = dmask.get().reshape(s2, order="F")
    return {"count": int(out[0]), "max": out[1], "area": out[2], "energy": out[3]}


def fluid_dry(fluid, dummy, height, t: int, hthresh: float = 0.055):
    """fluid_dry!(fluid, dummy, height, t) (src/measures.jl:23-28): dummy = h > hthresh
    and fluid[t, :] = vec(dummy) (t 1-based), the mask taken on the GPU."""
    shape = (height.Lx, height.Ly) if isinstance(height, HipState) else _shape2(np.asarray(height))
    m = np.zeros(shape, dtype=np.uint8, order="F")
    measure(height, hthresh, surface=False, mask=m)
    dummy[...] = m.astype(bool)
    fluid[t - 1, :] = m.ravel(order="F").astype(bool)


def surfacearea(area_lv, red_energy, height, θ, dhx=None, dhy=None, dgrad=None, surface=None,
                t: int = 1, htresh: float = 0.055):
    """surfacearea!(area_lv, red_energy, height, θ, ∇hx, ∇hy, dgrad, surface, t)
    (src/measures.jl:126-168): liquid-vapour area Σ sqrt(|∇h|² + 1) over
    h > htresh (∇f_simple! gradients) and the reduced energy, stored at t
    (1-based); θ a scalar or a field.  ∇hx / ∇hy, when given, receive the
    gradients; dgrad and surface are scratch here."""
    m = measure(height, htresh, θ=θ, surface=True)
    area_lv[t - 1] = m["area"]
    red_energy[t - 1] = m["energy"]
    if dhx is not None and dhy is not None:
        h = height.height if isinstance(height, HipState) else height
        grad_f_simple(dhx, dhy, h, dgrad)


def grad_f_simple(outx, outy, f, dgrad=None):
    """∇f_simple!(outputx, outputy, f, dgrad) (src/measures.jl:175-192)."""
    s2 = _shape2(f)
    R = _Raw(_dt([outx, outy, f]), s2)
    check(lib().swalbe_grad_simple_ptr(R.dtype, R.arg(outx, s2, True), R.arg(outy, s2, True),
                                       R.arg(f, s2), s2[0], s2[1], None), "grad_f_simple")
    R.finish()


def t0(hᵦ: float = 0.07, γ: float = 0.01, μ: float = 1 / 6, θ: float = 1 / 6, **kw):
    """t0(; hᵦ, γ, μ, θ) (src/measures.jl:56-61): the spinodal dewetting time
    scale 3μ/(γ q⁴) with q² = hᵦ(1 - cos πθ)(2 - 3hᵦ); returns (t0, q²)."""
    hᵦ = kw.pop("hb", hᵦ)
    γ = kw.pop("gamma", γ)
    μ = kw.pop("mu", μ)
    θ = kw.pop("theta", θ)
    if kw:
        raise TypeError(f"unknown keywords {sorted(kw)}")
    qsq = (hᵦ * (1 - cospi(θ))) * (2 - 3 * hᵦ)
    return (3 * μ) / (γ * (qsq * qsq)), qsq


def store_mode(Lx: int, Ly: int, fp64: bool, final_outputs_only: bool = False) -> int:
    """The store mode (swalbe_step_opts.nontemporal) measured fastest for a
    film on MI355X; every mode gives the same bits.  FP64 films: 2 (f and p
    stream past the caches).  FP32 films with more than 32 MB of inputs (h, ux,
    uy beyond the eight 4 MB L2s): 1 (every output streams) when every step
    writes its outputs, 0 (all cached) when only the last step does; smaller
    FP32 films the other way round."""
    if fp64:
        return 2
    big = Lx * Ly * 12 > (32 << 20)
    return (0 if final_outputs_only else 1) if big else (1 if final_outputs_only else 0)


def time_loop(sys: SysConst, state: HipState, *args, verbose: bool = False, thermal: bool = False,
              seed: int = 0, materialize: bool = True, tile_y: int = 0, mass: Optional[list] = None,
              snapshots: Optional[dict] = None, snapshot_every: int = 0,
              maxheight: Optional[list] = None, area_lv: Optional[list] = None,
              red_energy: Optional[list] = None, fluid: Optional[np.ndarray] = None,
              measure_thresh: float = 0.055, surface_theta_field: bool = False,
              final_outputs_only: bool = False, variant: int = 0,
              nontemporal: Optional[int] = None, **kw1d):
    """time_loop(sys, state), time_loop(sys, state, θ), time_loop(sys, state, Δh)
    and time_loop(sys, state, f, measure) (src/simulate.jl:6-96).

    The step is one fused kernel per time step.  Known callbacks run fused:
    wetted (on-device count per step) and inclination (body force).  Any other
    callable f(measure, state) runs the per-op kernels so that f sees the
    state between the force sum and equilibrium!, as the reference does.
    `thermal=True` subtracts the thermal! noise from F each step
    (scripts/Rivulet_stability.jl:123-124).  `snapshots` (a dict) receives
    "h_<t>" for every step t divisible by `snapshot_every`, as snapshot!
    (src/measures.jl:99-105) would record them, copied off the device while
    the following steps run.  maxheight / area_lv / red_energy (lists) and
    fluid (a (Tmax, Lx*Ly) bool array) receive wetted!'s maximum height,
    surfacearea!'s area and reduced energy and fluid_dry!'s mask of every
    step's input height, taken on the device without host synchronisation
    (threshold measure_thresh; energy from the θ field with
    surface_theta_field).  final_outputs_only (τ = 1) writes the
    distributions and the pressure on the last step only: the same results
    with 28 instead of 64 bytes per node and step in FP32.  A 1D state runs
    the 1D loop (its own keywords:
    variant, γ, grad_gamma_sign, snapshots + dumping).  nontemporal
    sets the stores' cache policy (0, 1, 2: the same bits); by default
    store_mode() picks the one measured fastest for the film."""
    if _is1d(state):
        return _one_d().time_loop(sys, state, *args, verbose=verbose, thermal=thermal, seed=seed,
                                  materialize=materialize, mass=mass, snapshots=snapshots,
                                  variant=variant, **kw1d)
    if kw1d:
        raise TypeError(f"unknown keywords {sorted(kw1d)}")
    P = sys.param
    prm = P.cparams()
    nsteps = P.Tmax
    o = _lib.StepOpts()
    o.thermal, o.seed, o.t0, o.materialize, o.tile_y = int(thermal), int(seed), int(state.t), int(materialize), tile_y
    o.variant = int(variant)  # FP32 kernel: 0 default, 1 / 2 LDS tiles, 3 strips
    o.nontemporal = int(nontemporal) if nontemporal is not None else store_mode(
        sys.Lx, sys.Ly, state.T is np.float64, final_outputs_only)
    keep = []
    dh = None
    f = measure = None
    if len(args) == 1:
        a = args[0]
        if isinstance(a, list):
            dh = np.zeros(nsteps)
            o.dh_series = dh.ctypes.data_as(C.POINTER(C.c_double))
        elif np.ndim(a) > 0:
            state.set("theta", a)
            o.use_theta_field = 1
        else:
            prm.theta = float(a)
    elif len(args) == 2:
        f, measure = args
    wet = None
    if f is not None:
        if f is wetted:
            wet = np.zeros(nsteps, dtype=np.int64)
            o.wet_series = wet.ctypes.data_as(C.POINTER(C.c_int64))
            o.wet_thresh = 0.055
        elif f is inclination:
            o.body, o.body_ax, o.body_ay, o.body_fac = 1, float(measure[0]), float(measure[1]), _incl_fac(1000, 0, 1)
        else:
            return _time_loop_callback(sys, state, f, measure, verbose, thermal, seed)
    ms = None
    if P.tdump > 0:
        ms = np.zeros(nsteps // P.tdump + 2)
        o.mass_series = ms.ctypes.data_as(C.POINTER(C.c_double))
    mser = {}
    for key, lst, field in (("maxh", maxheight, "maxh_series"), ("area", area_lv, "area_series"),
                            ("energy", red_energy, "energy_series")):
        if lst is not None:
            mser[key] = (np.zeros(max(1, nsteps)), lst)
            setattr(o, field, mser[key][0].ctypes.data_as(C.POINTER(C.c_double)))
    fl = None
    if fluid is not None:
        fl = np.zeros((max(1, nsteps), state.Lx * state.Ly), dtype=np.uint8)
        o.fluid_series = fl.ctypes.data_as(C.POINTER(C.c_uint8))
    o.surf_thresh, o.surf_theta_field = float(measure_thresh), int(bool(surface_theta_field))
    o.final_outputs_only = int(bool(final_outputs_only))
    snap_t = []
    if snapshots is not None and snapshot_every > 0:
        snap_t = [t for t in range(state.t + 1, state.t + nsteps + 1) if t % snapshot_every == 0]
        snap_buf = np.zeros((max(1, len(snap_t)), state.Lx * state.Ly))
        o.snap_every = int(snapshot_every)
        o.snap_series = snap_buf.ctypes.data_as(C.POINTER(C.c_double))
    check(lib().swalbe_time_loop(state.handle, C.byref(prm), nsteps, C.byref(o)), "time_loop")
    for k, t in enumerate(snap_t):
        snapshots[f"h_{t}"] = snap_buf[k].reshape((state.Lx, state.Ly), order="F").copy()
    t0 = state.t
    state.t = t0 + nsteps
    for arr, lst in mser.values():
        lst.extend(arr[:nsteps].tolist())
    if fl is not None:
        fluid[:nsteps, :] = fl[:nsteps].astype(bool)
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


def _time_loop_callback(sys, state, f: Callable, measure, verbose, thermal, seed):
    """Per-op loop for an arbitrary measurement callback (src/simulate.jl:69-96)."""
    P = sys.param
    for _ in range(P.Tmax):
        t = state.t + 1
        if P.tdump > 0 and t % P.tdump == 0 and verbose:
            print(f"Time step {t} mass is {round(state.reduce()['sum'], 3)}")
        filmpressure(state, sys)
        h_grad_p(state)
        slippage(state, sys)
        if thermal:
            check(lib().swalbe_thermal(state.handle, P.kbt, P.μ, P.δ, int(seed), int(t)), "thermal")
        force_sum(state, thermal)
        f(measure, state)
        equilibrium(state, sys)
        BGKandStream(state, sys)
        moments(state)
        state.t = t
    return state, measure
