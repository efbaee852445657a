"""ctypes view of the CPU FP64 oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY.  This module is the checker for the HIP product:
only tests/, the smoke check and the benchmark's CPU-baseline leg import it.

Arrays follow Julia's column-major layout: numpy arrays of shape (Lx, Ly) or
(Lx, Ly, 9) in Fortran order, so element (i, j, k) sits at i + Lx*j + Lx*Ly*k
exactly as in the reference.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_D = C.POINTER(C.c_double)
_I64 = C.POINTER(C.c_int64)


class Params(C.Structure):
    _fields_ = [
        ("tau", C.c_double), ("mu", C.c_double), ("delta", C.c_double), ("kbt", C.c_double),
        ("gamma", C.c_double), ("hmin", C.c_double), ("hcrit", C.c_double),
        ("theta", C.c_double), ("g", C.c_double), ("n", C.c_int32), ("m", C.c_int32),
        ("tdump", C.c_int64),
    ]


class StateC(C.Structure):
    _fields_ = [("Lx", C.c_int32), ("Ly", C.c_int32)] + [
        (name, _D) for name in (
            "fout", "ftemp", "feq", "height", "velx", "vely", "vsq", "pressure",
            "Fx", "Fy", "slipx", "slipy", "hgpx", "hgpy", "dgrad", "kbtx", "kbty")
    ]


class LoopOpts(C.Structure):
    _fields_ = [
        ("theta_field", _D), ("thermal", C.c_int32), ("body", C.c_int32),
        ("seed", C.c_uint64), ("t0", C.c_int64), ("body_ax", C.c_double),
        ("body_ay", C.c_double), ("body_fac", C.c_double), ("dh_series", _D),
        ("wet_series", _I64), ("wet_thresh", C.c_double), ("mass_series", _D),
    ]


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_log.restype = C.c_double
        L.orc_log.argtypes = [C.c_double]
        L.orc_cospi.restype = C.c_double
        L.orc_cospi.argtypes = [C.c_double]
        L.orc_sincospi.argtypes = [C.c_double, _D, _D]
        L.orc_power_broad.restype = C.c_double
        L.orc_power_broad.argtypes = [C.c_double, C.c_int]
        L.orc_fast_93.restype = C.c_double
        L.orc_fast_93.argtypes = [C.c_double]
        L.orc_fast_32.restype = C.c_double
        L.orc_fast_32.argtypes = [C.c_double]
        L.orc_uniform.restype = C.c_double
        L.orc_uniform.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32]
        L.orc_normal_pair.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32, _D, _D]
        L.orc_noise_index.restype = C.c_uint32
        L.orc_noise_index.argtypes = [C.c_int, C.c_int, C.c_int]
        L.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32)] * 3
        L.orc_filmpressure_raw.restype = C.c_int
        L.orc_time_loop.restype = C.c_int
        L.orc_time_loop.argtypes = [C.POINTER(StateC), C.POINTER(Params), C.c_int64,
                                    C.POINTER(LoopOpts)]
        _lib = L
    return _lib


def _f(a, shape=None):
    """Fresh Fortran-ordered float64 copy (inputs are never modified)."""
    a = np.asarray(a, dtype=np.float64)
    if shape is not None and a.shape != shape:
        a = np.broadcast_to(a, shape)
    return np.array(a, dtype=np.float64, order="F", copy=True)


def _p(a):
    return a.ctypes.data_as(_D) if a is not None else None


# ---------------------------------------------------------------- math
def philox4x32_10(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return tuple(int(v) for v in o)


def log(x):
    return lib().orc_log(float(x))


def sincospi(x):
    s, c = C.c_double(), C.c_double()
    lib().orc_sincospi(float(x), C.byref(s), C.byref(c))
    return s.value, c.value


def cospi(x):
    return lib().orc_cospi(float(x))


def normal_pair(seed, step, node, tag):
    a, b = C.c_double(), C.c_double()
    lib().orc_normal_pair(seed, step, node, tag, C.byref(a), C.byref(b))
    return a.value, b.value


def noise_index(x, y, Lx):
    """Stream index of node (x, y): vertical node pairs share one Philox call."""
    return lib().orc_noise_index(int(x), int(y), int(Lx))


def uniform(seed, step, node, tag):
    return lib().orc_uniform(seed, step, node, tag)


def power_broad(x, n):
    return lib().orc_power_broad(float(x), int(n))


def fast_93(x):
    return lib().orc_fast_93(float(x))


def fast_32(x):
    return lib().orc_fast_32(float(x))


# ---------------------------------------------------------------- ops
def circshift(src, shift):
    src = _f(src)
    Lx, Ly = src.shape
    dst = np.zeros((Lx, Ly), order="F")
    lib().orc_circshift(_p(dst), _p(src), Lx, Ly, int(shift[0]), int(shift[1]))
    return dst


def equilibrium(h, ux, uy, g):
    h = _f(h)
    Lx, Ly = h.shape
    ux, uy = _f(ux, h.shape), _f(uy, h.shape)
    feq = np.zeros((Lx, Ly, 9), order="F")
    vsq = np.zeros((Lx, Ly), order="F")
    lib().orc_equilibrium(_p(feq), _p(h), _p(ux), _p(uy), _p(vsq), C.c_double(g), Lx, Ly)
    return feq, vsq


def bgk_and_stream(feq, ftemp, Fx, Fy, tau):
    """Returns (fout, ftemp) after BGKandStream! (both hold the streamed f)."""
    feq = _f(feq)
    Lx, Ly, _ = feq.shape
    ftemp = np.array(_f(ftemp, feq.shape), order="F", copy=True)
    Fx, Fy = _f(Fx, (Lx, Ly)), _f(Fy, (Lx, Ly))
    fout = np.zeros((Lx, Ly, 9), order="F")
    lib().orc_bgk_and_stream(_p(fout), _p(feq), _p(ftemp), _p(Fx), _p(Fy), C.c_double(tau), Lx, Ly)
    return fout, ftemp


def moments(fout):
    fout = _f(fout)
    Lx, Ly, _ = fout.shape
    h = np.zeros((Lx, Ly), order="F")
    ux = np.zeros((Lx, Ly), order="F")
    uy = np.zeros((Lx, Ly), order="F")
    lib().orc_moments(_p(h), _p(ux), _p(uy), _p(fout), Lx, Ly)
    return h, ux, uy


def filmpressure(h, gamma, theta, n, m, hmin, hcrit, raw=False):
    """State form (power_broad) or, with raw=True, the fast_93/fast_32 form.
    theta may be a scalar or an (Lx, Ly) field.  Returns p (raises
    ValueError where the reference raises DomainError)."""
    h = _f(h)
    Lx, Ly = h.shape
    p = np.zeros((Lx, Ly), order="F")
    dgrad = np.zeros((Lx, Ly, 8), order="F")
    tf = None
    th = 0.0
    if np.ndim(theta) == 0:
        th = float(theta)
    else:
        tf = _f(theta, h.shape)
    args = (_p(p), _p(h), _p(dgrad), C.c_double(gamma), C.c_double(th), _p(tf),
            int(n), int(m), C.c_double(hmin), C.c_double(hcrit), Lx, Ly)
    if raw:
        if lib().orc_filmpressure_raw(*args) != 0:
            raise ValueError("DomainError: (n, m) not in {(9,3), (3,2)}")
    else:
        lib().orc_filmpressure(*args)
    return p


def grad(f, a=None):
    f = _f(f)
    Lx, Ly = f.shape
    ox = np.zeros((Lx, Ly), order="F")
    oy = np.zeros((Lx, Ly), order="F")
    af = None if a is None else _f(a, f.shape)
    lib().orc_grad(_p(ox), _p(oy), _p(f), _p(af), Lx, Ly)
    return ox, oy


def laplacian(f, gamma):
    f = _f(f)
    Lx, Ly = f.shape
    o = np.zeros((Lx, Ly), order="F")
    lib().orc_laplacian(_p(o), _p(f), C.c_double(gamma), Lx, Ly)
    return o


def hgradp(p, h):
    p = _f(p)
    Lx, Ly = p.shape
    h = _f(h, p.shape)
    ox = np.zeros((Lx, Ly), order="F")
    oy = np.zeros((Lx, Ly), order="F")
    dgrad = np.zeros((Lx, Ly, 8), order="F")
    lib().orc_hgradp(_p(ox), _p(oy), _p(p), _p(h), _p(dgrad), Lx, Ly)
    return ox, oy


def slippage(h, ux, uy, delta, mu):
    h = _f(h)
    Lx, Ly = h.shape
    ux, uy = _f(ux, h.shape), _f(uy, h.shape)
    sx = np.zeros((Lx, Ly), order="F")
    sy = np.zeros((Lx, Ly), order="F")
    lib().orc_slippage(_p(sx), _p(sy), _p(h), _p(ux), _p(uy), C.c_double(delta), C.c_double(mu), Lx, Ly)
    return sx, sy


def thermal(h, kbt, mu, delta, seed, step):
    h = _f(h)
    Lx, Ly = h.shape
    kx = np.zeros((Lx, Ly), order="F")
    ky = np.zeros((Lx, Ly), order="F")
    lib().orc_thermal(_p(kx), _p(ky), _p(h), C.c_double(kbt), C.c_double(mu), C.c_double(delta),
                      C.c_uint64(seed), C.c_uint64(step), Lx, Ly)
    return kx, ky


def inclination(Fx, Fy, h, ax, ay, fac):
    h = _f(h)
    Lx, Ly = h.shape
    Fx = np.array(_f(Fx, h.shape), order="F", copy=True)
    Fy = np.array(_f(Fy, h.shape), order="F", copy=True)
    lib().orc_inclination(_p(Fx), _p(Fy), _p(h), C.c_double(ax), C.c_double(ay), C.c_double(fac), Lx, Ly)
    return Fx, Fy


def force_sum(hgpx, hgpy, slipx, slipy, kbtx=None, kbty=None):
    hgpx = _f(hgpx)
    Lx, Ly = hgpx.shape
    hgpy, slipx, slipy = _f(hgpy, hgpx.shape), _f(slipx, hgpx.shape), _f(slipy, hgpx.shape)
    kx = None if kbtx is None else _f(kbtx, hgpx.shape)
    ky = None if kbty is None else _f(kbty, hgpx.shape)
    Fx = np.zeros((Lx, Ly), order="F")
    Fy = np.zeros((Lx, Ly), order="F")
    lib().orc_force_sum(_p(Fx), _p(Fy), _p(hgpx), _p(hgpy), _p(slipx), _p(slipy), _p(kx), _p(ky), Lx, Ly)
    return Fx, Fy


def init_uniform(Lx, Ly, seed, base, amp):
    h = np.zeros((Lx, Ly), order="F")
    lib().orc_init_uniform(_p(h), Lx, Ly, C.c_uint64(seed), C.c_double(base), C.c_double(amp))
    return h


def init_normal(Lx, Ly, seed, base, amp):
    h = np.zeros((Lx, Ly), order="F")
    lib().orc_init_normal(_p(h), Lx, Ly, C.c_uint64(seed), C.c_double(base), C.c_double(amp))
    return h


# ---------------------------------------------------------------- driver
FIELDS = ("fout", "ftemp", "feq", "height", "velx", "vely", "vsq", "pressure",
          "Fx", "Fy", "slipx", "slipy", "hgpx", "hgpy", "dgrad", "kbtx", "kbty")


class State:
    """State{Float64,3} + kbtx/kbty as Fortran-ordered numpy arrays, zeroed
    except height = 1 (as Sys does, src/initialize.jl:491-508)."""

    def __init__(self, Lx, Ly):
        self.Lx, self.Ly = Lx, Ly
        for name in FIELDS:
            if name in ("fout", "ftemp", "feq"):
                shape = (Lx, Ly, 9)
            elif name == "dgrad":
                shape = (Lx, Ly, 8)
            else:
                shape = (Lx, Ly)
            setattr(self, name, np.zeros(shape, order="F"))
        self.height[...] = 1.0

    def cstruct(self):
        s = StateC()
        s.Lx, s.Ly = self.Lx, self.Ly
        for name in FIELDS:
            arr = getattr(self, name)
            assert arr.flags.f_contiguous and arr.dtype == np.float64
            setattr(s, name, _p(arr))
        return s

    def set(self, name, value):
        arr = getattr(self, name)
        arr[...] = value


def params(tau=1.0, delta=1.0, kbt=0.0, gamma=0.01, n=9, m=3, hmin=0.1, hcrit=0.05,
           theta=1 / 9, g=0.0, tdump=100, mu=None):
    cs = 1 / np.sqrt(3.0)
    if mu is None:
        mu = cs ** 2 * (tau - 0.5)
    return Params(tau=tau, mu=mu, delta=delta, kbt=kbt, gamma=gamma, hmin=hmin, hcrit=hcrit,
                  theta=theta, g=g, n=n, m=m, tdump=tdump)


def time_loop(state: State, prm: Params, nsteps: int, *, theta_field=None, thermal=False,
              seed=0, t0=0, body=None, dh=False, wetted=None, mass=False):
    """Runs nsteps of the reference time loop.  Returns a dict of series."""
    o = LoopOpts()
    keep = []
    if theta_field is not None:
        tf = _f(theta_field, (state.Lx, state.Ly))
        keep.append(tf)
        o.theta_field = _p(tf)
    o.thermal = int(bool(thermal))
    o.seed = seed
    o.t0 = t0
    out = {}
    if body is not None:
        o.body = 1
        o.body_ax, o.body_ay, o.body_fac = body
    if dh:
        out["dh"] = np.zeros(nsteps)
        o.dh_series = _p(out["dh"])
    if wetted is not None:
        out["wetted"] = np.zeros(nsteps, dtype=np.int64)
        o.wet_series = out["wetted"].ctypes.data_as(_I64)
        o.wet_thresh = wetted
    if mass:
        out["mass"] = np.zeros(max(1, nsteps // max(1, prm.tdump) + 1))
        o.mass_series = _p(out["mass"])
    s = state.cstruct()
    rc = lib().orc_time_loop(C.byref(s), C.byref(prm), int(nsteps), C.byref(o))
    assert rc == 0
    return out


# ---------------------------------------------------------------- one dimension (D1Q3)
FIELDS_1D = ("fout", "ftemp", "feq", "height", "vel", "pressure", "F", "slip", "hgp", "kbt")


class State1DC(C.Structure):
    _fields_ = [("L", C.c_int32)] + [(name, _D) for name in FIELDS_1D]


class LoopOpts1D(C.Structure):
    _fields_ = [
        ("theta_field", _D), ("thermal", C.c_int32), ("body", C.c_int32),
        ("alpha", C.c_double), ("body_fac", C.c_double), ("border1", _D), ("border2", _D),
        ("seed", C.c_uint64), ("t0", C.c_int64), ("dh_series", _D), ("wet_series", _I64),
        ("wet_thresh", C.c_double), ("mass_series", _D), ("gamma_field", _D),
        ("grad_gamma", _D), ("grad_gamma_sign", C.c_int32), ("interior", _D),
    ]


class State1D:
    """State_1D{Float64} (src/initialize.jl:276-290) as numpy arrays: zeros
    except height = 1, as Sys(sysc::Consts_1D) allocates them (:587-600)."""

    def __init__(self, L):
        self.L = L
        for name in FIELDS_1D:
            shape = (L, 3) if name in ("fout", "ftemp", "feq") else (L,)
            setattr(self, name, np.zeros(shape, order="F"))
        self.height[...] = 1.0

    def cstruct(self):
        s = State1DC()
        s.L = self.L
        for name in FIELDS_1D:
            arr = getattr(self, name)
            assert arr.flags.f_contiguous and arr.dtype == np.float64
            setattr(s, name, _p(arr))
        return s


def equilibrium_1d(h, u, g):
    h = _f(h)
    L = h.shape[0]
    feq = np.zeros((L, 3), order="F")
    lib().orc1d_equilibrium(_p(feq), _p(h), _p(_f(u, (L,))), C.c_double(g), L)
    return feq


def bgk_and_stream_1d(feq, ftemp, F, tau, border1=None, border2=None):
    feq = _f(feq)
    L = feq.shape[0]
    ft = _f(ftemp, (L, 3))
    fout = np.zeros((L, 3), order="F")
    b1 = _f(border1, (L,)) if border1 is not None else None
    b2 = _f(border2, (L,)) if border2 is not None else None
    lib().orc1d_bgk_and_stream(_p(fout), _p(feq), _p(ft), _p(_f(F, (L,))), C.c_double(tau),
                               _p(b1), _p(b2), L)
    return fout, ft


def moments_1d(f):
    f = _f(f)
    L = f.shape[0]
    h, u = np.zeros(L), np.zeros(L)
    lib().orc1d_moments(_p(h), _p(u), _p(f), L)
    return h, u


def filmpressure_1d(h, gamma, theta, n, m, hmin, hcrit, theta_field=None, gamma_field=None):
    h = _f(h)
    L = h.shape[0]
    p = np.zeros(L)
    tf = _f(theta_field, (L,)) if theta_field is not None else None
    gf = _f(gamma_field, (L,)) if gamma_field is not None else None
    lib().orc1d_filmpressure(_p(p), _p(h), C.c_double(gamma), _p(gf), C.c_double(theta), _p(tf),
                             n, m, C.c_double(hmin), C.c_double(hcrit), L)
    return p


def grad_gamma_1d(gamma, h=None, delta=1.0, form=0):
    """∇γ!(state) (form 0) or ∇γ!(state, sys) (form 1)."""
    g = _f(gamma)
    L = g.shape[0]
    out = np.zeros(L)
    hh = _f(h, (L,)) if h is not None else np.ones(L)
    lib().orc1d_grad_gamma(_p(out), _p(g), _p(hh), C.c_double(delta), int(form), L)
    return out


def filmpressure_raw_1d(h, gamma, theta, n, m, hmin, hcrit):
    h = _f(h)
    L = h.shape[0]
    p = np.zeros(L)
    rc = lib().orc1d_filmpressure_raw(_p(p), _p(h), C.c_double(gamma), C.c_double(theta), n, m,
                                      C.c_double(hmin), C.c_double(hcrit), L)
    if rc != 0:
        raise ValueError("DomainError: (n, m) must be (9, 3) or (3, 2)")
    return p


def grad_1d(f, a=None):
    f = _f(f)
    L = f.shape[0]
    out = np.zeros(L)
    lib().orc1d_grad(_p(out), _p(f), _p(_f(a, (L,)) if a is not None else None), L)
    return out


def laplacian_1d(f):
    f = _f(f)
    out = np.zeros(f.shape[0])
    lib().orc1d_laplacian(_p(out), _p(f), f.shape[0])
    return out


def hgradp_1d(p, h):
    h = _f(h)
    L = h.shape[0]
    out = np.zeros(L)
    lib().orc1d_hgradp(_p(out), _p(_f(p, (L,))), _p(h), L)
    return out


def slippage_1d(h, u, delta, mu):
    h = _f(h)
    L = h.shape[0]
    out = np.zeros(L)
    lib().orc1d_slippage(_p(out), _p(h), _p(_f(u, (L,))), C.c_double(delta), C.c_double(mu), L)
    return out


def thermal_1d(h, kbt, mu, delta, seed, step):
    h = _f(h)
    L = h.shape[0]
    out = np.zeros(L)
    lib().orc1d_thermal(_p(out), _p(h), C.c_double(kbt), C.c_double(mu), C.c_double(delta),
                        C.c_uint64(seed), C.c_uint64(step), L)
    return out


def inclination_1d(F, h, alpha, fac):
    F = _f(F)
    lib().orc1d_inclination(_p(F), _p(_f(h, F.shape)), C.c_double(alpha), C.c_double(fac),
                            F.shape[0])
    return F


def obslist_1d(obs):
    """obslist1D(obs) (src/obstacle.jl:6-35): interior and the border masks
    [obsright, obsleft] of an obstacle vector, periodic neighbours."""
    obs = np.asarray(obs, dtype=float)
    right, left = np.roll(obs, -1), np.roll(obs, 1)  # obs[i+1], obs[i-1]
    interior = ((obs == 1) & (right == 1) & (left == 1)).astype(float)
    obsleft = ((obs == 1) & (left == 1)).astype(float)
    obsright = ((obs == 1) & (right == 1)).astype(float)
    return interior, (obsright, obsleft)


def time_loop_1d(state: State1D, prm: Params, nsteps: int, *, theta_field=None, thermal=False,
                 seed=0, t0=0, body=None, borders=None, dh=False, wetted=None, mass=False,
                 gamma_field=None, grad_gamma=None, grad_gamma_sign=-1, interior=None):
    """The 1D time loop; body = (α, fac); borders = (border1, border2)."""
    o = LoopOpts1D()
    keep = []
    if theta_field is not None:
        tf = _f(theta_field, (state.L,))
        keep.append(tf)
        o.theta_field = _p(tf)
    if gamma_field is not None:
        gf = _f(gamma_field, (state.L,))
        keep.append(gf)
        o.gamma_field = _p(gf)
    if grad_gamma is not None:
        gg = _f(grad_gamma, (state.L,))
        keep.append(gg)
        o.grad_gamma, o.grad_gamma_sign = _p(gg), int(grad_gamma_sign)
    if interior is not None:
        it = _f(interior, (state.L,))
        keep.append(it)
        o.interior = _p(it)
    o.thermal, o.seed, o.t0 = int(thermal), seed, t0
    if body is not None:
        o.body, o.alpha, o.body_fac = 1, body[0], body[1]
    if borders is not None:
        b1, b2 = _f(borders[0], (state.L,)), _f(borders[1], (state.L,))
        keep += [b1, b2]
        o.border1, o.border2 = _p(b1), _p(b2)
    out = {}
    if dh:
        out["dh"] = np.zeros(nsteps)
        o.dh_series = _p(out["dh"])
    if wetted is not None:
        out["wetted"] = np.zeros(nsteps, dtype=np.int64)
        o.wet_series = out["wetted"].ctypes.data_as(_I64)
        o.wet_thresh = wetted
    if mass:
        out["mass"] = np.zeros(max(1, nsteps // max(1, prm.tdump) + 1))
        o.mass_series = _p(out["mass"])
    L = lib()
    L.orc1d_time_loop.restype = C.c_int
    L.orc1d_time_loop.argtypes = [C.POINTER(State1DC), C.POINTER(Params), C.c_int64,
                                  C.POINTER(LoopOpts1D)]
    cs = state.cstruct()
    L.orc1d_time_loop(C.byref(cs), C.byref(prm), nsteps, C.byref(o))
    return out


def time_loop_1d_f32(h, prm: Params, nsteps: int, gamma_field=None, grad_gamma=None,
                     grad_gamma_sign=0, dtype=np.float32):
    """The 1D loop (src/simulate.jl:98-116, run_gamma's force :548) in numpy
    with every operation rounded to `dtype` in the library's order (periodic,
    τ = 1, g = 0, no noise): float32 is the reference for FP32 states, and
    float64 reproduces the C oracle to the bit (which pins this restatement).
    Returns (h, u) after nsteps from rest.  κ is formed in FP64 and rounded
    once, as the library does."""
    f32 = dtype
    assert prm.tau == 1.0 and prm.g == 0.0
    L = len(h)
    h = np.asarray(h, dtype=f32).copy()
    u = np.zeros(L, dtype=f32)
    kap = f32((((1.0 - cospi(prm.theta)) * float(prm.n - 1)) * float(prm.m - 1)) /
              (float(prm.n - prm.m) * prm.hmin))
    gam = np.full(L, f32(prm.gamma)) if gamma_field is None else np.asarray(gamma_field, dtype=f32)
    gg = None if grad_gamma is None else np.asarray(grad_gamma, dtype=f32)
    hmin, hcrit = f32(prm.hmin), f32(prm.hcrit)
    c6mu, c6d, c3dd = f32(6.0 * prm.mu), f32(6.0 * prm.delta), f32(3.0 * (prm.delta * prm.delta))
    zero, half, one, two = f32(0), f32(0.5), f32(1), f32(2)
    for _ in range(nsteps):
        hm, hp = np.roll(h, 1), np.roll(h, -1)
        a = hmin / (h + hcrit)
        pn = np.full(L, one)
        for _k in range(prm.n):
            pn = pn * a
        pm_ = np.full(L, one)
        for _k in range(prm.m):
            pm_ = pm_ * a
        p = -gam * (kap * (pn - pm_)) - gam * ((hm - two * h) + hp)
        hgp = (h * f32(-0.5)) * (np.roll(p, 1) - np.roll(p, -1))
        F = -hgp - (c6mu * h * u) / ((two * (h * h) + c6d * h) + c3dd)
        if gg is not None:
            F = F - gg if grad_gamma_sign < 0 else F + gg
        uu = u * u
        e0 = h * ((one - zero * h) - uu)
        e1 = h * ((zero * h + half * u) + half * uu)
        e2 = h * ((zero * h - half * u) + half * uu)
        f0 = zero * zero + one * e0
        f1 = np.roll((zero * zero + one * e1) + half * F, 1)
        f2 = np.roll((zero * zero + one * e2) - half * F, -1)
        h = (f0 + f1) + f2
        u = (f1 - f2) / h
    return h.astype(np.float64), u.astype(np.float64)


def grad_simple(f):
    """∇f_simple! (src/measures.jl:175-192) of an (Lx, Ly) array."""
    f = _f(f)
    Lx, Ly = f.shape
    gx, gy = np.zeros((Lx, Ly), order="F"), np.zeros((Lx, Ly), order="F")
    lib().orc_grad_simple(_p(gx), _p(gy), _p(f), Lx, Ly)
    return gx, gy


def measure(h, theta_field=None, theta=0.0, thresh=0.055):
    """wetted! / fluid_dry! / surfacearea! of a height field: returns
    ({count, max, area, energy}, mask)."""
    h = _f(h)
    Lx, Ly = h.shape
    out = np.zeros(4)
    mask = np.zeros((Lx, Ly), dtype=np.uint8, order="F")
    tf = _f(theta_field, (Lx, Ly)) if theta_field is not None else None
    lib().orc_measure(_p(h), _p(tf), C.c_double(theta), Lx, Ly, C.c_double(thresh), _p(out),
                      mask.ctypes.data_as(C.POINTER(C.c_ubyte)))
    return {"count": out[0], "max": out[1], "area": out[2], "energy": out[3]}, mask


def t0(hb=0.07, gamma=0.01, mu=1 / 6, theta=1 / 6):
    out = np.zeros(2)
    lib().orc_t0(C.c_double(hb), C.c_double(gamma), C.c_double(mu), C.c_double(theta), _p(out))
    return out[0], out[1]


def fused_time_loop(state: "State", prm: Params, nsteps: int, thermal=False, seed=0, t0=0):
    """The multi-core CPU baseline (swalbe_cpu_fused.c): the same step as
    time_loop in three OpenMP sweeps; the same bits."""
    L = lib()
    L.orc_fused_time_loop.restype = C.c_int
    L.orc_fused_time_loop.argtypes = [C.POINTER(StateC), C.POINTER(Params), C.c_int64, C.c_int,
                                      C.c_uint64, C.c_int64]
    cs = state.cstruct()
    rc = L.orc_fused_time_loop(C.byref(cs), C.byref(prm), int(nsteps), int(bool(thermal)),
                               int(seed), int(t0))
    if rc != 0:
        raise MemoryError("orc_fused_time_loop: allocation failed")


def fused_threads() -> int:
    return int(lib().orc_fused_threads())
