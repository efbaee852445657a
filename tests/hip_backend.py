"""KAT backends over the HIP library: the positional (raw-array) op forms
and the state forms.  Same numpy-in / numpy-out interface as the oracle
backend of test_oracle_golden.py."""
import ctypes as C

import numpy as np

import swalbe as S
from swalbe import _lib


class RawBackend:
    """Positional-array forms (equilibrium!(feq, h, ux, uy, vsq, g), ...)."""

    cospi = staticmethod(S.cospi)
    power_broad = staticmethod(S.power_broad)
    singledroplet = staticmethod(S.singledroplet)

    def __init__(self, T=np.float64):
        self.T = T

    def _z(self, shape):
        return np.zeros(shape, dtype=self.T, order="F")

    def grad(self, f, a):
        ox, oy = self._z(f.shape), self._z(f.shape)
        if a is None:
            S.grad_f(ox, oy, f.astype(self.T))
        else:
            S.grad_f(ox, oy, f.astype(self.T), a.astype(self.T))
        return ox, oy

    def laplacian(self, f, gamma):
        o = self._z(f.shape)
        S.laplacian_f(o, f.astype(self.T), gamma)
        return o

    def filmpressure(self, h, gamma, theta, n, m, hmin, hcrit, raw=False):
        if raw:
            p = self._z(h.shape)
            S.filmpressure(p, h.astype(self.T), None, gamma, theta, n, m, hmin, hcrit)
            return p
        return StateBackend(self.T).filmpressure(h, gamma, theta, n, m, hmin, hcrit, raw=False)

    def hgradp(self, p, h):
        ox, oy = self._z(p.shape), self._z(p.shape)
        S.h_grad_p_raw(ox, oy, p.astype(self.T), h.astype(self.T))
        return ox, oy

    def slippage(self, h, ux, uy, delta, mu):
        sx, sy = self._z(h.shape), self._z(h.shape)
        S.slippage(sx, sy, h.astype(self.T), ux.astype(self.T), uy.astype(self.T), delta, mu)
        return sx, sy

    def thermal(self, h, kbt, mu, delta, seed, step):
        kx, ky = self._z(h.shape), self._z(h.shape)
        S.thermal(kx, ky, h.astype(self.T), kbt, mu, delta, seed=seed, step=step)
        return kx, ky

    def inclination(self, Fx, Fy, h, ax, ay, fac):
        dt = _lib.F64 if self.T == np.float64 else _lib.F32
        a, b, c = (S.DeviceArray.from_numpy(np.asfortranarray(x, dtype=self.T)) for x in (Fx, Fy, h))
        _lib.check(_lib.lib().swalbe_inclination_ptr(dt, a.ptr, b.ptr, c.ptr, ax, ay, fac,
                                                     h.shape[0], h.shape[1], None))
        return a.get(), b.get()

    def equilibrium(self, h, ux, uy, g):
        feq = self._z(h.shape + (9,))
        vsq = self._z(h.shape)
        S.equilibrium(feq, h.astype(self.T), ux.astype(self.T), uy.astype(self.T), vsq, g)
        return feq, vsq

    def bgk_and_stream(self, feq, ftemp, Fx, Fy, tau):
        fout = self._z(feq.shape)
        ft = np.asfortranarray(ftemp, dtype=self.T).copy()
        S.BGKandStream(fout, feq.astype(self.T), ft, Fx.astype(self.T), Fy.astype(self.T), tau)
        return fout, ft

    def moments(self, f):
        s2 = f.shape[:2]
        h, ux, uy = self._z(s2), self._z(s2), self._z(s2)
        S.moments(h, ux, uy, f.astype(self.T))
        return h, ux, uy

    def run_flat(self, Lx, Ly, Tmax, tdump):
        sys = S.SysConst(Lx=Lx, Ly=Ly, param=S.Taumucs(Tmax=Tmax, tdump=tdump))
        return S.run_flat(sys, "GPU", verbos=False, T=self.T)


class StateBackend(RawBackend):
    """State forms: upload into a HipState, run the op, download."""

    def _state(self, shape, **fields):
        st = S.HipState(shape[0], shape[1], T=self.T)
        for k, v in fields.items():
            st.set(k, v)
        return st

    def filmpressure(self, h, gamma, theta, n, m, hmin, hcrit, raw=False):
        st = self._state(h.shape, height=h)
        if raw:
            _lib.check(_lib.lib().swalbe_filmpressure(st.handle, gamma, float(theta), n, m, hmin,
                                                      hcrit, 0, 1))
        else:
            sys = S.SysConst(h.shape[0], h.shape[1], S.Taumucs())
            S.filmpressure(st, sys, θ=theta, γ=gamma, n=n, m=m, hmin=hmin, hcrit=hcrit)
        return st.pressure

    def hgradp(self, p, h):
        st = self._state(p.shape, pressure=p, height=h)
        S.h_grad_p(st)
        return st.hgradpx, st.hgradpy

    def slippage(self, h, ux, uy, delta, mu):
        st = self._state(h.shape, height=h, velx=ux, vely=uy)
        S.slippage(st, S.SysConst(h.shape[0], h.shape[1], S.Taumucs(δ=delta, μ=mu)))
        return st.slipx, st.slipy

    def thermal(self, h, kbt, mu, delta, seed, step):
        st = self._state(h.shape, height=h)
        S.thermal(st, S.SysConst(h.shape[0], h.shape[1], S.Taumucs(kbt=kbt, μ=mu, δ=delta)),
                  seed=seed, step=step)
        return st.kbtx, st.kbty

    def inclination(self, Fx, Fy, h, ax, ay, fac):
        st = self._state(h.shape, Fx=Fx, Fy=Fy, height=h)
        _lib.check(_lib.lib().swalbe_inclination(st.handle, ax, ay, fac))
        return st.Fx, st.Fy

    def equilibrium(self, h, ux, uy, g):
        st = self._state(h.shape, height=h, velx=ux, vely=uy)
        S.equilibrium(st, S.SysConst(h.shape[0], h.shape[1], S.Taumucs(g=g)))
        return st.feq, st.vsq

    def bgk_and_stream(self, feq, ftemp, Fx, Fy, tau):
        st = self._state(feq.shape[:2], feq=feq, ftemp=ftemp, Fx=Fx, Fy=Fy)
        S.BGKandStream(st, S.SysConst(feq.shape[0], feq.shape[1], S.Taumucs(τ=tau)))
        return st.fout, st.ftemp

    def moments(self, f):
        st = self._state(f.shape[:2], fout=f)
        S.moments(st)
        return st.height, st.velx, st.vely
