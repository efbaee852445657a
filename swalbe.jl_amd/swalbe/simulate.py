"""The run_* drivers of src/simulate.jl:236-560 on a GPU-resident state; a
SysConst_1D selects the 1D driver of the same name."""
from __future__ import annotations

import math

import numpy as np

from . import core, one_d
from .core import Sys, SysConst, equilibrium, inclination, time_loop, wetted
from .initialvalues import randinterface, singledroplet


def run_flat(sys: SysConst, device: str = "GPU", verbos: bool = True, T=np.float64):
    """run_flat (src/simulate.jl:236-245): a flat film stays flat.  Returns
    state.height (the state's live height field)."""
    if isinstance(sys, one_d.SysConst_1D):
        return one_d.run_flat(sys, verbos=verbos, T=T)
    print("Simulating a flat interface without driving forces (nothing should happen) in two dimensions")
    state = Sys(sys, device, T=T)
    state.height[...] = 1.0
    time_loop(sys, state, verbose=verbos)
    return state.height


def run_random(sys: SysConst, device: str = "GPU", h0: float = 1.0, eps: float = 0.01,
               verbos: bool = True, seed: int = 0, T=np.float64):
    """run_random (src/simulate.jl:286-294): a randomly undulated film
    (randinterface! writes into state.height)."""
    if isinstance(sys, one_d.SysConst_1D):
        return one_d.run_random(sys, h0=h0, eps=eps, verbos=verbos, seed=seed, T=T)
    print("Simulating a random undulated interface in two dimensions")
    state = Sys(sys, device, T=T)
    randinterface(state.height, h0, eps, seed=seed)
    equilibrium(state, sys)
    time_loop(sys, state, 1 / 9, verbose=verbos)
    return state.height


def run_rayleightaylor(sys: SysConst, device: str = "GPU", kx: int = 15, ky: int = 18, h0: float = 1.0,
                       eps: float = 0.001, verbos: bool = True, T=np.float64, **kw):
    """run_rayleightaylor (src/simulate.jl:338-358): the sine undulation is
    written node by node into state.height, as the reference's loop does;
    returns (height, Δh series)."""
    if isinstance(sys, one_d.SysConst_1D):
        return one_d.run_rayleightaylor(sys, h0=h0, eps=eps, verbos=verbos, T=T, **kw)
    print("Simulating the Rayleigh Taylor instability in two dimensions")
    state = Sys(sys, device, T=T)
    for i in range(1, sys.Lx + 1):
        for j in range(1, sys.Ly + 1):
            state.height[i - 1, j - 1] = h0 * (1 + eps * math.sin(2 * math.pi * kx * i / (sys.Lx - 1))
                                               * math.sin(2 * math.pi * ky * j / (sys.Ly - 1)))
    diff = []
    equilibrium(state, sys)
    time_loop(sys, state, diff, verbose=verbos)
    return state.height, diff


def run_dropletrelax(sys: SysConst, device: str = "GPU", radius=20, θ0=1 / 6, center=None,
                     verbos: bool = True, T=np.float64):
    """run_dropletrelax (src/simulate.jl:384-399): singledroplet writes into
    state.height; returns (height, wetted area series)."""
    if isinstance(sys, one_d.SysConst_1D):
        return one_d.run_dropletrelax(sys, radius=radius, θ0=θ0, center=center, verbos=verbos, T=T)
    print("Simulating an out of equilibrium droplet in two dimensions")
    center = center or (sys.Lx // 2, sys.Ly // 2)
    state = Sys(sys, device, T=T)
    singledroplet(state.height, radius, θ0, center)
    equilibrium(state, sys)
    area = []
    time_loop(sys, state, wetted, area, verbose=verbos)
    return state.height, area


def run_dropletpatterned(sys: SysConst, device: str = "GPU", radius=20, θ0=1 / 6, center=None,
                         θs=None, verbos: bool = True, T=np.float64):
    """run_dropletpatterned (src/simulate.jl:422-439): contact-angle field θₛ."""
    if isinstance(sys, one_d.SysConst_1D):
        return one_d.run_dropletpatterned(sys, radius=radius, θ0=θ0, center=center, θs=θs,
                                          verbos=verbos, T=T)
    print("Simulating a droplet on a patterned substrate in two dimensions")
    center = center or (sys.Lx // 2, sys.Ly // 2)
    θs = np.full((sys.Lx, sys.Ly), 1 / 9) if θs is None else θs
    state = Sys(sys, device, T=T)
    singledroplet(state.height, radius, θ0, center)
    equilibrium(state, sys)
    time_loop(sys, state, θs, verbose=verbos)
    return state.height


def run_dropletforced(sys: SysConst, device: str = "GPU", radius=20, θ0=1 / 6, center=None, fx=0.0,
                      fy=0.0, verbos: bool = True, T=np.float64, **kw):
    """run_dropletforced (src/simulate.jl:462-484): a droplet under a body force;
    returns (height, velx, vely)."""
    if isinstance(sys, one_d.SysConst_1D):
        return one_d.run_dropletforced(sys, radius=radius, θ0=θ0, center=center, verbos=verbos,
                                       T=T, **kw)
    bodyforce = np.zeros(2)
    bodyforce[0] = fx
    bodyforce[1] = fy
    print("Simulating a sliding droplet in two dimensions")
    center = center or (sys.Lx // 2, sys.Ly // 2)
    state = Sys(sys, device, T=T)
    singledroplet(state.height, radius, θ0, center)
    equilibrium(state, sys)
    print("Starting the lattice Boltzmann time loop")
    time_loop(sys, state, inclination, bodyforce, verbose=verbos)
    return state.height, state.velx, state.vely


run_gamma = one_d.run_gamma

__all__ = ["run_flat", "run_random", "run_rayleightaylor", "run_dropletrelax",
           "run_dropletpatterned", "run_dropletforced", "run_gamma", "core"]
