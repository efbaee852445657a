"""Host-side initial conditions of src/initialvalues.jl used by the 2D drivers.

Random fields use the library's counter-based Philox stream (keyed by seed
and node) in place of Julia's randn!/rand!, so a GPU-generated and a
host-generated field with the same seed are identical.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from ._lib import check, lib
from .core import cospi

TAG_INIT_NORMAL = 1
TAG_INIT_UNIFORM = 2


def normals(n: int, seed: int, step: int = 0, tag: int = TAG_INIT_NORMAL):
    z0 = np.empty(n)
    z1 = np.empty(n)
    check(lib().swalbe_host_normals(int(seed), int(step), int(tag), int(n),
                                    z0.ctypes.data_as(C.POINTER(C.c_double)),
                                    z1.ctypes.data_as(C.POINTER(C.c_double))), "host_normals")
    return z0, z1


def noise_index(Lx: int, Ly: int) -> np.ndarray:
    """Stream index of every node (x, y) of an (Lx, Ly) film, Fortran order:
    2 (x + Lx floor(y/2)) + y mod 2 (vertical node pairs share a Philox call)."""
    x = np.arange(Lx)[:, None]
    y = np.arange(Ly)[None, :]
    return (2 * (x + Lx * (y // 2)) + (y % 2)).astype(np.int64)


def lattice_normals(Lx: int, Ly: int, seed: int, step: int = 0, tag: int = TAG_INIT_NORMAL):
    """The two normals of every node of an (Lx, Ly) film, shape (Lx, Ly) each."""
    idx = noise_index(Lx, Ly)
    z0, z1 = normals(int(idx.max()) + 1, seed, step, tag)
    return z0[idx], z1[idx]


def uniforms(n: int, seed: int, step: int = 0, tag: int = TAG_INIT_UNIFORM):
    u = np.empty(n)
    check(lib().swalbe_host_uniforms(int(seed), int(step), int(tag), int(n),
                                     u.ctypes.data_as(C.POINTER(C.c_double))), "host_uniforms")
    return u


def randinterface(height: np.ndarray, h0: float, eps: float, seed: int = 0) -> None:
    """randinterface!(height, h₀, ϵ) (src/initialvalues.jl:23-33):
    height .= h₀ (1 + ϵ N(0,1)), normals from the counter stream."""
    if height.ndim == 1:
        from .one_d import randinterface as r1
        return r1(height, h0, eps, seed=seed)
    z0, _ = lattice_normals(height.shape[0], height.shape[1], seed)
    height[...] = h0 * (1.0 + eps * z0)


def random_film(Lx: int, Ly: int, seed: int, base: float, amp: float, dist: str = "normal"):
    """h = base + amp * N(0,1) (or U(0,1)) per node, Fortran order."""
    if dist == "normal":
        z, _ = lattice_normals(Lx, Ly, seed)
        return base + amp * z
    z = uniforms(Lx * Ly, seed)
    return (base + amp * z).reshape((Lx, Ly), order="F")


def singledroplet(height: np.ndarray, radius, θ, center, precursor: float = 0.05):
    """singledroplet(height, radius, θ, center) (src/initialvalues.jl:203-224):
    a spherical cap h = (cos(asin(r/R)) - cos πθ) R, precursor elsewhere.
    `center` is 1-based as in the reference; the precursor (0.05 there) is a
    parameter here.  A vector height is the 1D method (hcrit clamp, scalar
    center)."""
    if height.ndim == 1:
        from .one_d import singledroplet as s1
        return s1(height, radius, θ, center, hcrit=precursor)
    lx, ly = height.shape
    i = np.arange(1, lx + 1)[:, None]
    j = np.arange(1, ly + 1)[None, :]
    circ = np.sqrt((i - center[0]) ** 2.0 + (j - center[1]) ** 2.0)
    inside = circ <= radius
    cap = (np.cos(np.arcsin(np.minimum(circ / radius, 1.0))) - cospi(θ)) * radius
    h = np.where(inside, cap, precursor)
    h = np.where(h < 0, precursor, h)
    height[...] = h
    return height


def two_droplets_2d(Lx: int, Ly: int, r1, r2, θ1, θ2, c1, c2, precursor: float = 0.05):
    """Two spherical caps on a 2D film: cap₁ + cap₂ - precursor, the 2D
    analogue of two_droplets (src/initialvalues.jl:277-323)."""
    a = singledroplet(np.zeros((Lx, Ly), order="F"), r1, θ1, c1, precursor)
    b = singledroplet(np.zeros((Lx, Ly), order="F"), r2, θ2, c2, precursor)
    h = a + b - precursor
    return np.asfortranarray(np.where(h < 0, precursor, h))


def rivulet(Lx: int, Ly: int, radius, θ, orientation, center, hmin: float = 0.05):
    """rivulet(Lx, Ly, radius, θ, orientation, center, hmin) (src/initialvalues.jl:69-118),
    without the randn() surface noise (noise = 0 default)."""
    idx = np.arange(1, (Lx if orientation == "y" else Ly) + 1)
    circ = np.abs(idx - center).astype(float)
    prof = np.where(circ <= radius,
                    (np.cos(np.arcsin(np.minimum(circ / radius, 1.0))) - cospi(θ)) * radius, hmin)
    prof = np.where(prof <= hmin, hmin, prof)
    if orientation == "y":
        return np.asfortranarray(np.repeat(prof[:, None], Ly, axis=1))
    return np.asfortranarray(np.repeat(prof[None, :], Lx, axis=0))


def torus(lx: int, ly: int, r1, R2, θ, center, hmin: float = 0.05):
    """torus(lx, ly, r₁, R₂, θ, center, hmin) (src/initialvalues.jl:151-168), noise = 0."""
    i = np.arange(1, lx + 1)[:, None]
    j = np.arange(1, ly + 1)[None, :]
    coord = np.sqrt((i - center[0]) ** 2.0 + (j - center[1]) ** 2.0)
    half = r1 ** 2 - (coord - R2) ** 2
    h = np.where(half <= 0.0, hmin, np.sqrt(np.maximum(half, 0.0)))
    corr = h - r1 * cospi(θ)
    return np.asfortranarray(np.where(corr < hmin, hmin, corr))


def sinewave(Lx: int, Ly: int, h0: float, kx: int, ky: int, eps: float):
    """The initial height of run_rayleightaylor (src/simulate.jl:350-353)."""
    i = np.arange(1, Lx + 1)[:, None]
    j = np.arange(1, Ly + 1)[None, :]
    return np.asfortranarray(
        h0 * (1 + eps * np.sin(2 * math.pi * kx * i / (Lx - 1)) * np.sin(2 * math.pi * ky * j / (Ly - 1))))
