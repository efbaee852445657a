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


def sinewave1D(height: np.ndarray, h0: float, n: int, eps: float, dims: int) -> None:
    """sinewave1D!(height, h₀, n, ϵ, dims) (src/initialvalues.jl:6-16): n sine
    waves along dimension `dims` (1-based as in the reference), amplitude
    h₀ ϵ, shifted by -h₀ ϵ / size(height, dims)."""
    ax = int(dims) - 1
    size = height.shape[ax]
    acc = np.cumsum(np.ones(height.shape), axis=ax)  # accumulate(+, ones, dims)
    hd = 2 * math.pi * n * acc / (size - 1)
    height[...] = h0 * (1.0 + eps * (np.sin(hd) - 1 / size))


# ---------------------------------------------------------------- contact-angle patterns


class VPolygon:
    """A convex polygon given by its vertices (the LazySets.VPolygon the
    pattern functions return); `p in P` includes the boundary."""

    def __init__(self, vertices):
        v = np.asarray(vertices, dtype=float)
        c = v.mean(axis=0)
        order = np.argsort(np.arctan2(v[:, 1] - c[1], v[:, 0] - c[0]))
        self.vertices = v[order]  # counter-clockwise

    def __contains__(self, p) -> bool:
        return bool(self.contains(np.asarray(p, dtype=float)[None, :])[0])

    def contains(self, pts: np.ndarray) -> np.ndarray:
        """Membership of many points (rows of pts), boundary included with a
        relative tolerance as LazySets' comparisons use."""
        v = self.vertices
        inside = np.ones(len(pts), dtype=bool)
        for k in range(len(v)):
            a, b = v[k], v[(k + 1) % len(v)]
            cross = (b[0] - a[0]) * (pts[:, 1] - a[1]) - (b[1] - a[1]) * (pts[:, 0] - a[0])
            scale = np.abs(b - a).max() * (np.abs(pts - a).max(axis=1) + 1.0)
            inside &= cross >= -1.5e-8 * scale
        return inside

    def __repr__(self):
        return f"VPolygon({self.vertices.tolist()})"


class Ellipsoid:
    """{x : (x - c)ᵀ Q⁻¹ (x - c) ≤ 1} with a diagonal shape matrix Q (LazySets.Ellipsoid)."""

    def __init__(self, center, q_diag):
        self.center = np.asarray(center, dtype=float)
        self.q = np.asarray(q_diag, dtype=float)

    def __contains__(self, p) -> bool:
        return bool(self.contains(np.asarray(p, dtype=float)[None, :])[0])

    def contains(self, pts: np.ndarray) -> np.ndarray:
        d = pts - self.center
        r = (d ** 2 / self.q).sum(axis=1)
        return r <= 1.0 + 1.5e-8

    def __repr__(self):
        return f"Ellipsoid(center={self.center.tolist()}, Q=diag{self.q.tolist()})"


def _clamp_vertices(xc, yc, shape):
    """The reference's "vertices on the grid" rule: an x outside [1, size 1]
    moves to 1, else a y outside [1, size 2] moves to 1."""
    for k in range(len(xc)):
        if xc[k] < 1 or xc[k] > shape[0]:
            xc[k] = 1
        elif yc[k] < 1 or yc[k] > shape[1]:
            yc[k] = 1


def _paint(θ, θ0, δ, P):
    """θ[j, i] = θ₀ + δₐ where the point (i, j) (1-based) is in P, else θ₀,
    as the reference's loops assign (note the swapped indices)."""
    n1, n2 = θ.shape
    i, j = np.meshgrid(np.arange(1, n1 + 1, dtype=float), np.arange(1, n2 + 1, dtype=float),
                       indexing="ij")
    inside = P.contains(np.stack([i.ravel(), j.ravel()], axis=1)).reshape(n1, n2)
    val = np.where(inside, θ0 + δ, θ0)
    # θ[j, i] = val[i, j]: a transpose (the reference fills square maps)
    θ[...] = val.T if θ.shape == val.T.shape else val.T[: θ.shape[0], : θ.shape[1]]
    return θ


def boxpattern(θ: np.ndarray, θ0, center=None, δa=1 / 36, side=20, **kw):
    """boxpattern(θ, θ₀; center, δₐ, side) (src/initialvalues.jl:423-470):
    a square patch of contact angle θ₀ + δₐ.  Returns (θ, polygon)."""
    δa = kw.pop("δₐ", δa)
    if kw:
        raise TypeError(f"unknown keywords {sorted(kw)}")
    center = (θ.shape[0] // 2, θ.shape[1] // 2) if center is None else center
    xc = [center[0] - side / 2, center[0] - side / 2, center[0] + side / 2, center[0] + side / 2]
    yc = [center[1] - side / 2, center[1] + side / 2, center[1] - side / 2, center[1] + side / 2]
    _clamp_vertices(xc, yc, θ.shape)
    P = VPolygon(list(zip(xc, yc)))
    return _paint(θ, θ0, δa, P), P


def ellipsepattern(θ: np.ndarray, θ0, center=None, δa=1 / 36, a=10, b=5, **kw):
    """ellipsepattern(θ, θ₀; center, δₐ, a, b) (src/initialvalues.jl:494-525):
    an elliptic patch with half axes a (first index) and b.  Returns (θ, ellipse)."""
    δa = kw.pop("δₐ", δa)
    if kw:
        raise TypeError(f"unknown keywords {sorted(kw)}")
    center = (θ.shape[0] // 2, θ.shape[1] // 2) if center is None else center
    P = Ellipsoid([float(center[0]), float(center[1])], [float(a) ** 2, float(b) ** 2])
    return _paint(θ, θ0, δa, P), P


def trianglepattern(θ: np.ndarray, θ0, center=None, δa=1 / 36, side=60, **kw):
    """trianglepattern(θ, θ₀; center, δₐ, side) (src/initialvalues.jl:550-581):
    an equilateral triangle of side `side` centred on `center`.  Returns (θ, polygon)."""
    δa = kw.pop("δₐ", δa)
    if kw:
        raise TypeError(f"unknown keywords {sorted(kw)}")
    center = (θ.shape[0] // 2, θ.shape[1] // 2) if center is None else center
    height = math.sqrt(3) / 2 * side
    xc = [center[0] - side / 2, center[0] + side / 2, center[0]]
    yc = [center[1] - height / 3, center[1] - height / 3, center[1] + 2 * height / 3]
    _clamp_vertices(xc, yc, θ.shape)
    P = VPolygon(list(zip(xc, yc)))
    return _paint(θ, θ0, δa, P), P
