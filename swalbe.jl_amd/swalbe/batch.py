"""Batches of independent FP32 films on one GPU (swalbe_batch_* of the C ABI).

Each film keeps its own parameters, size and step budget; every
Batch.step(n) advances all resident films in one launch per time step.
add/remove/step only enqueue work; get() waits for the film's own last step.
A film's result is bit-identical to the same film stepped alone by
time_loop on an FP32 HipState (both run the same tile kernel).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib
from ._lib import FIELDS, FilmInit, check, lib
from .core import SysConst, Taumucs

INIT_FLAT, INIT_NORMAL, INIT_CAP, INIT_TWO_CAPS = 0, 1, 2, 3


def film_init(kind=INIT_FLAT, *, thermal=False, seed=0, base=1.0, amp=0.0, precursor=0.05,
              href=0.0, cap1=None, cap2=None, body=None, clamp=False) -> FilmInit:
    """cap = (radius, θ in units of π, 1-based centre x, centre y);
    body = (αx, αy, fac) for the inclination! force; clamp raises heights
    below the precursor to it (the reference only replaces negative ones)."""
    fi = FilmInit(kind=kind, thermal=int(bool(thermal)), seed=int(seed), base=base, amp=amp,
                  precursor=precursor, href=href, clamp=int(bool(clamp)))
    if cap1 is not None:
        fi.r1, fi.theta1, fi.c1x, fi.c1y = cap1
    if cap2 is not None:
        fi.r2, fi.theta2, fi.c2x, fi.c2y = cap2
    if body is not None:
        fi.body = 1
        fi.body_ax, fi.body_ay, fi.body_fac = body
    return fi


def device_memory():
    """(free, total) bytes of device memory."""
    f, t = C.c_size_t(), C.c_size_t()
    check(lib().swalbe_device_memory(C.byref(f), C.byref(t)), "device_memory")
    return f.value, t.value


def tune_shape(param, Lx: int, Ly: int, init) -> dict:
    """The autotuner's choice for a film shape on the current device: tile
    height (8 or 16 rows) and the milliseconds per batched step of ~4 M nodes
    of copies at each height (measured once, then cached)."""
    prm = param.cparams()
    ty, m8, m16 = C.c_int32(), C.c_double(), C.c_double()
    check(lib().swalbe_batch_tune_shape(C.byref(prm), int(Lx), int(Ly), C.byref(init),
                                        C.byref(ty), C.byref(m8), C.byref(m16)), "tune_shape")
    return {"tile_y": ty.value, "ms_rows8": m8.value, "ms_rows16": m16.value}


def film_bytes(Lx: int, Ly: int, tau: float = 1.0) -> int:
    out = C.c_size_t()
    check(lib().swalbe_batch_film_bytes(int(Lx), int(Ly), float(tau), C.byref(out)), "film_bytes")
    return out.value


class Batch:
    """A device arena of `capacity_bytes` holding up to `max_films` films."""

    def __init__(self, capacity_bytes: int, max_films: int = 4096, graph_steps: int = 0,
                 final_outputs_only: bool = False, autotune: bool = False):
        h = C.c_void_p()
        check(lib().swalbe_batch_create(_lib.F32, int(capacity_bytes), int(max_films), C.byref(h)),
              "swalbe_batch_create")
        self._h = h
        self.dims = {}
        if final_outputs_only:
            self.set_final_outputs_only(True)
        if autotune:
            self.set_autotune(True)
        if graph_steps:
            self.set_graph(graph_steps)

    def set_final_outputs_only(self, on: bool = True) -> None:
        """Films added afterwards (τ = 1) write distributions and pressure on
        their last step only; heights and velocities every step."""
        check(lib().swalbe_batch_set_final_outputs_only(self._h, int(bool(on))),
              "set_final_outputs_only")

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().swalbe_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, param: Taumucs, Lx: int, Ly: int, init: FilmInit, nsteps: int) -> int:
        slot = C.c_int32()
        prm = param.cparams()
        check(lib().swalbe_batch_add(self._h, C.byref(prm), int(Lx), int(Ly), C.byref(init),
                                       int(nsteps), C.byref(slot)), "swalbe_batch_add")
        self.dims[slot.value] = (int(Lx), int(Ly))
        return slot.value

    def remove(self, slot: int) -> None:
        check(lib().swalbe_batch_remove(self._h, int(slot)), "swalbe_batch_remove")
        self.dims.pop(slot, None)

    def step(self, nsteps: int) -> None:
        check(lib().swalbe_batch_step(self._h, int(nsteps)), "swalbe_batch_step")

    def step_timed(self, nsteps: int) -> float:
        ms = C.c_double()
        check(lib().swalbe_batch_step_timed(self._h, int(nsteps), C.byref(ms)), "step_timed")
        return ms.value

    def set_graph(self, steps: int) -> None:
        check(lib().swalbe_batch_set_graph(self._h, int(steps)), "set_graph")

    def set_autotune(self, on: bool = True) -> None:
        """Films added afterwards take the tile height (8 or 16 rows) measured
        fastest for their shape; the heights give the same bits."""
        check(lib().swalbe_batch_set_autotune(self._h, int(bool(on))), "set_autotune")

    def set_tile_y(self, rows: int) -> None:
        """Films added afterwards use tiles of `rows` rows (8 or 16); 0
        restores the default (16, or the tuned height with autotune on)."""
        check(lib().swalbe_batch_set_tile_y(self._h, int(rows)), "set_tile_y")

    def tile_y(self, slot: int) -> int:
        t = C.c_int32()
        check(lib().swalbe_batch_film_tile_y(self._h, int(slot), C.byref(t)), "film_tile_y")
        return t.value

    def set_nontemporal(self, mode: int) -> None:
        """Stores of the films (the same bits in every mode): 0 cached,
        1 every output streaming, 2 distributions and pressure streaming
        (the default: 3 % faster than cached on the 256-film benchmark)."""
        check(lib().swalbe_batch_set_nontemporal(self._h, int(mode)), "set_nontemporal")

    def wait(self, steps_behind: int = 0) -> None:
        """Block until at most `steps_behind` enqueued steps are pending."""
        check(lib().swalbe_batch_wait(self._h, int(steps_behind)), "wait")

    def sync(self) -> None:
        check(lib().swalbe_batch_sync(self._h), "sync")

    def status(self, slot: int):
        done = C.c_int64()
        fin = C.c_int32()
        check(lib().swalbe_batch_film_status(self._h, int(slot), C.byref(done), C.byref(fin)),
              "status")
        return done.value, bool(fin.value)

    def get(self, slot: int, field: str = "height", dtype=np.float64) -> np.ndarray:
        Lx, Ly = self.dims[slot]
        shape = (Lx, Ly, 9) if field == "fout" else (Lx, Ly)
        out = np.empty(shape, dtype=dtype, order="F")
        check(lib().swalbe_batch_download(self._h, int(slot), FIELDS[field], out.ctypes.data,
                                            out.nbytes), "batch download")
        return out

    def stats(self):
        nf, nn = C.c_int32(), C.c_int64()
        bu, bf = C.c_size_t(), C.c_size_t()
        check(lib().swalbe_batch_stats(self._h, C.byref(nf), C.byref(nn), C.byref(bu),
                                         C.byref(bf)), "stats")
        return {"films": nf.value, "nodes": nn.value, "bytes_used": bu.value, "bytes_free": bf.value}


__all__ = ["Batch", "FilmInit", "film_init", "film_bytes", "device_memory", "INIT_FLAT", "INIT_NORMAL",
           "INIT_CAP", "INIT_TWO_CAPS", "SysConst"]
