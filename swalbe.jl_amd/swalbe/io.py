"""Snapshots and restarts.

- `snapshot(snap, field, t; dumping)` is the reference's host-side snapshot!
  (src/measures.jl:99-105).  The fused loop takes snapshots itself:
  `time_loop(..., snapshots=d, snapshot_every=k)` fills d["h_<t>"] from an
  on-device ring whose device-to-host copies overlap the following steps.
- `save_snapshots` / `restart_from_height` store and read heights keyed
  "h_<t>" as the reference does with JLD2/BSON (src/initialvalues.jl:358-382).
  JLD2 and BSON need Julia, so the container format here is NumPy's .npz
  (each entry vec(h), column-major, as the reference stores them).
- `save_state` / `load_state` keep a state's stored representation (FP32
  deviations included), so a run restarted from them continues bit for bit.
"""
from __future__ import annotations

import ctypes as C
import re
from typing import Dict, Optional, Sequence

import numpy as np

from . import _lib
from ._lib import FIELDS, check, lib
from .core import HipState

DIST = ("fout", "ftemp", "feq")
# what the fused loop reads (h, u, f_t when τ ≠ 1) plus the usual outputs
RESTART_FIELDS = ("height", "velx", "vely", "pressure", "fout", "ftemp")


def snapshot(snap: np.ndarray, field, t: int, dumping: int = 1000) -> None:
    """snapshot!(snap, field, t; dumping): row t ÷ dumping (1-based in Julia,
    so row t // dumping - 1 here) becomes vec(field) when t % dumping == 0."""
    if t % dumping == 0:
        a = field.get("height") if isinstance(field, HipState) else np.asarray(field)
        snap[t // dumping - 1, :] = a.ravel(order="F")


def save_snapshots(path: str, snaps: Dict[str, np.ndarray]) -> None:
    """Heights keyed "h_<t>", each stored as vec(h)."""
    np.savez(path, **{k: np.asarray(v).ravel(order="F") for k, v in snaps.items()})


def _steps(keys: Sequence[str]):
    out = []
    for k in keys:
        m = re.fullmatch(r"h_(\d+)", k)
        if m:
            out.append((int(m.group(1)), k))
    return sorted(out)


def restart_from_height(data: str, kind: str = "npz", timestep: int = 0,
                        size=(512, 512)) -> np.ndarray:
    """restart_from_height(data; kind, timestep, size) (src/initialvalues.jl:358-382):
    the height "h_<timestep>" of a snapshot file, or the last one when
    timestep == 0, reshaped to `size`."""
    if kind != "npz":
        raise ValueError("snapshot files are .npz here (JLD2/BSON need Julia)")
    with np.load(data, allow_pickle=False) as z:
        steps = _steps(z.files)
        if not steps:
            raise KeyError(f"{data}: no h_<t> entries")
        key = steps[-1][1] if timestep == 0 else f"h_{timestep}"
        return np.asarray(z[key], dtype=np.float64).reshape(size, order="F").copy()


def save_state(state: HipState, path: str, fields: Sequence[str] = RESTART_FIELDS) -> None:
    """The stored representation of `fields`, the element type, the FP32
    reference height and the step counter, in one .npz."""
    L = lib()
    out = {}
    esz = 8 if state.T is np.float64 else 4
    dt = np.float64 if esz == 8 else np.float32
    for name in fields:
        shape = (state.Lx, state.Ly, 9) if name in DIST else (state.Lx, state.Ly)
        a = np.empty(shape, dtype=dt, order="F")
        check(L.swalbe_download_raw(state.handle, FIELDS[name], a.ctypes.data, a.nbytes),
              f"download_raw {name}")
        out["raw_" + name] = a.ravel(order="F")
    href, g = C.c_double(), C.c_double()
    check(L.swalbe_get_reference_height(state.handle, C.byref(href), C.byref(g)), "href")
    meta = np.array([state.Lx, state.Ly, esz, state.t], dtype=np.int64)
    np.savez(path, meta=meta, enc=np.array([href.value, g.value]),
             kind=np.array([state.kind == "thermal"], dtype=np.int8), **out)


def load_state(path: str) -> HipState:
    """A new state holding exactly what save_state stored."""
    L = lib()
    with np.load(path, allow_pickle=False) as z:
        Lx, Ly, esz, t = (int(v) for v in z["meta"])
        T = np.float64 if esz == 8 else np.float32
        st = HipState(Lx, Ly, T=T, kind="thermal" if int(z["kind"][0]) else "simple")
        href, g = (float(v) for v in z["enc"])
        if T is np.float32:
            st.set_reference_height(href, g)
        for key in z.files:
            if not key.startswith("raw_"):
                continue
            name = key[4:]
            a = np.ascontiguousarray(z[key])
            check(L.swalbe_upload_raw(st.handle, FIELDS[name], a.ctypes.data, a.nbytes),
                  f"upload_raw {name}")
    object.__setattr__(st, "t", t)
    return st


__all__ = ["snapshot", "save_snapshots", "restart_from_height", "save_state", "load_state",
           "RESTART_FIELDS"]
