"""Fields of a GPU-resident state that read and write through, as the
reference's state arrays do.

In the reference, `state.height` is the state's own array: the drivers and
user scripts write into it in place and the next op sees the writes
(`state.height .= 1.0`, src/simulate.jl:241; `randinterface!(state.height,
h₀, ϵ)`, :289; `state.height[i, j] = …`, :350-353; `singledroplet(state.height,
…)`, :394; `state.ftemp .= state.fout` in scripts/Rivulet_stability.jl:94-95).

Here `state.height` is a `HipField`, a view of the device array:

* reads (indexing, numpy functions, arithmetic) see the device's current
  values;
* writes (item and slice assignment, in-place arithmetic, ufuncs with
  `out=`, `np.copyto`, `fill`) reach the device before the next op on the
  state runs.

A state keeps a host mirror per field that has been read or written.  A
mirror is downloaded on first use, edited in place by writes and marked
dirty.  Every library call on the state goes through `state.handle`, which
first uploads the dirty mirrors and then marks every mirror stale (the call
may change any field), so the next read downloads again.  Whole-field
assignments (`state.height[...] = 1.0`, `state.height = h`) go to the device
directly, with no mirror.

Mirrors hold float64 for both storage types: an FP32 state's deviation
encoding round-trips exactly through it (decoded in FP64, re-encoded to the
same FP32 bits), so uploading an untouched part of a mirror changes nothing.
"""
from __future__ import annotations

import numpy as np

# ndarray methods that would change the array in place: on a HipField they
# would change a copy and be lost, so they are refused (fill is supported)
_MUTATORS = frozenset(("sort", "put", "itemset", "resize", "partition", "setfield",
                       "setflags", "byteswap"))


def _whole(idx) -> bool:
    """True for an index that selects the whole array: ..., :, or (:, :, ...)."""
    if idx is Ellipsis:
        return True
    if isinstance(idx, slice):
        return idx == slice(None)
    if isinstance(idx, tuple):
        return all(i is Ellipsis or (isinstance(i, slice) and i == slice(None)) for i in idx)
    return False


class _Entry:
    __slots__ = ("arr", "version", "dirty")

    def __init__(self, arr, version):
        self.arr = arr
        self.version = version
        self.dirty = False


class HipField:
    """`state.<name>`: a live view of one field of a GPU-resident state (see
    the module docstring).  `np.asarray(field)` and `field[...]` on the
    right of an assignment give host copies; writes through the field reach
    the device."""

    __slots__ = ("_st", "_name")

    def __init__(self, state, name: str):
        object.__setattr__(self, "_st", state)
        object.__setattr__(self, "_name", name)

    # ---- the mirror
    def _m(self) -> np.ndarray:
        return self._st._mirror(self._name)

    def _wrote(self) -> None:
        self._st._mark_dirty(self._name)

    # ---- array protocol
    @property
    def shape(self):
        return self._st._field_shape(self._name)

    @property
    def ndim(self) -> int:
        return len(self.shape)

    @property
    def size(self) -> int:
        return int(np.prod(self.shape))

    @property
    def dtype(self):
        return np.dtype(self._st.T)

    @property
    def nbytes(self) -> int:
        return self.size * self.dtype.itemsize

    def __len__(self) -> int:
        return self.shape[0]

    def __array__(self, dtype=None, copy=None):
        return np.array(self._m(), dtype=dtype or self.dtype, order="F")

    def __iter__(self):
        return iter(np.asarray(self))

    def __getitem__(self, idx):
        v = self._m()[idx]
        if isinstance(v, np.ndarray):
            return v.astype(self.dtype)
        return self.dtype.type(v)

    def __setitem__(self, idx, value):
        if isinstance(value, HipField):
            value = np.asarray(value, dtype=np.float64)
        if _whole(idx) and (np.ndim(value) == 0 or np.shape(value) == self.shape):
            self._st.set(self._name, value)  # the whole field: straight to the device
            return
        m = self._m()
        m[idx] = value
        self._wrote()

    def fill(self, value) -> None:
        self._st.set(self._name, float(value))

    def copy(self, order="F"):
        return np.array(self, order=order)

    # ---- in-place arithmetic writes through
    def _inplace(self, op, other):
        m = self._m()
        op(m, np.asarray(other, dtype=np.float64) if isinstance(other, HipField) else other,
           out=m)
        self._wrote()
        return self

    def __iadd__(self, o):
        return self._inplace(np.add, o)

    def __isub__(self, o):
        return self._inplace(np.subtract, o)

    def __imul__(self, o):
        return self._inplace(np.multiply, o)

    def __itruediv__(self, o):
        return self._inplace(np.true_divide, o)

    def __ipow__(self, o):
        return self._inplace(np.power, o)

    # ---- numpy dispatch: inputs are read as copies, outputs write through
    def __array_ufunc__(self, ufunc, method, *inputs, out=None, **kw):
        args = [np.asarray(a) if isinstance(a, HipField) else a for a in inputs]
        if out is None:
            return getattr(ufunc, method)(*args, **kw)
        outs = tuple(o._m() if isinstance(o, HipField) else o for o in out)
        res = getattr(ufunc, method)(*args, out=outs, **kw)
        for o in out:
            if isinstance(o, HipField):
                o._wrote()
        return out[0] if len(out) == 1 else out

    def __array_function__(self, func, types, args, kwargs):
        if func is np.copyto and args and isinstance(args[0], HipField):
            dst = args[0]
            src = args[1] if len(args) > 1 else kwargs.pop("src")
            if kwargs.get("where") is None and len(args) < 4:
                dst[...] = np.broadcast_to(np.asarray(src, dtype=np.float64), dst.shape)
            else:
                m = dst._m()
                np.copyto(m, np.asarray(src) if isinstance(src, HipField) else src,
                          *args[2:], **kwargs)
                dst._wrote()
            return None

        def conv(a):
            if isinstance(a, HipField):
                return np.asarray(a)
            if isinstance(a, (list, tuple)):
                return type(a)(conv(b) for b in a)
            return a
        return func(*conv(args), **{k: conv(v) for k, v in kwargs.items()})

    # ---- operators give host arrays
    def _bin(self, op, o, rev=False):
        a = np.asarray(self)
        o = np.asarray(o) if isinstance(o, HipField) else o
        return op(o, a) if rev else op(a, o)

    def __add__(self, o):
        return self._bin(np.add, o)

    def __radd__(self, o):
        return self._bin(np.add, o, True)

    def __sub__(self, o):
        return self._bin(np.subtract, o)

    def __rsub__(self, o):
        return self._bin(np.subtract, o, True)

    def __mul__(self, o):
        return self._bin(np.multiply, o)

    def __rmul__(self, o):
        return self._bin(np.multiply, o, True)

    def __truediv__(self, o):
        return self._bin(np.true_divide, o)

    def __rtruediv__(self, o):
        return self._bin(np.true_divide, o, True)

    def __pow__(self, o):
        return self._bin(np.power, o)

    def __rpow__(self, o):
        return self._bin(np.power, o, True)

    def __matmul__(self, o):
        return self._bin(np.matmul, o)

    def __neg__(self):
        return -np.asarray(self)

    def __pos__(self):
        return np.asarray(self)

    def __abs__(self):
        return np.abs(np.asarray(self))

    def __lt__(self, o):
        return self._bin(np.less, o)

    def __le__(self, o):
        return self._bin(np.less_equal, o)

    def __gt__(self, o):
        return self._bin(np.greater, o)

    def __ge__(self, o):
        return self._bin(np.greater_equal, o)

    def __eq__(self, o):
        return self._bin(np.equal, o)

    def __ne__(self, o):
        return self._bin(np.not_equal, o)

    __hash__ = None

    # ---- everything else an ndarray offers, on a copy
    def __getattr__(self, attr):
        if attr.startswith("__"):  # numpy probes __array_interface__ etc.: not an ndarray
            raise AttributeError(attr)
        if attr in _MUTATORS:
            raise TypeError(f"state.{self._name}.{attr}() would change a host copy; "
                            f"assign through state.{self._name}[...] instead")
        return getattr(np.asarray(self), attr)

    def __setattr__(self, attr, value):
        raise AttributeError(f"cannot set attribute {attr!r} of state.{self._name}")

    def __repr__(self):
        return f"HipField({self._name}, {np.asarray(self)!r})"


class LiveFields:
    """Mixin of the GPU-resident states: `state.<field>` gives a HipField,
    `state.<field> = value` uploads (or fills), and every library call made
    through `state.handle` first uploads the fields written on the host.

    The class provides `_names` (field name -> id), `_h` (the library
    handle), `T`, `_field_shape(name)`, `_download(name, dtype)` and
    `_upload(name, value)`."""

    def _init_fields(self) -> None:
        object.__setattr__(self, "_cache", {})
        object.__setattr__(self, "_version", 0)

    def _canon(self, name: str) -> str:
        fid = self._names[name]
        for k, v in self._names.items():
            if v == fid:
                return k
        return name

    def _mirror(self, name: str) -> np.ndarray:
        name = self._canon(name)
        ent = self._cache.get(name)
        if ent is None or (ent.version != self._version and not ent.dirty):
            ent = _Entry(self._download(name, np.float64), self._version)
            self._cache[name] = ent
        return ent.arr

    def _mark_dirty(self, name: str) -> None:
        self._cache[self._canon(name)].dirty = True

    def _flush(self) -> None:
        for name, ent in self._cache.items():
            if ent.dirty:
                self._upload(name, ent.arr)
                ent.dirty = False
                ent.version = self._version

    @property
    def handle(self):
        """The library handle, for a call that may read or change the state:
        uploads the fields written on the host, then marks every host mirror
        stale."""
        self._flush()
        object.__setattr__(self, "_version", self._version + 1)
        return self._h

    def get(self, name: str, dtype=None) -> np.ndarray:
        """Download a field (a host copy); FP32 states can return FP64
        (decoded exactly)."""
        self._flush()
        return self._download(self._canon(name), dtype)

    def set(self, name: str, value) -> None:
        """Upload an array or fill with a scalar (the whole field)."""
        name = self._canon(name)
        if isinstance(value, HipField):
            if value._st is self and value._name == name:
                return
            value = np.asarray(value, dtype=np.float64)
        self._flush()
        self._cache.pop(name, None)
        self._upload(name, value)

    def __getattr__(self, name):
        if name in type(self)._names:
            return HipField(self, self._canon(name))
        raise AttributeError(name)

    def __setattr__(self, name, value):
        if name in type(self)._names:
            self.set(name, value)
        else:
            object.__setattr__(self, name, value)

    # deterministic release: `with Sys(...) as state:` or state.close(), rather
    # than waiting for the garbage collector to run the finaliser
    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False
