"""CPU tests of the write-through state fields (swalbe/fields.py) against a
simulated device: a state class whose 'device' is a dict of host arrays and
which counts uploads.  The reference's in-place statements on a state
(state.height .= 1.0, state.height[i, j] = v, randinterface!(state.height,
…), singledroplet(state.height, …), state.ftemp .= state.fout, F .+= …)
must reach the device before the next op, and reads must see the device's
current values."""
import math

import numpy as np
import pytest

from swalbe.fields import HipField, LiveFields


class FakeState(LiveFields):
    _names = {"height": 0, "fout": 1, "ftemp": 2, "Fx": 3, "hgradpx": 4, "h∇px": 4}

    def __init__(self, Lx, Ly, T=np.float64):
        self._init_fields()
        object.__setattr__(self, "Lx", Lx)
        object.__setattr__(self, "Ly", Ly)
        object.__setattr__(self, "T", T)
        object.__setattr__(self, "_h", "handle")
        object.__setattr__(self, "dev", {})
        object.__setattr__(self, "uploads", 0)
        for n in ("height", "fout", "ftemp", "Fx", "hgradpx"):
            self.dev[n] = np.zeros(self._field_shape(n), dtype=T, order="F")

    def _field_shape(self, name):
        return (self.Lx, self.Ly, 9) if name in ("fout", "ftemp") else (self.Lx, self.Ly)

    def _download(self, name, dtype=None):
        return np.array(self.dev[name], dtype=dtype or self.T, order="F")

    def _upload(self, name, value):
        object.__setattr__(self, "uploads", self.uploads + 1)
        self.dev[name][...] = value

    def op(self):
        """A library call: goes through the handle, then changes the height."""
        assert self.handle == "handle"
        self.dev["height"] += 1.0


def test_whole_field_writes_go_straight_to_the_device():
    st = FakeState(5, 4)
    st.height[...] = 1.0
    assert np.all(st.dev["height"] == 1.0) and st.uploads == 1
    st.height[:, :] = np.arange(20.0).reshape(5, 4)
    assert st.dev["height"][4, 3] == 19.0
    st.height = 2.0
    assert np.all(st.dev["height"] == 2.0)
    st.height.fill(3.0)
    assert np.all(st.dev["height"] == 3.0)
    assert not st._cache  # no mirror was needed


def test_node_writes_reach_the_device_before_the_next_op():
    st = FakeState(6, 5)
    ref = np.zeros((6, 5))
    for i in range(1, 7):
        for j in range(1, 6):
            v = 1 + 0.1 * math.sin(i) * math.cos(j)
            st.height[i - 1, j - 1] = v
            ref[i - 1, j - 1] = v
    assert st.uploads == 0                      # edits wait in the mirror ...
    assert np.array_equal(np.asarray(st.height), ref)
    st.op()                                     # ... and are uploaded before the op
    assert st.uploads == 1
    assert np.array_equal(st.dev["height"], ref + 1.0)
    assert np.array_equal(np.asarray(st.height), ref + 1.0)  # reads see the op's result


def test_a_kept_view_stays_live_across_ops():
    st = FakeState(3, 3)
    h = st.height
    assert np.all(np.asarray(h) == 0.0)
    st.op()
    assert np.all(np.asarray(h) == 1.0)
    assert h.max() == 1.0 and h.shape == (3, 3) and h.ndim == 2


def test_in_place_arithmetic_ufuncs_and_copyto_write_through():
    st = FakeState(4, 3)
    st.height = 1.0
    st.height += 0.5
    st.height *= 2.0
    np.add(st.height, 1.0, out=st.height)
    st.op()
    assert np.all(st.dev["height"] == 5.0)
    np.copyto(st.Fx, np.full((4, 3), 0.25))
    assert np.all(st.dev["Fx"] == 0.25)
    st.Fx += st.height
    st.op()
    assert np.all(st.dev["Fx"] == 5.25)


def test_field_to_field_assignment_and_aliases():
    st = FakeState(3, 2)
    st.dev["fout"][...] = np.arange(54.0).reshape(3, 2, 9)
    st.ftemp[...] = st.fout                    # state.ftemp .= state.fout
    assert np.array_equal(st.dev["ftemp"], st.dev["fout"])
    getattr(st, "h∇px")[1, 1] = 7.0            # an alias of hgradpx
    st.op()
    assert st.dev["hgradpx"][1, 1] == 7.0


def test_functions_that_write_into_their_argument():
    from swalbe.initialvalues import singledroplet
    st = FakeState(40, 30)
    singledroplet(st.height, 10, 1 / 6, (20, 15))
    want = singledroplet(np.ones((40, 30), order="F"), 10, 1 / 6, (20, 15))
    assert np.array_equal(st.dev["height"], want)


def test_reads_are_copies_and_in_place_methods_are_refused():
    st = FakeState(3, 3)
    st.height = 1.0
    row = st.height[0]
    row[:] = 9.0                               # a copy, as Julia's slicing gives
    st.op()
    assert np.all(st.dev["height"] == 2.0)
    with pytest.raises(TypeError):
        st.height.sort()
    a = np.asarray(st.height)
    assert a.dtype == np.float64 and a.flags.f_contiguous
    assert isinstance(st.height, HipField)
    assert np.array_equal(st.height + 1, a + 1) and np.array_equal(2 * st.height, 2 * a)


def test_fp32_state_mirrors_keep_fp64():
    st = FakeState(2, 2, T=np.float32)
    st.height[0, 0] = 1.0 + 2 ** -30            # below FP32 resolution of the value
    assert st.height.dtype == np.float32
    assert st._mirror("height")[0, 0] == 1.0 + 2 ** -30
    st.op()
    assert st.dev["height"].dtype == np.float32
