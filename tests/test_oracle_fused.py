"""The multi-core CPU baseline (oracle/swalbe_cpu_fused.c) equals the oracle's
time loop bit for bit: it is the same step in three OpenMP sweeps."""
import numpy as np
import pytest

import oracle as O


@pytest.mark.parametrize("tau,thermal,shape,n", [(1.0, True, (37, 21), 7), (0.8, False, (16, 40), 6),
                                                (1.0, False, (64, 64), 4)])
def test_fused_cpu_loop_equals_oracle(tau, thermal, shape, n):
    Lx, Ly = shape
    h0 = O.init_normal(Lx, Ly, 5, 1.0, 0.05)
    prm = O.params(tau=tau, kbt=1e-5, tdump=10 ** 6)
    out = []
    for fused in (False, True):
        st = O.State(Lx, Ly)
        st.height[...] = h0
        st.feq[...], _ = O.equilibrium(st.height, st.velx, st.vely, 0.0)
        st.ftemp[...] = st.feq
        if fused:
            O.fused_time_loop(st, prm, n, thermal=thermal, seed=9)
        else:
            O.time_loop(st, prm, n, thermal=thermal, seed=9)
        out.append((st.height.copy(), st.velx.copy(), st.vely.copy(), st.fout.copy()))
    for a, b in zip(*out):
        assert np.array_equal(a, b)
    assert O.fused_threads() >= 1
