"""Pins the CPU FP64 oracle against the reference's known-answer tests."""
import numpy as np
import pytest

import kat
import oracle as O


class OracleBackend:
    grad = staticmethod(O.grad)
    laplacian = staticmethod(O.laplacian)
    hgradp = staticmethod(O.hgradp)
    slippage = staticmethod(O.slippage)
    inclination = staticmethod(O.inclination)
    equilibrium = staticmethod(O.equilibrium)
    bgk_and_stream = staticmethod(O.bgk_and_stream)
    moments = staticmethod(O.moments)
    power_broad = staticmethod(O.power_broad)
    cospi = staticmethod(O.cospi)

    @staticmethod
    def filmpressure(h, gamma, theta, n, m, hmin, hcrit, raw=False):
        return O.filmpressure(h, gamma, theta, n, m, hmin, hcrit, raw=raw)

    @staticmethod
    def thermal(h, kbt, mu, delta, seed, step):
        return O.thermal(h, kbt, mu, delta, seed, step)

    @staticmethod
    def run_flat(Lx, Ly, Tmax, tdump):
        st = O.State(Lx, Ly)
        O.time_loop(st, O.params(tdump=tdump), Tmax)
        return st.height


CASES = [c for c in kat.load_cases() if c["op"] != "singledroplet"]


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_oracle_kat(case):
    kat.run_case(OracleBackend, case)


def test_raw_filmpressure_domain_error():
    # src/pressure.jl:101-107: exponents other than (9,3), (3,2) throw DomainError
    with pytest.raises(ValueError):
        O.filmpressure(np.ones((5, 5)), 0.01, 1 / 9, 4, 2, 0.1, 0.05, raw=True)
    # the State form accepts any (n, m) through power_broad (src/pressure.jl:141-147)
    O.filmpressure(np.ones((5, 5)), 0.01, 1 / 9, 4, 2, 0.1, 0.05)
