"""Contact-angle patterns and the 1D sine wave against the reference's
tests (test/initialvalues.jl:2-8, 104-170) and its doctests."""
import math

import numpy as np
import pytest

T9, D = 1 / 9, 1 / 36


@pytest.fixture(scope="module")
def S():
    import swalbe
    return swalbe


def test_sinewave1D(S):
    # test/initialvalues.jl:2-8
    h = np.ones((10, 10))
    S.sinewave1D(h, 2.0, 1, 1, 1)
    assert abs(h.max() - 4.0) <= 0.5 and abs(h.min() - 0.0) <= 0.5


def test_ellipse_patterns(S):
    # test/initialvalues.jl:104-130 (1-based indices: [50,50] -> [49,49])
    th, P = S.ellipsepattern(np.ones((100, 100)), T9)
    assert th.shape == (100, 100) and th[49, 49] == T9 + D and th[0, 0] == T9
    th, _ = S.ellipsepattern(np.ones((100, 100)), T9, center=(20, 20))
    assert th[19, 19] == T9 + D and th[99, 99] == T9
    th, _ = S.ellipsepattern(np.ones((100, 100)), T9, δa=-1 / 36)
    assert th[49, 49] == T9 - D and th[99, 99] == T9
    th, _ = S.ellipsepattern(np.ones((100, 100)), T9, a=10, b=20)
    assert th[39, 49] == T9 + D and th[69, 49] == T9 + D and th[99, 99] == T9


def test_triangle_patterns(S):
    # test/initialvalues.jl:133-151
    lx, ly = 100, 100
    th, _ = S.trianglepattern(np.ones((100, 100)), T9)
    assert np.all(th > 0) and th[0, 0] == T9 and th[lx // 2 - 1, ly // 2 - 1] == T9 + D
    th, _ = S.trianglepattern(np.ones((100, 100)), T9, center=(lx / 3, 2 * ly / 3))
    assert th[0, 0] == T9 and th[round(lx / 3) - 1, round(2 * ly / 3) - 1] == T9
    th, _ = S.trianglepattern(np.ones((100, 100)), T9, center=(0, 2 * ly / 3))
    assert th[0, 0] == T9 and th[round(lx / 3) - 1, round(2 * ly / 3) - 1] == T9
    assert th[59, 0] == T9 + D  # vertices off the grid stick to the origin
    th, _ = S.trianglepattern(np.ones((100, 100)), T9, δa=-1 / 36)
    assert th[0, 0] == T9 and th[lx // 2 - 1, ly // 2 - 1] == T9 - D
    th, _ = S.trianglepattern(np.ones((100, 100)), T9, side=30)
    assert th[0, 0] == T9 and th[lx // 2 - 1, ly // 2 - 1] == T9 + D and th[49, 59] == T9 + D


def test_triangle_doctest_vertices(S):
    # src/initialvalues.jl:534-545: side 20 on 50 x 50
    th, P = S.trianglepattern(np.ones((50, 50)), T9, side=20)
    want = {(15.0, 19.226497308103742), (35.0, 19.226497308103742), (25.0, 36.547005383792516)}
    got = {tuple(np.round(v, 12)) for v in P.vertices}
    assert got == {tuple(np.round(w, 12)) for w in want}
    assert th[24, 24] == 0.1388888888888889


def test_box_patterns(S):
    # test/initialvalues.jl:153-170 and the doctest of :405-414
    lx, ly = 100, 100
    th, P = S.boxpattern(np.ones((100, 100)), T9)
    assert np.all(th > 0) and th[0, 0] == T9 and th[lx // 2 - 1, ly // 2 - 1] == T9 + D
    assert {tuple(v) for v in P.vertices} == {(40.0, 40.0), (60.0, 40.0), (60.0, 60.0), (40.0, 60.0)}
    th, _ = S.boxpattern(np.ones((100, 100)), T9, center=(lx / 3, 2 * ly / 3))
    assert th[0, 0] == T9 and th[round(2 * ly / 3) - 1, round(lx / 3) - 1] == T9 + D
    th, _ = S.boxpattern(np.ones((100, 100)), T9, center=(0, 2 * ly / 3))
    assert th[19, 0] == T9 and th[59, 0] == T9 + D
    th, _ = S.boxpattern(np.ones((100, 100)), T9, δa=-1 / 36)
    assert th[0, 0] == T9 and th[lx // 2 - 1, ly // 2 - 1] == T9 - D
    th, _ = S.boxpattern(np.ones((100, 100)), T9, side=30)
    assert th[0, 0] == T9 and th[lx // 2 - 1, ly // 2 - 1] == T9 + D
    assert th[lx // 2 + 15 - 1, ly // 2 + 15 - 1] == T9 + D  # the corner is on the patch


def test_pattern_is_transposed_as_the_reference_fills_it(S):
    # θ[j, i] = θ₀ + δₐ for the point (i, j): an off-diagonal box lands transposed
    th, P = S.boxpattern(np.ones((100, 100)), T9, center=(20, 70), side=10)
    assert th[69, 19] == T9 + D and th[19, 69] == T9
    assert (15, 65) in P and (65, 15) not in P
