"""Hand-worked cases for the two parts whose only other check is a second
restatement: CNRGarray::SetDegen (src/NRGarrayClass.cpp:244-268) and the
omega-grid drivers CalcSpecDM_NRG_FixedOmegas / CalcSpec_ManyOmegas
(src/SpecFuncClass.cpp:1932-2341).  Every expected value below was worked out
by hand from those lines; the library (C-ABI, host code: no GPU needed) and the
oracle must both reproduce them.

SetDegen walks ii = 0 .. n-1 and, for each jj > ii with |E_jj - E_ii| < MyEPS
(strict, dEqualPrec src/dEqual.cpp:31-34), lowers the larger of the two to the
smaller one; E_ii keeps its new value for the rest of the jj loop.  The cases
use MyEPS = 0.5 and binary-exact levels so that the comparisons are exact.

The grid drivers use Lambda = 2: D_N = 0.75 * 2^(-(N-1)/2), Lambda^1.25 = 2^1.25,
so the recorded omegas are 0.75 * 2^(1.25 - (N-1)/2 - i/(2 Nw)).
"""
import math

import numpy as np
import pytest

import pyoracle as po
import twochnrg_amd as t

EPS = 0.5
DEGEN_CASES = [
    # ascending: 0.375 snaps to 0; 0.75 is 0.75 from the (new) 0 and stays
    ([0.0, 0.375, 0.75], [0.0, 0.0, 0.75]),
    # descending: E_0 is lowered twice inside its jj loop, then E_1 follows
    ([0.75, 0.375, 0.0], [0.0, 0.0, 0.0]),
    ([0.375, 0.0, 0.75], [0.0, 0.0, 0.75]),
    # E_1 drops to E_0 = 0.375, E_0 then drops to 0, and E_1 follows it on its own pass
    ([0.375, 0.75, 0.0], [0.0, 0.0, 0.0]),
    ([1.0, 1.25, 0.75], [0.75, 0.75, 0.75]),
    # |dE| == MyEPS does not snap
    ([0.0, 0.5], [0.0, 0.5]),
    # a chain in index order snaps pairwise, not transitively
    ([0.0, 0.375, 0.75, 1.125], [0.0, 0.0, 0.75, 0.75]),
    # the same levels in the opposite order all collapse
    ([1.125, 0.75, 0.375, 0.0], [0.0, 0.0, 0.0, 0.0]),
]


@pytest.mark.parametrize("e,expect", DEGEN_CASES)
def test_setdegen_hand_worked(e, expect):
    e = np.array(e)
    assert np.array_equal(po.setdegen(e, EPS), np.array(expect))
    e2, _, _, _ = t.truncate(e, len(e), EPS)
    assert np.array_equal(e2, np.array(expect))


F = 2.0 ** 1.25


def D(N):
    return 0.75 * 2.0 ** (-(N - 1) / 2.0)


def _both_plans(nshell, windowed, Nw):
    a = t.spectrum_plan(2.0, nshell, windowed, Nw)
    b = po.grid_plan(2.0, nshell, windowed, Nw)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    return a


def test_fixed_omegas_plan_dmnrg():
    # DM-NRG, NshellMax 3: even shells 0, 2 (negative up, positive down), then odd shell 1
    sh, wb, w, par = _both_plans(3, False, 0)
    assert sh.tolist() == [0, 2, 2, 0, 1, 1]
    assert par.tolist() == [0, 0, 0, 0, 1, 1]
    assert wb.tolist() == [-F, -F, F, F, -F, F]
    expect = [-0.75 * 2 ** 1.75, -0.75 * 2 ** 0.75, 0.75 * 2 ** 0.75, 0.75 * 2 ** 1.75, -0.75 * 2 ** 1.25,
              0.75 * 2 ** 1.25]
    np.testing.assert_allclose(w, expect, rtol=1e-15, atol=0)


def test_fixed_omegas_plan_windowed():
    # CFS/FDM start two shells later: even 2, 4 and odd 3, 5 of NshellMax 6
    sh, wb, w, par = _both_plans(6, True, 0)
    assert sh.tolist() == [2, 4, 4, 2, 3, 5, 5, 3]
    assert par.tolist() == [0, 0, 0, 0, 1, 1, 1, 1]
    np.testing.assert_allclose(w, [-F * D(2), -F * D(4), F * D(4), F * D(2), -F * D(3), -F * D(5), F * D(5),
                                   F * D(3)], rtol=1e-15, atol=0)


def test_many_omegas_plan():
    # Nw = 2, NshellMax 3: negative N = 0, 1, 2 with i = 0, 1; positive N = 2, 1, 0 with i = 1, 0.
    # recorded w = -+2^(-i/4) 2^1.25 D_N; the DM-NRG shell sum is evaluated at -+2^1.25 2^(+i/4)
    sh, wb, w, par = _both_plans(3, False, 2)
    q = 2.0 ** 0.25
    assert sh.tolist() == [0, 0, 1, 1, 2, 2, 2, 2, 1, 1, 0, 0]
    assert par.tolist() == [-1] * 12
    np.testing.assert_allclose(wb, [-F, -F * q, -F, -F * q, -F, -F * q, F * q, F, F * q, F, F * q, F], rtol=1e-15)
    np.testing.assert_allclose(w, [-F * D(0), -F * D(0) / q, -F * D(1), -F * D(1) / q, -F * D(2), -F * D(2) / q,
                                   F * D(2) / q, F * D(2), F * D(1) / q, F * D(1), F * D(0) / q, F * D(0)],
                               rtol=1e-15)
    # windowed: shells 2 .. NshellMax-1 only
    sh, wb, w, par = _both_plans(4, True, 1)
    assert sh.tolist() == [2, 3, 3, 2]
    np.testing.assert_allclose(w, [-F * D(2), -F * D(3), F * D(3), F * D(2)], rtol=1e-15)


def _both_combines(w, par, v, Nw):
    a = t.spectrum_combine(w, par, v, Nw)
    b = po.grid_combine(w, par, v, Nw)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    return a


def test_fixed_omegas_interpolation_hand_worked():
    """Even values a, b, c, d at -F D0, -F D2, F D2, F D0 and odd values p, q at
    -+F D1.  On the sorted union the even set is read between its knots with
    the weights 2 - sqrt(2) (at -F D1) and sqrt(2) - 1 (at F D1), the odd set
    with (1 -+ 1/sqrt(2))/2 (at -+F D2); the outermost points lie outside the
    odd set and take the even value alone."""
    _, _, w, par = _both_plans(3, False, 0)
    a, b, c, d, p, q = 1.0, 2.0, 3.0, 4.0, 5.0, 7.0
    v = np.array([a, b, c, d, p, q])
    wo, ro = _both_combines(w, par, v, 0)
    r2 = math.sqrt(2.0)
    np.testing.assert_allclose(wo, [-F * D(0), -F * D(1), -F * D(2), F * D(2), F * D(1), F * D(0)], rtol=1e-15)
    expect = [a,
              0.5 * (a + (2 - r2) * (b - a) + p),
              0.5 * (b + p + 0.5 * (1 - 1 / r2) * (q - p)),
              0.5 * (c + p + 0.5 * (1 + 1 / r2) * (q - p)),
              0.5 * (c + (r2 - 1) * (d - c) + q),
              d]
    np.testing.assert_allclose(ro, expect, rtol=1e-14, atol=0)


@pytest.mark.parametrize("nshell,windowed", [(9, False), (12, True)])
def test_fixed_omegas_reproduce_linear_values(nshell, windowed):
    """Values exactly linear in omega on both shell sets are reproduced by the
    interpolate-and-average at every point of the grid."""
    _, _, w, par = _both_plans(nshell, windowed, 0)
    alpha, beta = 0.3, -1.7
    wo, ro = _both_combines(w, par, alpha + beta * w, 0)
    np.testing.assert_allclose(ro, alpha + beta * wo, rtol=1e-13, atol=1e-15)
    # a linear function on the even shells and another on the odd ones average
    v = np.where(par == 0, alpha + beta * w, -alpha + 2.0 * beta * w)
    wo, ro = _both_combines(w, par, v, 0)
    inner = (wo >= w[par == 1].min()) & (wo <= w[par == 1].max()) & (wo >= w[par == 0].min()) & (
        wo <= w[par == 0].max())
    np.testing.assert_allclose(ro[inner], 0.5 * ((alpha + beta * wo) + (-alpha + 2 * beta * wo))[inner], rtol=1e-13,
                               atol=1e-15)


def test_many_omegas_record_values_as_they_are():
    _, _, w, par = _both_plans(5, False, 3)
    v = np.arange(len(w), dtype=float) * 0.25
    wo, ro = _both_combines(w, par, v, 3)
    assert np.array_equal(wo, w) and np.array_equal(ro, v)


def test_plan_arguments():
    with pytest.raises(t.NrgError):
        t.spectrum_plan(1.0, 3, False, 0)
    with pytest.raises(t.NrgError):
        t.spectrum_plan(2.0, 0, False, 0)
