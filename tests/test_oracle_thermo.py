"""Impurity thermodynamics with the chain subtraction (CalcThermo
src/NRGcodehandlerClass.cpp:627-652, CNRGthermo::AddValue / SaveNValue
src/NRGthermoClass.cpp:10-151) on the oracle:

- with the impurity decoupled (Gamma = 0) and the chain untruncated, the
  subtracted values are the atomic impurity's, in closed form, at every
  (N, betabar_nc) row;
- for the single-channel Anderson model of the Ncut=400 workload the three
  regimes appear: free orbital (T chi = 1/8, S = ln 4) at T >> U, the local
  moment (T chi ~ 1/4, S ~ ln 2) and Kondo screening (both -> 0) below T_K.
"""
import numpy as np
import pytest

import pyoracle as po


@pytest.mark.parametrize("U,ed", [(0.5, -0.25), (0.3, -0.05)])
def test_decoupled_impurity_closed_form(U, ed):
    r = po.onechqs_thermo(U=U, Gamma=0.0, ed=ed, Lambda=2.5, nsitesmax=4, ncut=100000)
    assert r.shape == (16, 7)
    T = r[:, 0]
    # atomic levels: empty 0, singly occupied ed (twice, Sz = +-1/2), double 2 ed + U
    w1 = np.exp(-ed / T)
    w2 = np.exp(-(2 * ed + U) / T)
    Z = 1 + 2 * w1 + w2
    tchi = 0.5 * w1 / Z
    S = np.log(Z) + (2 * ed * w1 + (2 * ed + U) * w2) / (T * Z)
    np.testing.assert_allclose(r[:, 3], tchi, rtol=1e-11, atol=1e-14)
    np.testing.assert_allclose(r[:, 6], S, rtol=1e-11, atol=1e-14)
    # the columns are Tot, Chain, Tot - Chain
    np.testing.assert_allclose(r[:, 3], r[:, 1] - r[:, 2], rtol=0, atol=1e-15)
    np.testing.assert_allclose(r[:, 6], r[:, 4] - r[:, 5], rtol=0, atol=1e-15)


def test_kondo_regimes():
    r = po.onechqs_thermo(U=0.5, Gamma=0.02, ed=-0.25, Lambda=2.5, nsitesmax=40, ncut=400)
    assert r.shape == (160, 7)
    T, tchi, S = r[:, 0], r[:, 3], r[:, 6]
    assert np.all(np.diff(T) < 0)
    # free orbital
    assert abs(tchi[0] - 0.125) < 0.01 and abs(S[0] - np.log(4.0)) < 0.01
    # local moment
    lm = (T > 1e-4) & (T < 1e-2)
    assert np.all((tchi[lm] > 0.18) & (tchi[lm] < 0.25))
    assert np.all((S[lm] > np.log(2.0) - 0.03) & (S[lm] < 0.8))   # ln 2 with the betabar_nc oscillation
    # Kondo screening
    assert tchi[-1] < 0.01 and S[-1] < 0.15
    low = T < 1e-2
    assert np.all(np.diff(S[low][::8]) < 0)   # one betabar_nc, even N: no even/odd oscillation
