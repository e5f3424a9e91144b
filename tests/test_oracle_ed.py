"""Pin the oracle's OneChQS and TwoChQS chains against brute-force Fock-space
exact diagonalisation of the same Wilson-chain Hamiltonians (untruncated
steps): every rescaled level, multiplets expanded to 2S+1 states."""
import numpy as np
import pytest

import pyoracle as po
from ed_util import expand_multiplets, onechqs_chain_spectrum, twochqs_chain_spectrum


@pytest.mark.parametrize("pars", [(0.5, 0.03, -0.2, 2.5, True), (0.3, 0.05, -0.15, 3.0, False),
                                  (0.0, 0.02, 0.0, 2.5, True)])
def test_onechqs_vs_ed(pars):
    U, G, ed, L, alam = pars
    for N in range(0, 4):
        st = po.run_onechqs(U=U, Gamma=G, ed=ed, Lambda=L, nsitesmax=N + 1, ncut=100000, calcdens=2,
                            disc_scheme=0 if alam else 1)
        s = st[-1]
        e_nrg = expand_multiplets(s["en"], s["qn"], s["bbe"])
        e_ed = onechqs_chain_spectrum(U, G, ed, L, N, a_lambda=alam)
        assert len(e_nrg) == len(e_ed) == 4 ** (N + 2)
        assert np.max(np.abs(e_nrg - e_ed)) < 1e-12 * max(1.0, e_ed.max())


@pytest.mark.parametrize("pars", [(0.4, 0.03, 0.03, -0.2, 3.0), (0.5, 0.03, 0.01, -0.1, 2.5)])
def test_twochqs_vs_ed(pars):
    U, G1, G2, ed, L = pars
    for N in range(0, 2):
        st = po.run_twochqs(U=U, Gamma1=G1, Gamma2=G2, ed=ed, Lambda=L, nsitesmax=N, ncut=100000)
        s = st[-1]
        e_nrg = expand_multiplets(s["en"], s["qn"], s["bbe"])
        e_ed = twochqs_chain_spectrum(U, G1, G2, ed, L, N)
        assert len(e_nrg) == len(e_ed) == 4 * 16 ** (N + 1)
        assert np.max(np.abs(e_nrg - e_ed)) < 1e-12 * max(1.0, e_ed.max())


def test_cgordan_closed_forms():
    # <j1 m1; 1/2 +-1/2 | j m> textbook values and orthogonality
    for j1 in (0.5, 1.0, 1.5, 2.0):
        for m in np.arange(-j1 + 0.5, j1 + 0.51, 1.0):
            j = j1 + 0.5
            up = po.cgordan(j1, m - 0.5, 0.5, 0.5, j, m)
            assert abs(up - np.sqrt((j1 + m + 0.5) / (2 * j1 + 1))) < 1e-14
            dn = po.cgordan(j1, m + 0.5, 0.5, -0.5, j, m)
            assert abs(dn - np.sqrt((j1 - m + 0.5) / (2 * j1 + 1))) < 1e-14
    # sum_m1 |<j1 m1 j2 m-m1|j m>|^2 = 1
    for j1, j2, j in [(1.0, 0.5, 1.5), (1.5, 1.0, 1.5), (2.0, 1.0, 1.0)]:
        s = sum(po.cgordan(j1, m1, j2, j - m1, j, j) ** 2 for m1 in np.arange(-j1, j1 + 0.1, 1.0))
        assert abs(s - 1.0) < 1e-13
    assert po.cgordan(0.5, 0.5, 0.5, 0.5, 0.0, 1.0) == 0.0  # |M| > J
