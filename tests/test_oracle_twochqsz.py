"""Pin the oracle's TwoChQSz chain (orc_run_twochqsz: the standalone driver
src/TwoChQSz/TwoChQSz.cpp:103-425, (Q, Sz) blocks, no SU(2)) by

- brute-force Fock-space ED of the same two-channel Wilson chain at N = 0, 1
  (untruncated): every rescaled level;
- the SU(2) restatement of the same Hamiltonian (orc_run_twochqs, itself pinned
  by ED): the (Q, Sz, E) multiset equals the TwoChQS multiplets expanded to
  Sz = -S..S, and <Sz^2> equals sum (2S+1) S(S+1)/3 weights (T chi) and the
  entropies agree, on untruncated steps;
- the (Q, Sz) quantum numbers of every state at N = 0 counted in the Fock space.
"""
import numpy as np
import pytest

import pyoracle as po
from ed_util import twochqs_chain_spectrum

PARS = [(0.4, 0.03, 0.03, -0.2, 3.0), (0.5, 0.03, 0.01, -0.25, 2.5), (0.0, 0.02, 0.05, 0.0, 2.5)]


@pytest.mark.parametrize("pars", PARS)
def test_twochqsz_vs_ed(pars):
    U, G1, G2, ed, L = pars
    for N in range(0, 2):
        st = po.run_twochqsz(U=U, Gamma1=G1, Gamma2=G2, ed=ed, Lambda=L, nsitesmax=N, ncut=100000)
        s = st[-1]
        e = np.sort(s["en"])
        e_ed = twochqs_chain_spectrum(U, G1, G2, ed, L, N)
        assert len(e) == len(e_ed) == 4 * 16 ** (N + 1)
        assert np.max(np.abs(e - e_ed)) < 1e-12 * max(1.0, e_ed.max())


def _labels_qsz(s):
    out = []
    for b in range(len(s["qn"])):
        Q, Sz = s["qn"][b]
        for i in range(s["bbe"][b, 0], s["bbe"][b, 1] + 1):
            out.append((int(round(Q)), int(round(2 * Sz)), s["en"][i]))
    return sorted(out)


def _labels_qs(s):
    out = []
    for b in range(len(s["qn"])):
        Q, S = s["qn"][b]
        for i in range(s["bbe"][b, 0], s["bbe"][b, 1] + 1):
            for m2 in range(-int(round(2 * S)), int(round(2 * S)) + 1, 2):
                out.append((int(round(Q)), m2, s["en"][i]))
    return sorted(out)


@pytest.mark.parametrize("pars", PARS[:2])
def test_twochqsz_matches_su2_chain(pars):
    U, G1, G2, ed, L = pars
    sz = po.run_twochqsz(U=U, Gamma1=G1, Gamma2=G2, ed=ed, Lambda=L, nsitesmax=1, ncut=100000)
    qs = po.run_twochqs(U=U, Gamma1=G1, Gamma2=G2, ed=ed, Lambda=L, nsitesmax=1, ncut=100000)
    assert len(sz) == len(qs) == 2
    for a, b in zip(sz, qs):
        la, lb = _labels_qsz(a), _labels_qs(b)
        assert len(la) == len(lb)
        # same (Q, 2Sz) multiset level by level
        ka = sorted((q, m, round(e, 8)) for q, m, e in la)
        kb = sorted((q, m, round(e, 8)) for q, m, e in lb)
        assert [(q, m) for q, m, _ in ka] == [(q, m) for q, m, _ in kb]
        ea = np.array([e for _, _, e in sorted(la, key=lambda x: (x[0], x[1], x[2]))])
        eb = np.array([e for _, _, e in sorted(lb, key=lambda x: (x[0], x[1], x[2]))])
        np.testing.assert_allclose(ea, eb, atol=1e-11)
        # T chi: <Sz^2> = sum (2S+1) S(S+1)/3 e^{-bE}/Z, and the entropy
        assert abs(a["suscep"][0] - b["suscep"][0]) <= 1e-12 * abs(b["suscep"][0])
        assert abs(a["entropy"][0] - b["entropy"][0]) <= 1e-12 * abs(b["entropy"][0])


def _reversed_pairs(q):
    """Cut block pairs (i, j) that f connects (Q_j = Q_i + 1, Sz_j = Sz_i +- 1/2)
    with j before i: the reference's update visits only jbl > ibl
    (src/TwoChQSz/TwoChQSz_UpdateMatrices.cpp:212), so f between them is never
    stored, in either order, and the next H_N misses those couplings."""
    return [(i, j) for i in range(len(q)) for j in range(i)
            if abs(q[j][0] - q[i][0] - 1) < 1e-9 and abs(abs(q[j][1] - q[i][1]) - 0.5) < 1e-9]


def _sz_asymmetry(s):
    spec = {}
    for b in range(len(s["qn"])):
        Q, Sz = s["qn"][b]
        spec[(int(round(Q)), int(round(2 * Sz)))] = np.sort(s["en"][s["bbe"][b, 0]:s["bbe"][b, 1] + 1])
    return max(np.max(np.abs(e - spec[(q, -m)])) for (q, m), e in spec.items())


@pytest.mark.parametrize("G2", [0.0, 0.01])
def test_twochqsz_truncated_run(G2):
    # truncated runs to N = 4 (states, not multiplets, are counted).  Without a
    # reversed block pair the (Q, Sz) and (Q, -Sz) spectra coincide, as the
    # Hamiltonian is spin symmetric; the step after a reversed pair loses that
    # symmetry, which is the reference's rule at work, not rounding
    st = po.run_twochqsz(U=0.5, Gamma1=0.03, Gamma2=G2, ed=-0.25, Lambda=2.5, nsitesmax=4, ncut=300)
    assert len(st) == 5
    broken = False
    for n, s in enumerate(st):
        assert s["en"].min() == 0.0
        assert int(s["kept"].sum()) >= min(300, len(s["en"]))
        for b in range(len(s["qn"])):
            e = s["en"][s["bbe"][b, 0]:s["bbe"][b, 1] + 1]
            assert np.all(np.diff(e) >= -1e-12)
        asym = _sz_asymmetry(s)
        if not broken:
            assert asym < 1e-10, (n, asym)
        else:
            assert asym > 1e-6, (n, asym)
        broken = broken or len(_reversed_pairs(s["cut_qn"])) > 0
    assert broken == (G2 > 0.0)
