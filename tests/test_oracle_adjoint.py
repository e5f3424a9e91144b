"""The oracle's derivative of the fixed-truncation NRG_main loop (forward
mode, tangents along (U, Gamma, ed)) against central finite differences of
the same oracle with every step's Kept set frozen at the base point, and
against the particle-hole-symmetry invariant <n_d> = 1 on the line
ed = -U/2.  The reference has no derivative; its own "gradient" is the
finite difference of its loop, which the frozen-Kept oracle runs restate."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import pyoracle as po  # noqa: E402

KEYS = ("U", "Gamma", "ed")

CASES = [
    # truncated chain off particle-hole symmetry (cuts from N = 3 on)
    dict(U=0.5, Gamma=0.02, ed=-0.2, Lambda=2.5, nsitesmax=9, ncut=120),
    # truncated chain at particle-hole symmetry
    dict(U=0.5, Gamma=0.03, ed=-0.25, Lambda=2.5, nsitesmax=8, ncut=100),
    # untruncated, the A_Lambda-off convention (NRG_main -D 1)
    dict(U=0.3, Gamma=0.05, ed=-0.1, Lambda=3.0, nsitesmax=4, ncut=10000, disc_scheme=1),
]


def _fd(kw, kept, key, h):
    kp, km = dict(kw), dict(kw)
    kp[key] += h
    km[key] -= h
    sp, _ = po.run_onechqs_jvp(kept=kept, **kp)
    sm, _ = po.run_onechqs_jvp(kept=kept, **km)
    return sp, sm


def _cluster_sums(st, v, tol=1e-7):
    out = []
    for b0, b1 in st["bbe"]:
        e, x, i = st["en"][b0:b1 + 1], v[b0:b1 + 1], 0
        while i < len(e):
            j = i
            while j + 1 < len(e) and abs(e[j + 1] - e[i]) < tol:
                j += 1
            out.append(x[i:j + 1].sum())
            i = j + 1
    return np.array(out)


@pytest.mark.parametrize("case", range(len(CASES)))
def test_tangent_matches_central_fd(case):
    kw = dict(CASES[case], calcdens=2, betabar=0.727)
    base = po.run_onechqs(**kw)
    kept = po.kept_arrays(base)
    st, tn = po.run_onechqs_jvp(kept=kept, dpar=np.eye(3), **kw)
    assert any(not s["kept"].all() for s in st) == (kw["ncut"] < 1000)
    for a, b in zip(base, st):   # the frozen run is the plain run
        assert np.array_equal(a["en"], b["en"]) and np.array_equal(a["opavg"], b["opavg"])
    for t, key in enumerate(KEYS):
        h = 1e-4 * abs(kw[key])
        # Richardson: D = (4 D(h/2) - D(h)) / 3 removes the h^2 term of the central difference
        runs = [_fd(kw, kept, key, hh) for hh in (h, h / 2)]
        for fld, tf in (("suscep", "dsuscep"), ("entropy", "dentropy"), ("opavg", "dopavg"), ("en", "den")):
            fds, ans = [], []
            for q, s in enumerate(st):
                if fld in ("suscep", "entropy") and s["N"] < 0:
                    continue
                d1 = (runs[0][0][q][fld] - runs[0][1][q][fld]) / (2 * h)
                d2 = (runs[1][0][q][fld] - runs[1][1][q][fld]) / h
                fd, an = (4 * d2 - d1) / 3, tn[q][tf][t]
                if fld == "en":
                    # single levels of a degenerate cluster move by the eigenvalues of G
                    # restricted to it: compare sums over degenerate clusters
                    fd, an = _cluster_sums(s, fd), _cluster_sums(s, an)
                fds.append(fd)
                ans.append(an)
            fd, an = np.concatenate(fds), np.concatenate(ans)
            # derivatives that vanish by symmetry (d/ded of T chi at ed = -U/2) are
            # compared in absolute terms
            err = np.max(np.abs(fd - an)) / max(np.max(np.abs(fd)), 1e-3)
            assert err <= 1e-6, (case, key, fld, err)


def test_particle_hole_line_keeps_nd():
    """On ed = -U/2, <n_d> = 1 for every U and Gamma: the directional derivative
    along (dU, dGamma, ded) = (1, 0, -1/2) and the Gamma derivative vanish."""
    kw = dict(U=0.5, Gamma=0.03, ed=-0.25, Lambda=2.5, nsitesmax=8, ncut=100, calcdens=2)
    base = po.run_onechqs(**kw)
    st, tn = po.run_onechqs_jvp(kept=po.kept_arrays(base), dpar=[[1, 0, -0.5], [0, 1, 0], [0, 0, 1]], **kw)
    dmax = max(abs(t["dopavg"][2][0]) for t in tn)
    assert dmax > 0.05   # <n_d> does move off the line
    for s, t in zip(st, tn):
        assert abs(s["opavg"][0] - 1.0) < 1e-12
        assert abs(t["dopavg"][0][0]) < 1e-9 and abs(t["dopavg"][1][0]) < 1e-9, (s["N"], t["dopavg"])


def test_h_minus1_closed_form():
    """At N = -1 the tangent of H_{-1} is (U/2, ed + U/2, 2 ed + 3U/2)' / (Lambda HLF)
    (src/Main/Main_SetH0.cpp:1847-1849) and f, n_d are constants."""
    L = 2.5
    hlf = 0.5 * (1 + 1 / L)
    st, tn = po.run_onechqs_jvp(U=0.5, Gamma=0.02, ed=-0.2, Lambda=L, nsitesmax=1, ncut=100,
                                dpar=np.eye(3))
    den = tn[0]["den"]
    assert np.allclose(den[0], np.array([0.5, 0.5, 1.5]) / (L * hlf), rtol=0, atol=1e-15)
    assert np.allclose(den[1], 0.0, atol=0)
    assert np.allclose(den[2], np.array([0, 1, 2]) / (L * hlf), rtol=0, atol=1e-15)


def _dm_fd(kw, kept, key, h, shells, ws):
    kp, km = dict(kw), dict(kw)
    kp[key] += h
    km[key] -= h
    dp, dm = po.DmRun(kept=kept, **kp), po.DmRun(kept=kept, **km)
    return {(N, w): (dp.spec_shell(N, w) - dm.spec_shell(N, w)) / (2 * h) for N in shells for w in ws}


@pytest.mark.parametrize("ed", [-0.2, -0.25])
def test_dmnrg_shell_sums_match_central_fd(ed):
    """d/dtheta of the DM-NRG shell sums (DMNRG_SpecDens_M_STL, Costi form at the
    last shell) against Richardson central differences of frozen-Kept oracle
    runs, every shell, 7 omegas per shell on both sides of the BroadDelta /
    LorentzDelta switch."""
    kw = dict(U=0.5, Gamma=0.03, ed=ed, Lambda=2.5, nsitesmax=11, ncut=150, Mtemp=8, betabar=0.9)
    base = po.run_onechqs(U=kw["U"], Gamma=kw["Gamma"], ed=ed, Lambda=2.5, nsitesmax=11, ncut=150, calcdens=1)
    assert any(not s["kept"].all() for s in base)
    kept = po.kept_arrays(base)
    D = po.DmRun(kept=kept, dpar=np.eye(3), **kw)
    shells = list(range(D.spec_nshell_max))
    ws = [0.3, -0.7, 1.5, -2.2, 3.0, 0.05, -0.02]
    an = {(N, w): D.spec_shell_jvp(N, w) for N in shells for w in ws}
    for t, key in enumerate(KEYS):
        # levels 4e-4 apart at the particle-hole point make larger steps nonlinear
        h = 1e-5 * abs(kw[key])
        d1 = _dm_fd(kw, kept, key, h, shells, ws)
        d2 = _dm_fd(kw, kept, key, h / 2, shells, ws)
        fd = {k: (4 * d2[k] - d1[k]) / 3 for k in d1}
        scale = max(abs(v) for v in fd.values())
        err = max(abs(fd[k] - an[k][t]) for k in fd)
        assert err <= 1e-8 * scale, (ed, key, err, scale)
    if ed == -0.25:
        # particle-hole symmetry keeps A(w) = A(-w) along (1, 0, -1/2) and Gamma
        Dp = po.DmRun(kept=kept, dpar=[[1, 0, -0.5], [0, 1, 0]], **kw)
        for N in shells:
            for w in (0.3, 1.5, 0.05):
                a, b = Dp.spec_shell_jvp(N, w), Dp.spec_shell_jvp(N, -w)
                assert np.allclose(a, b, rtol=1e-9, atol=1e-9 * max(1.0, np.max(np.abs(a)))), (N, w, a, b)
