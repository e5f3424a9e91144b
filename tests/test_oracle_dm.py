"""DM-NRG oracle (oracle/nrg_oracle.c orc_dm_*) pinned by brute-force Fock-space
ED on untruncated chains, by invariants of the backward sweep, and by an
independent recomputation of the fixed-omega grid interpolation.

- rho_N: the reduced density matrix of DM_NRG_CalcRhoN_withSU2
  (src/DM_NRG/DM_NRG_CalcRhoN.cpp:252-449) equals <i, m=S| Tr_site rho_{N+1} |j, m=S>
  in the H_N eigenbasis; rho_Nmax = e^{-betabar E}/Z_N with Z_N from the shifted
  temperature Temp/D_{Nmax-1} (src/DM_NRG/DM_NRG.cpp:203-214).
- DMNRG_SpecDens_M_STL (src/SpecFunc_allSTL.cpp:223-692) and CalcSpecCosti_T_N
  (src/SpecFuncClass.cpp:417-662) equal the Lehmann sums of c_{d,up} over all
  spin components with the same broadening.
Compared gauge-invariantly: eigenvector phases and rotations inside degenerate
levels are arbitrary on both sides.
"""
import numpy as np
import pytest

import pyoracle as po
from ed_util import onechqs_chain_fock, partial_trace_last_orbital

LAM = 2.5
PARS = [dict(U=0.5, Gamma=0.05, ed=-0.25), dict(U=0.6, Gamma=0.08, ed=-0.15)]


def _sectors(nmode):
    dim = 1 << nmode
    ne = np.array([bin(s).count("1") for s in range(dim)])
    up = sum(1 << m for m in range(0, nmode, 2))
    nup = np.array([bin(s & up).count("1") for s in range(dim)])
    return ne, nup - (ne - nup)   # n_e, 2 Sz


def _ed_chain(p, nmax, betabar):
    """rho_N (full Fock space of d + sites 0..N) for N = nmax .. 0 with the DM-NRG
    conventions, and the H_N, annihilators per shell."""
    HLF = 0.5 * (1 + 1 / LAM)

    def dn(N):
        return HLF * LAM ** (-(N - 1) / 2.0)
    H, C = {}, {}
    for N in range(nmax + 1):
        H[N], C[N] = onechqs_chain_fock(p["U"], p["Gamma"], p["ed"], LAM, N, a_lambda=True)
    temp = dn(nmax) / betabar
    w, V = np.linalg.eigh(H[nmax])
    ZN = np.sum(np.exp(-w * dn(nmax - 1) / temp))          # Z at Temp/D_{Nmax-1}
    rho = {nmax: (V * (np.exp(-betabar * w) / ZN)) @ V.T}  # TempBar of rho = 1/betabar
    for N in range(nmax - 1, -1, -1):
        rho[N] = partial_trace_last_orbital(rho[N + 1], 2 * (N + 3))
    return H, C, rho, temp


def _hw_basis(Hn, C, N, Q, S):
    """Highest-weight states (S+ |.> = 0) of sector (Q, Sz=S) in the H_N eigenbasis."""
    nm = 2 * (N + 2)
    ne, sz2 = _sectors(nm)
    sel = np.where((ne - (N + 2) == int(round(Q))) & (sz2 == int(round(2 * S))))[0]
    Sp = sum(C[2 * o].T @ C[2 * o + 1] for o in range(N + 2))
    up = np.where((ne - (N + 2) == int(round(Q))) & (sz2 == int(round(2 * S)) + 2))[0]
    M = Sp[np.ix_(up, sel)] if len(up) else np.zeros((0, len(sel)))
    if M.shape[0]:
        u, s, vt = np.linalg.svd(M)
        null = vt[np.sum(s > 1e-9):].T
    else:
        null = np.eye(len(sel))
    Hs = null.T @ Hn[np.ix_(sel, sel)] @ null
    e, W = np.linalg.eigh(Hs)
    return sel, null @ W, e


def _clusters(e, tol=1e-8):
    out, cur = [], [0]
    for i in range(1, len(e)):
        if e[i] - e[i - 1] <= tol:
            cur.append(i)
        else:
            out.append(cur)
            cur = [i]
    out.append(cur)
    return out


@pytest.mark.parametrize("p", PARS)
def test_rho_sweep_matches_partial_trace(p):
    nsites, mtemp, betabar = 4, 2, 1.3
    d = po.DmRun(U=p["U"], Gamma=p["Gamma"], ed=p["ed"], Lambda=LAM, nsitesmax=nsites, ncut=100000,
                 Mtemp=mtemp, betabar=betabar)
    assert d.nshell_max == mtemp
    H, C, rho, temp = _ed_chain(p, mtemp, betabar)
    assert abs(d.temp - temp) <= 1e-14 * temp
    for N in range(mtemp, -1, -1):
        sh = d.shell(N)
        for b in range(len(sh["rho"])):
            Q, S = sh["qn"][b]
            sel, V, e = _hw_basis(H[N], C[N], N, Q, S)
            en = sh["en"][sh["bbe"][b, 0]:sh["bbe"][b, 1] + 1]
            assert len(e) == len(en)
            np.testing.assert_allclose(e, en, atol=1e-10)
            red = V.T @ rho[N][np.ix_(sel, sel)] @ V
            R = sh["rho"][b]
            for c1 in _clusters(en):
                for c2 in _clusters(en):
                    a = np.linalg.norm(R[np.ix_(c1, c2)])
                    x = np.linalg.norm(red[np.ix_(c1, c2)])
                    assert abs(a - x) <= 1e-10 * max(1.0, x), (N, b, c1, c2, a, x)
        # the printed trace: sum (2S+1) rho_ii equals the full trace of the ED rho
        assert abs(sh["trace"] - np.trace(rho[N])) <= 1e-12 * abs(np.trace(rho[N]))


def _lehmann(Hn, c, rho, w, temp, dn, twindow=2.0, bt=0.8, b=0.5 * np.log(LAM)):
    e, V = np.linalg.eigh(Hn)
    cm = V.T @ c @ V
    r = V.T @ rho @ V
    dE = e[None, :] - e[:, None]          # E_b - E_a  (rows a, cols b)
    tdn = temp / dn

    def bd(x):
        x = np.where(x == 0.0, 1e-100, x)
        lg = np.exp(-b * b / 4) / (b * np.abs(x) * np.sqrt(np.pi)) * np.exp(-(np.log(np.abs(w / x)) / b) ** 2)
        lg = np.where(w * x < 0, 0.0, lg)
        lor = (1 / (2 * np.pi)) * bt / ((w - x) ** 2 + bt ** 2)
        return np.where(np.abs(x) < twindow * tdn, lor, lg)
    D = bd(dE)
    # term 1: sum_{a,a',b} rho_{a'a} <a|c|b> <a'|c|b> BD(E_b - E_a')
    t1 = np.sum((r.T @ cm) * cm * D)
    # term 2: sum_{a,b,b'} rho_{bb'} <a|c|b'> <a|c|b> BD(E_b - E_a)
    t2 = np.sum((cm @ r) * cm * D)
    return t1 + t2


@pytest.mark.parametrize("p", PARS)
def test_shell_spectral_sums_match_lehmann(p):
    nsites, mtemp, betabar = 4, 2, 1.3
    d = po.DmRun(U=p["U"], Gamma=p["Gamma"], ed=p["ed"], Lambda=LAM, nsitesmax=nsites, ncut=100000,
                 Mtemp=mtemp, betabar=betabar)
    H, C, rho, temp = _ed_chain(p, mtemp, betabar)
    HLF = 0.5 * (1 + 1 / LAM)
    for N in range(0, mtemp):          # DMNRG form (shell below the last)
        dn = HLF * LAM ** (-(N - 1) / 2.0)
        for w in (-3.1, -0.7, -0.05, 0.02, 0.4, 1.9):
            ref = _lehmann(H[N], C[N][0], rho[N], w, temp, dn)
            got = d.spec_shell(N, w)
            assert abs(got - ref) <= 1e-10 * max(1e-6, abs(ref)), (N, w, got, ref)
    # last shell: Costi form with the input betabar
    N = mtemp
    dn = HLF * LAM ** (-(N - 1) / 2.0)
    e = np.linalg.eigvalsh(H[N])
    Z = np.sum(np.exp(-betabar * e))
    ev, V = np.linalg.eigh(H[N])
    costi_rho = (V * np.exp(-betabar * ev)) @ V.T / Z
    for w in (-1.3, -0.2, 0.3, 2.2):
        # Costi: sum |<a|c|b>|^2 (e^{-bE_a} + e^{-bE_b}) BD / Z == Lehmann with the thermal rho
        ref = _lehmann(H[N], C[N][0], costi_rho, w, temp, dn)
        got = d.spec_shell(N, w)
        assert abs(got - ref) <= 1e-10 * max(1e-6, abs(ref)), (w, got, ref)


def test_trace_conserved_with_truncation():
    # truncated chain: sum_i (2S+1) rho_ii is passed down the sweep unchanged when
    # the rotation, the child map and the CG weights are right
    d = po.DmRun(U=0.5, Gamma=0.02, ed=-0.25, Lambda=LAM, nsitesmax=12, ncut=120, Mtemp=10, betabar=1.0)
    tr = [d.shell(N)["trace"] for N in range(d.nshell_max + 1)]
    assert np.ptp(tr) <= 1e-12 * abs(tr[0])


def test_fixed_omega_grid_interpolation():
    # CalcSpecDM_NRG_FixedOmegas: even/odd shell sets, each linearly interpolated
    # on the sorted union and averaged (outside one set's range: the other set)
    d = po.DmRun(U=0.5, Gamma=0.02, ed=-0.2, Lambda=LAM, nsitesmax=11, ncut=150, Mtemp=9, betabar=1.1)
    w, r = d.spectrum()
    HLF = 0.5 * (1 + 1 / LAM)
    fac = LAM ** 1.25
    sets = []
    for par in (0, 1):
        Ns = list(range(par, d.spec_nshell_max, 2))
        xs = [-fac * HLF * LAM ** (-(N - 1) / 2.0) for N in Ns] + \
             [fac * HLF * LAM ** (-(N - 1) / 2.0) for N in Ns[::-1]]
        ys = [d.spec_shell(N, -fac) / (HLF * LAM ** (-(N - 1) / 2.0)) for N in Ns] + \
             [d.spec_shell(N, fac) / (HLF * LAM ** (-(N - 1) / 2.0)) for N in Ns[::-1]]
        sets.append((np.array(xs), np.array(ys)))
    allw = np.sort(np.concatenate([s[0] for s in sets]))
    np.testing.assert_allclose(w, allw, rtol=1e-15)
    exp = []
    for x in allw:
        v = []
        for xs, ys in sets:
            v.append(np.interp(x, xs, ys) if xs[0] <= x <= xs[-1] else None)
        if v[0] is None:
            v[0] = v[1]
        if v[1] is None:
            v[1] = v[0]
        exp.append(0.5 * (v[0] + v[1]))
    np.testing.assert_allclose(r, exp, rtol=1e-12, atol=1e-300)
