"""The oracle's DM-NRG / FDM post-processing of the TwoChQSz chain (the
reference never hands a TwoChQSz run to DM_NRG, so this part restates its
OneChQS path on (Q, Sz) blocks: degeneracy 1, no Clebsch-Gordan factors, d_up
on every state of a shell, FDM site dimension 16).

Pinned by a Fock-space construction of the same two-channel chain (impurity +
both channels' sites 0..N, the legacy TwoChQS(z) rescaling): the operator d_up
in the shell eigenbasis (the impurity table of TwoChQSz_SetH0Anderson and the
impurity-only update with the new site's fermion sign) through the thermal
Lehmann sum of the last shell, and the backward sweep rho_0 = Tr_site rho_1 by
the Fock-space partial trace, both untruncated.  With truncation: the trace of
rho_N is conserved down the sweep and the FDM shell weights sum to 1."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import pyoracle as po  # noqa: E402
import ed_util as ed  # noqa: E402

L = 2.5
HLF = 0.5 * (1 + 1 / L)
CASES = [dict(U=0.5, G1=0.03, G2=0.01, ed=-0.2), dict(U=0.3, G1=0.05, G2=0.05, ed=-0.15)]


def _fock(p, N):
    """Dense rescaled H_N (minus its minimum) and the annihilators: orbital 0 = d,
    orbital 1 + 2n + ch = site n of channel ch (mode 2 orbital + spin)."""
    v = [np.sqrt(g * (2 / np.pi) / HLF ** 2 / L) for g in (p["G1"], p["G2"])]

    def chi(n):
        return (1 - L ** -float(n)) / np.sqrt((1 - L ** -(2.0 * n - 1)) * (1 - L ** -(2.0 * n + 1)))
    E = ed.FockED(1 + 2 * (N + 1))
    s = L ** (N / 2.0) / (HLF * np.sqrt(L))
    E.eps(0, p["ed"] * s)
    E.U(0, p["U"] * s)
    E.const = 0.5 * p["U"] * s
    for ch in range(2):
        E.hop(0, 1 + ch, v[ch] * L ** (N / 2.0))
        for n in range(1, N + 1):
            E.hop(1 + 2 * (n - 1) + ch, 1 + 2 * n + ch, chi(n) * L ** ((N - n) / 2.0))
    H, C = ed.fock_hamiltonian(E)
    return H - np.linalg.eigvalsh(H).min() * np.eye(len(H)), C


def _loggauss(w, E, b):
    if E == 0.0:
        E = 1e-100
    if w * E < 0:
        return 0.0
    a1 = np.log(abs(w / E)) / b
    return np.exp(-(a1 - b / 4) ** 2) / (b * abs(E) * np.sqrt(np.pi))


def _gauss(w, E, b):
    return np.exp(-((w - E) / b) ** 2) / (b * np.sqrt(np.pi))


@pytest.mark.parametrize("p", CASES)
def test_last_shell_is_the_lehmann_sum(p):
    """FDM on an untruncated 2-shell chain: the window sum is the Costi form of the
    last shell, the thermal Lehmann sum of d_up with LogGaussDelta / GaussDelta."""
    bb = 0.9
    D = po.DmRun(U=p["U"], Gamma=p["G1"], ed=p["ed"], Lambda=L, nsitesmax=1, ncut=10 ** 6, Mtemp=0, betabar=bb,
                 method=2, model=1, Gamma2=p["G2"], bbroad=0.6)
    assert D.nsh == 2
    N = D.spec_nshell_max - 1
    H, C = _fock(p, N)
    ev, V = np.linalg.eigh(H)
    c = V.T @ C[0] @ V
    x = np.exp(-bb * ev)
    Z = x.sum()
    tdn = D.temp / (HLF * L ** (-(N - 1) / 2.0))
    for w in (-2.1, -0.6, -0.05, 0.03, 0.4, 1.7):
        ref = 0.0
        for a in range(len(ev)):
            for b in np.nonzero(np.abs(c[a]) > 1e-14)[0]:
                dE = ev[b] - ev[a]
                bd = _gauss(w, dE, 0.8) if abs(dE) < 2.0 * tdn else _loggauss(w, dE, 0.6)
                ref += c[a, b] ** 2 * (x[a] + x[b]) / Z * bd
        got = D.spec_shell(N, w)
        assert abs(got - ref) <= 1e-12 * max(abs(ref), 1e-6), (w, got, ref)


def _trace_out_site(rho, nmode):
    """Trace out the highest 4 modes (site N of both channels)."""
    r = ed.partial_trace_last_orbital(rho, nmode)
    return ed.partial_trace_last_orbital(r, nmode - 2)


@pytest.mark.parametrize("p", CASES)
def test_sweep_is_the_partial_trace(p):
    """rho_0 = Tr_site rho_1 with rho_1 = exp(-betabar H_1)/Z (FDM, untruncated):
    block by block against the Fock-space partial trace, compared through
    Frobenius norms between degenerate clusters of H_0 (gauge invariant)."""
    bb = 0.9
    D = po.DmRun(U=p["U"], Gamma=p["G1"], ed=p["ed"], Lambda=L, nsitesmax=1, ncut=10 ** 6, Mtemp=0, betabar=bb,
                 method=2, model=1, Gamma2=p["G2"], bbroad=0.6)
    H1, _ = _fock(p, 1)
    w, V = np.linalg.eigh(H1)
    rho1 = (V * np.exp(-bb * w)) @ V.T / np.sum(np.exp(-bb * w))
    rho0 = _trace_out_site(rho1, 10)
    H0, _ = _fock(p, 0)
    nm = 6
    ne = np.array([bin(s).count("1") for s in range(1 << nm)])
    sz2 = np.array([sum(((s >> (2 * o)) & 1) - ((s >> (2 * o + 1)) & 1) for o in range(nm // 2)) for s in range(1 << nm)])
    sh = D.shell(0)
    assert abs(sh["trace"] - 1.0) <= 1e-12 and abs(np.trace(rho0) - 1.0) <= 1e-12
    for b in range(len(sh["rho"])):
        Q, Sz = sh["qn"][b]
        sel = np.where((ne - 3 == int(round(Q))) & (sz2 == int(round(2 * Sz))))[0]
        e, U = np.linalg.eigh(H0[np.ix_(sel, sel)])
        en = sh["en"][sh["bbe"][b, 0]:sh["bbe"][b, 1] + 1]
        np.testing.assert_allclose(e, en, atol=1e-10)
        red = U.T @ rho0[np.ix_(sel, sel)] @ U
        R = sh["rho"][b]
        cl, cur = [], [0]
        for i in range(1, len(en)):
            if abs(en[i] - en[cur[0]]) < 1e-8:
                cur.append(i)
            else:
                cl.append(cur)
                cur = [i]
        cl.append(cur)
        for c1 in cl:
            for c2 in cl:
                a, x = np.linalg.norm(R[np.ix_(c1, c2)]), np.linalg.norm(red[np.ix_(c1, c2)])
                assert abs(a - x) <= 1e-11, (b, c1, c2, a, x)


def test_truncated_run_invariants():
    """A TwoChQSz FDM run with truncation from N = 2: the printed traces
    sum rho_ii are conserved down the sweep, the FDM weights sum to 1, and the
    window sums are finite and non-negative."""
    D = po.DmRun(U=0.5, Gamma=0.03, ed=-0.25, Lambda=2.5, nsitesmax=6, ncut=400, Mtemp=4, betabar=0.779,
                 method=2, model=1, Gamma2=0.01, bbroad=0.6)
    kept = [D.shell(N)["kept"] for N in range(D.nshell_max + 1)]
    assert any(not k.all() for k in kept)
    wn = np.array([D.shell(N)["wn"] for N in range(D.nshell_max + 1)])
    assert abs(wn.sum() - 1.0) <= 1e-12
    for w in (1e-3, -1e-3, 0.05, -0.3, 0.8):
        a = D.spec_omega(w)
        assert np.isfinite(a) and a >= -1e-12, (w, a)


def test_fdm_site_dimension():
    """The FDM environment count d enters w_N only as Z_N^disc d^{Nmax-N}: the
    unnormalised weights w_N Z_tot of d = 16 and d = 4 differ by 4^{Nmax-N} on
    the same chain, and both sets of w_N sum to 1."""
    c = dict(U=0.5, Gamma=0.03, ed=-0.25, Lambda=2.5, nsitesmax=4, ncut=200, Mtemp=3, betabar=0.779,
             method=2, model=1, Gamma2=0.01, bbroad=0.6)
    D16 = po.DmRun(**c)
    D4 = po.DmRun(**dict(c, fdm_site_dim=4.0))
    NM = D16.nshell_max
    w16 = np.array([D16.shell(N)["wn"] for N in range(NM + 1)]) * D16.ztot
    w4 = np.array([D4.shell(N)["wn"] for N in range(NM + 1)]) * D4.ztot
    assert w4[0] == 0.0 == w16[0] and (w4[1:] > 0).all()   # N = 0 keeps every state
    np.testing.assert_allclose(w16[1:] / w4[1:], 4.0 ** (NM - np.arange(1, NM + 1)), rtol=1e-12)
    for D in (D4, D16):
        assert abs(sum(D.shell(N)["wn"] for N in range(NM + 1)) - 1.0) <= 1e-12


@pytest.mark.parametrize("site_dim", [16.0, 4.0])
def test_exact_fdm_sum_rule(site_dim):
    """The full density matrix (every FDM quirk of the reference switched off,
    fdm_fix = 63) on a truncated TwoChQSz chain with Mtemp < Nmax-1: CalcNorm is
    the sum rule <{d, d^dag}> = 1 and the density matrix has unit trace
    (sum over shells of the discarded diagonal plus rho_Nmax), whatever the
    environment count d.  Each quirk that touches the weights, switched back
    on alone, breaks the sum rule: rho_Nmax's exponent at D_Mtemp, the Costi
    last shell, the skipped all-discarded block pairs and CalcNorm's double
    count of the discarded-discarded terms."""
    c = dict(U=0.5, Gamma=0.03, ed=-0.25, Lambda=2.5, nsitesmax=5, ncut=120, Mtemp=2, betabar=0.779,
             method=2, model=1, Gamma2=0.01, bbroad=0.6, fdm_site_dim=site_dim)
    D = po.DmRun(**c, fdm_fix=po.FDM_EXACT)
    NM = D.nshell_max
    sh = [D.shell(N) for N in range(NM + 1)]
    assert any(not s["kept"].all() for s in sh[:NM])
    assert abs(D.norm() - 1.0) <= 1e-10, D.norm()
    tr = sh[NM]["trace"]
    for N in range(NM):
        dsc = sh[N]["kept"] == 0
        tr += sum(np.diag(r)[dsc[b0:b1 + 1]].sum() for r, (b0, b1) in zip(sh[N]["rho"], sh[N]["bbe"]))
    assert abs(tr - 1.0) <= 1e-12, tr
    for bit in (po.FDM_FIX_RHOMAX, po.FDM_FIX_LAST, po.FDM_FIX_PAIRS, po.FDM_FIX_NORM):
        n = po.DmRun(**c, fdm_fix=po.FDM_EXACT & ~bit).norm()
        assert abs(n - 1.0) > 1e-4, (bit, n)
    # the reference's own FDM (no bit) is not normalised on this chain
    assert abs(po.DmRun(**c).norm() - 1.0) > 1e-2


def _lehmann_dmnrg(Hn, c, rho, w, temp, dn, twindow=2.0, bt=0.8, b=0.5 * np.log(L)):
    """DMNRG_SpecDens_M_STL's shell sum in an orthonormal basis:
    sum_{a,a',b} rho_{a'a} c_ab c_a'b BD(E_b - E_a') + sum_{a,b,b'} rho_bb' c_ab' c_ab BD(E_b - E_a),
    BD = BroadDelta, the Lorentzian BDeltaTemp for |dE| < twindow T / D_N."""
    e, V = np.linalg.eigh(Hn)
    cm = V.T @ c @ V
    r = V.T @ rho @ V
    dE = e[None, :] - e[:, None]
    x = np.where(dE == 0.0, 1e-100, dE)
    lg = np.exp(-b * b / 4) / (b * np.abs(x) * np.sqrt(np.pi)) * np.exp(-(np.log(np.abs(w / x)) / b) ** 2)
    lg = np.where(w * x < 0, 0.0, lg)
    lor = (1 / (2 * np.pi)) * bt / ((w - x) ** 2 + bt ** 2)
    D = np.where(np.abs(x) < twindow * temp / dn, lor, lg)
    return np.sum((r.T @ cm) * cm * D) + np.sum((cm @ r) * cm * D)


@pytest.mark.parametrize("p", CASES)
def test_dmnrg_shell_sum_is_the_reduced_lehmann_sum(p):
    """Plain DM-NRG (UseCFS 0) on the untruncated 2-shell TwoChQSz chain at a
    finite temperature: with Mtemp >= Nsitesmax-1 the reference takes
    T = D_{Nsitesmax+199} / betabar (Nsitesmax = 2 shells here), so betabar is
    chosen to give D_1 / T = 0.9.
    rho_1 = exp(-E D_1/T) / sum exp(-E D_0/T) (DM_NRG's two scales), rho_0 its
    Fock-space partial trace over site 1, and shell 0's sum equals the
    DMNRG_SpecDens_M_STL expression in the Fock space of the first shell with
    c = d_up, BroadDelta and the Lorentzian window."""
    def dn(N):
        return HLF * L ** (-(N - 1) / 2.0)
    target = 0.9
    bb = target * dn(2 + 199) / dn(1)
    D = po.DmRun(U=p["U"], Gamma=p["G1"], ed=p["ed"], Lambda=L, nsitesmax=1, ncut=10 ** 6, Mtemp=1, betabar=bb,
                 method=0, model=1, Gamma2=p["G2"])
    assert D.nsh == 2 and D.nshell_max == 1
    T = D.temp
    assert abs(dn(1) / T - target) <= 1e-12 * target
    H1, _ = _fock(p, 1)
    w1, V1 = np.linalg.eigh(H1)
    rho1 = (V1 * np.exp(-w1 * dn(1) / T)) @ V1.T / np.sum(np.exp(-w1 * dn(0) / T))
    rho0 = _trace_out_site(rho1, 10)
    H0, C0 = _fock(p, 0)
    assert abs(D.shell(0)["trace"] - np.trace(rho0)) <= 1e-12 * abs(np.trace(rho0))
    for w in (-2.3, -0.6, -0.04, 0.02, 0.35, 1.8):
        ref = _lehmann_dmnrg(H0, C0[0], rho0, w, T, dn(0))
        got = D.spec_shell(0, w)
        assert abs(got - ref) <= 1e-10 * max(1e-6, abs(ref)), (w, got, ref)


def test_dmnrg_truncated_trace_conserved():
    """Plain DM-NRG on a truncated TwoChQSz chain: sum rho_ii is passed down the
    sweep unchanged (the child map and the (Q, Sz) weights), and the spectrum
    is finite."""
    D = po.DmRun(U=0.5, Gamma=0.03, ed=-0.25, Lambda=2.5, nsitesmax=6, ncut=300, Mtemp=5, betabar=0.9,
                 method=0, model=1, Gamma2=0.01)
    kept = [D.shell(N)["kept"] for N in range(D.nshell_max + 1)]
    assert any(not k.all() for k in kept)
    tr = [D.shell(N)["trace"] for N in range(D.nshell_max + 1)]
    assert np.ptp(tr) <= 1e-12 * abs(tr[0])
    w, r = D.spectrum()
    assert np.all(np.isfinite(r)) and len(w) > 0
