"""FDM and CFS spectral sums of the oracle (orc_dm_* with method 2 / 1 on a chain
run with UpdateBefCut 1) pinned by an independent Fock-space construction of the
same truncated NRG.

The ED side builds, shell by shell, the NRG eigenstates as vectors of the full
Fock space of d + sites 0..N: the shell-N space is (kept states of N-1) x
(site N), on which the Wilson-chain H_N is diagonalised per (Q, S) from its
highest-weight states.  Kept/discarded labels are the oracle's (energies are
compared first).  On that basis:

- rho_Nmax = e^{-E/TempBar}/Z_tot on every state of the last shell (the exponent
  Temp/D_Mtemp of src/DM_NRG/DM_NRG.cpp:102-118 and :273-297, Z_tot from the
  discarded partition functions of :221-271);
- rho_N = Tr_{site N+1} rho_{N+1} plus, for FDM, e^{-E D_{N-1}/Temp} 4^{Nmax-N}/Z_tot
  on the discarded states (src/DM_NRG/DM_NRG_CalcRhoN.cpp:393-398);
- the shell sums are Lehmann sums of c_{d,up} over all spin components,
  restricted to the kept/discarded classes of FDM_SpecDens_M_STL
  (src/SpecFunc_allSTL.cpp:1253-2354) / CFS_SpecDens_M_STL (:702-1240) with the
  reference's block-level conditions, LogGauss/Gauss (FDM) or
  BroadDelta/LorentzAnders (CFS) broadening; the last shell is the Costi form;
- CalcSpecDM_NRG's shell window and CalcNorm follow from those.
Comparisons are gauge-invariant (Frobenius norms between degenerate clusters,
Lehmann sums)."""
import numpy as np
import pytest

import pyoracle as po
from ed_util import fock_annihilate, onechqs_chain_fock, partial_trace_last_orbital

LAM = 2.5
HLF = 0.5 * (1 + 1 / LAM)


def dn(N):
    return HLF * LAM ** (-(N - 1) / 2.0)


def _sectors(nmode):
    dim = 1 << nmode
    ne = np.array([bin(s).count("1") for s in range(dim)])
    up = sum(1 << m for m in range(0, nmode, 2))
    nup = np.array([bin(s & up).count("1") for s in range(dim)])
    return ne, nup - (ne - nup)


def _label(v, ne, sz2):
    nz = np.abs(v) > 1e-12
    a, b = set(ne[nz]), set(sz2[nz])
    assert len(a) == 1 and len(b) == 1
    return a.pop(), b.pop()


def _nrg_fock(p, d):
    """Per shell: dict(V=m-resolved eigenstates (columns), E, kept, blk, S, hw=m=S states per block)."""
    shells = []
    Kprev = np.eye(4)                     # H_{-1}: all four impurity states kept
    for N in range(d.nshell_max + 1):
        H, C = onechqs_chain_fock(p["U"], p["Gamma"], p["ed"], LAM, N, a_lambda=True)
        nm = 2 * (N + 2)
        dim, dl = 1 << nm, 1 << (nm - 2)
        ne, sz2 = _sectors(nm)
        nk = Kprev.shape[1]
        B = np.zeros((dim, 4 * nk))
        for h in range(4):
            B[h * dl:(h + 1) * dl, h * nk:(h + 1) * nk] = Kprev
        lab = np.array([_label(B[:, c], ne, sz2) for c in range(B.shape[1])])
        Sp = sum(C[2 * o].T @ C[2 * o + 1] for o in range(N + 2))
        sh = d.shell(N)
        raw, hws = [], []
        for b in range(len(sh["qn"])):
            Q, S = sh["qn"][b]
            cols = np.where((lab[:, 0] - (N + 2) == int(round(Q))) & (lab[:, 1] == int(round(2 * S))))[0]
            V = B[:, cols]
            M = Sp @ V
            u, s, vt = np.linalg.svd(M, full_matrices=True)
            null = vt[np.sum(s > 1e-9):].T
            X = V @ null
            e, W = np.linalg.eigh(X.T @ H @ X)
            raw.append(e)
            hws.append(X @ W)
        emin = min(e.min() for e in raw)
        Sm = Sp.T
        Vs, Es, kp, blk, Ss = [], [], [], [], []
        for b in range(len(sh["qn"])):
            Q, S = sh["qn"][b]
            i0, i1 = sh["bbe"][b]
            en = sh["en"][i0:i1 + 1]
            kb = sh["kept"][i0:i1 + 1]
            e = raw[b] - emin
            assert len(e) == len(en)
            np.testing.assert_allclose(e, en, atol=1e-10)
            kc = int(kb.sum())
            assert np.all(kb[:kc] == 1) and np.all(kb[kc:] == 0)   # kept = energy prefix
            for i in range(len(e)):
                v = hws[b][:, i]
                m = S
                while m >= -S - 1e-9:
                    Vs.append(v.copy())
                    Es.append(en[i]); kp.append(1 if i < kc else 0); blk.append(b); Ss.append(S)
                    if m - 1 >= -S - 1e-9:
                        v = Sm @ v
                        v = v / np.linalg.norm(v)
                    m -= 1.0
        V = np.array(Vs).T
        shells.append(dict(V=V, E=np.array(Es), kept=np.array(kp), blk=np.array(blk), S=np.array(Ss),
                           hw=hws, C=C, nblk=len(sh["qn"]),
                           kc=np.array([int(sh["kept"][sh["bbe"][b, 0]:sh["bbe"][b, 1] + 1].sum())
                                        for b in range(len(sh["qn"]))]),
                           nb=np.array([int(sh["bbe"][b, 1] - sh["bbe"][b, 0] + 1) for b in range(len(sh["qn"]))])))
        Kprev = V[:, np.array(kp) == 1]
    return shells


def _rho_fock(sh, d, mtemp, betabar, fdm, exact=False):
    """exact: the full density matrix of Weichselbaum and von Delft, every shell's
    Boltzmann factors at e^{-E D_N/T} (rho_Nmax included); otherwise the
    reference's D_{N-1} scales and rho_Nmax exponent betabar = D_Mtemp/T."""
    NM = d.nshell_max
    temp = dn(mtemp) / betabar
    sc = (lambda N: dn(N)) if exact else (lambda N: dn(N - 1))
    ZN = np.sum(np.exp(-sh[NM]["E"] * sc(NM) / temp))
    ztot = ZN
    for N in range(NM - 1, -1, -1):
        dsc = sh[N]["kept"] == 0
        ztot += np.sum(np.exp(-sh[N]["E"][dsc] * sc(N) / temp)) * 4.0 ** (NM - N)
    norm = ztot if fdm else ZN
    V = sh[NM]["V"]
    bmax = sc(NM) / temp if exact else betabar
    rho = {NM: (V * (np.exp(-bmax * sh[NM]["E"]) / norm)) @ V.T}
    for N in range(NM - 1, -1, -1):
        R = partial_trace_last_orbital(rho[N + 1], 2 * (N + 3))
        if fdm:
            dsc = sh[N]["kept"] == 0
            Vd = sh[N]["V"][:, dsc]
            w = np.exp(-sh[N]["E"][dsc] * sc(N) / temp) * 4.0 ** (NM - N) / ztot
            R = R + (Vd * w) @ Vd.T
        rho[N] = R
    return rho, temp, (ztot if fdm else 0.0)


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


def _bd(method, w, x, tdn, twindow=2.0, bt=0.8, b=0.5 * np.log(LAM)):
    with np.errstate(divide="ignore"):
        return _bd_(method, w, x, tdn, twindow, bt, b)


def _bd_(method, w, x, tdn, twindow, bt, b):
    x = np.where(x == 0.0, 1e-100, x)
    if method == 2:
        a1 = np.log(np.abs(w / x)) / b
        lg = 1.0 / (b * np.abs(x) * np.sqrt(np.pi)) * np.exp(-(a1 - b / 4) ** 2)
        tw = 1.0 / (bt * np.sqrt(np.pi)) * np.exp(-((w - x) / bt) ** 2)
    else:
        lg = np.exp(-b * b / 4) / (b * np.abs(x) * np.sqrt(np.pi)) * np.exp(-(np.log(np.abs(w / x)) / b) ** 2)
        tw = (1 / np.pi) * bt / ((w - x) ** 2 + bt ** 2)
    lg = np.where(w * x < 0, 0.0, lg)
    return np.where(np.abs(x) < twindow * tdn, tw, lg)


def _shell_sum(s, rho, method, w, temp, N, norm=False, exact=False):
    """exact: every pole with a discarded state at either end, weight
    (rho_a + rho_b)|c_ab|^2, at both signs of omega (the broadening picks the side)."""
    V = s["V"]
    cm = V.T @ s["C"][0] @ V
    r = V.T @ rho @ V
    W1 = (r.T @ cm) * cm
    W2 = (cm @ r.T) * cm
    E, kp, blk = s["E"], s["kept"], s["blk"]
    dE = E[None, :] - E[:, None]
    Dl = np.ones_like(dE) if norm else _bd(method, w, dE, temp / dn(N))
    kc, nb = s["kc"], s["nb"]
    dc = nb - kc
    b1, b2 = blk[:, None], blk[None, :]
    skip = ((kc[b1] == 0) & (kc[b2] == 0)) | ((dc[b1] == 0) & (dc[b2] == 0))
    aK, aD = (kp == 1)[:, None], (kp == 0)[:, None]
    bK, bD = (kp == 1)[None, :], (kp == 0)[None, :]
    if exact:
        return np.sum(np.where(aK & bK, 0.0, W1 + W2) * Dl)
    if method == 2:
        pos = np.where(aK & bD, W1 + np.where(dc[b1] > 0, W2, 0.0), 0.0) + np.where(aD & bD, W1 + W2, 0.0)
        neg = np.where(aD & bK, W2 + np.where(dc[b2] > 0, W1, 0.0), 0.0) + np.where(aD & bD, W1 + W2, 0.0)
    else:
        pos = np.where(aK & bD, W1, 0.0)
        neg = np.where(aD & bK, W2, 0.0)
    pos = np.sum(np.where(skip, 0.0, pos) * Dl)
    neg = np.sum(np.where(skip, 0.0, neg) * Dl)
    if norm:
        return pos + neg
    return pos if w >= 0 else neg


def _costi(s, method, w, temp, N, betabar, norm=False, ztot=None):
    """ztot given: the last shell of the exact FDM, rho = e^{-E D_N/T}/Z_tot."""
    V = s["V"]
    cm = V.T @ s["C"][0] @ V
    E = s["E"]
    x = np.exp(-(dn(N) / temp if ztot else betabar) * E)
    Z = ztot if ztot else np.sum(x)
    dE = E[None, :] - E[:, None]
    Dl = np.ones_like(dE) if norm else _bd(method, w, dE, temp / dn(N))
    return np.sum(cm * cm * (x[:, None] + x[None, :]) * Dl) / Z


CASES = [
    # FDM, Mtemp = Nmax-1: rho_Nmax exponent and Z_N at the same shell
    dict(p=dict(U=0.5, Gamma=0.05, ed=-0.25), method=2, nsites=4, ncut=5, mtemp=2, betabar=1.3),
    # FDM, Mtemp < Nmax-1: rho_Nmax exponent from D_Mtemp (reference quirk), off particle-hole symmetry
    dict(p=dict(U=0.6, Gamma=0.07, ed=-0.2), method=2, nsites=4, ncut=6, mtemp=1, betabar=0.9),
    # CFS: rho from the plain sweep over kept + discarded states, the chain cut at Mtemp+1
    dict(p=dict(U=0.5, Gamma=0.05, ed=-0.25), method=1, nsites=6, ncut=5, mtemp=3, betabar=1.1),
    # exact FDM (every quirk switched off), Mtemp < Nmax-1, off particle-hole symmetry
    dict(p=dict(U=0.6, Gamma=0.07, ed=-0.2), method=2, nsites=4, ncut=6, mtemp=1, betabar=0.9, fix=63),
    # exact FDM, Mtemp = Nmax-1
    dict(p=dict(U=0.5, Gamma=0.05, ed=-0.25), method=2, nsites=4, ncut=5, mtemp=2, betabar=1.3, fix=63),
]


@pytest.fixture(scope="module", params=range(len(CASES)))
def case(request):
    c = CASES[request.param]
    p = c["p"]
    d = po.DmRun(U=p["U"], Gamma=p["Gamma"], ed=p["ed"], Lambda=LAM, nsitesmax=c["nsites"], ncut=c["ncut"],
                 Mtemp=c["mtemp"], betabar=c["betabar"], method=c["method"], update_bef_cut=1,
                 fdm_fix=c.get("fix", 0))
    sh = _nrg_fock(p, d)
    rho, temp, ztot = _rho_fock(sh, d, c["mtemp"], c["betabar"], c["method"] == 2, c.get("fix", 0) == 63)
    return c, d, sh, rho, temp, ztot


def test_fdm_shells_and_weights(case):
    c, d, sh, rho, temp, ztot = case
    if c["method"] == 2:
        assert d.nshell_max == c["nsites"] - 1 and d.spec_nshell_max == c["nsites"]
        assert abs(d.ztot - ztot) <= 1e-12 * ztot
    else:
        assert d.nshell_max == c["mtemp"]
    assert abs(d.temp - temp) <= 1e-14 * temp
    # the chain is truncated at every shell below the last
    for N in range(d.nshell_max):
        assert 0 < sh[N]["kept"].sum() < len(sh[N]["kept"])


def test_rho_matches_fock_space(case):
    c, d, sh, rho, temp, ztot = case
    for N in range(d.nshell_max, -1, -1):
        o = d.shell(N)
        assert abs(o["trace"] - np.trace(rho[N])) <= 1e-12 * max(1e-30, abs(np.trace(rho[N]))), N
        for b in range(len(o["rho"])):
            X = sh[N]["hw"][b]
            red = X.T @ rho[N] @ X
            R = o["rho"][b]
            en = o["en"][o["bbe"][b, 0]:o["bbe"][b, 1] + 1]
            for c1 in _clusters(en):
                for c2 in _clusters(en):
                    a = np.linalg.norm(R[np.ix_(c1, c2)])
                    x = np.linalg.norm(red[np.ix_(c1, c2)])
                    assert abs(a - x) <= 1e-10 * x + 1e-14, (N, b, c1, c2, a, x)


def test_shell_sums_match_fock_space(case):
    c, d, sh, rho, temp, ztot = case
    NM = d.nshell_max
    ex = c.get("fix", 0) == 63
    zl = ztot if ex else None
    for N in range(NM):
        for w in (-2.3, -0.6, -0.04, 0.0, 0.03, 0.5, 1.7):
            ref = _shell_sum(sh[N], rho[N], c["method"], w, temp, N, exact=ex)
            got = d.spec_shell(N, w)
            assert abs(got - ref) <= 1e-10 * max(1e-8, abs(ref)), (N, w, got, ref)
    for w in (-1.1, -0.1, 0.2, 2.0):
        ref = _costi(sh[NM], c["method"], w, temp, NM, c["betabar"], ztot=zl)
        got = d.spec_shell(NM, w)
        assert abs(got - ref) <= 1e-10 * max(1e-8, abs(ref)), (w, got, ref)


def test_norm_and_window_sum(case):
    c, d, sh, rho, temp, ztot = case
    NM = d.nshell_max
    ex = c.get("fix", 0) == 63
    zl = ztot if ex else None
    nrm = sum(_shell_sum(sh[N], rho[N], c["method"], 0.0, temp, N, norm=True, exact=ex) for N in range(NM))
    nrm += _costi(sh[NM], c["method"], 0.0, temp, NM, c["betabar"], norm=True, ztot=zl)
    assert abs(d.norm() - nrm) <= 1e-11 * abs(nrm)
    if ex:   # the full density matrix: sum rule <{c, c^dag}> = 1 and unit trace
        assert abs(nrm - 1.0) <= 1e-12, nrm
        tr = sum(np.trace(rho[N]) - np.trace(partial_trace_last_orbital(rho[N + 1], 2 * (N + 3)))
                 for N in range(NM)) + np.trace(rho[NM])
        assert abs(tr - 1.0) <= 1e-12, tr
    # CalcSpecDM_NRG: shells [N_w-8, N_w+14) of the spectral chain, A_N(w/D_N)/D_N
    for w in (-0.9, -0.05, 0.004, 0.3, 1.2):
        Nw = 1 - int(2.0 * np.log(2.0 * abs(w) / HLF) / np.log(LAM))
        Nw = max(Nw, 0)
        ref = 0.0
        for N in range(max(Nw - 8, 0), min(Nw + 14, d.spec_nshell_max)):
            if N == NM:
                ref += _costi(sh[N], c["method"], w / dn(N), temp, N, c["betabar"], ztot=zl) / dn(N)
            else:
                ref += _shell_sum(sh[N], rho[N], c["method"], w / dn(N), temp, N, exact=ex) / dn(N)
        got = d.spec_omega(w)
        assert abs(got - ref) <= 1e-10 * max(1e-8, abs(ref)), (w, got, ref)


def test_grid_uses_window_sums():
    # CFS/FDM grids: the fixed-omega grid starts two shells later and each value is CalcSpecDM_NRG(+-w)
    d = po.DmRun(U=0.5, Gamma=0.02, ed=-0.25, Lambda=LAM, nsitesmax=12, ncut=60, Mtemp=9, betabar=1.0,
                 method=2, update_bef_cut=1)
    w, r = d.spectrum()
    fac = LAM ** 1.25
    NMs = d.spec_nshell_max
    expect_w = sorted([s * fac * dn(N) for N in range(2, NMs) for s in (-1, 1)])
    np.testing.assert_allclose(w, expect_w, rtol=1e-15)
    # the first (most negative) and last points are shell-2 values of the even set,
    # not interpolated: equal to the window sum at those omegas
    assert abs(r[0] - d.spec_omega(w[0])) <= 1e-12 * abs(r[0])
    assert abs(r[-1] - d.spec_omega(w[-1])) <= 1e-12 * abs(r[-1])


def test_window_bounds_on_a_long_chain():
    # 22 shells: the window [N_w-8, N_w+14) is cut at one end or both for these
    # omegas; the sum over it equals the shell values assembled here, and the
    # shells left out carry no weight at the tolerances used (the window is a
    # work bound, not an approximation)
    d = po.DmRun(U=0.5, Gamma=0.02, ed=-0.25, Lambda=LAM, nsitesmax=22, ncut=40, Mtemp=20, betabar=1.0,
                 method=2, update_bef_cut=1)
    cut = 0
    for w in (1.5, -0.8, 0.05, 2.0e-3, -1.0e-4):
        Nw = max(1 - int(2.0 * np.log(2.0 * abs(w) / HLF) / np.log(LAM)), 0)
        lo, hi = max(Nw - 8, 0), min(Nw + 14, d.spec_nshell_max)
        cut += (lo > 0) + (hi < d.spec_nshell_max)
        ref = sum(d.spec_shell(N, w / dn(N)) / dn(N) for N in range(lo, hi))
        assert abs(d.spec_omega(w) - ref) <= 1e-13 * abs(ref), w
        full = sum(d.spec_shell(N, w / dn(N)) / dn(N) for N in range(d.spec_nshell_max))
        assert abs(full - ref) <= 1e-12 * abs(ref), w
    assert cut >= 4
