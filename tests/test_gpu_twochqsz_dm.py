"""DM-NRG, CFS and FDM of the TwoChQSz two-channel chain on the device (nrg_dm_create
with model 1: the chain keeps every state and d_up per shell, the sweep runs
over (Q, Sz) blocks with degeneracy 1 and no Clebsch-Gordan factors, the FDM
weights count 16 states per later site, or 4 as the reference's one-channel
code) against the CPU oracle on the same
inputs: Z_tot and w_N, rho_N traces and gauge-invariant block norms, the shell
sums, the shell-window sums and CalcNorm, and the spectral function on both
omega grids.  Tolerance: 1e-8 relative for spectral values, 1e-10 for traces
and weights."""
import numpy as np
import pytest

import pyoracle as po

pytestmark = pytest.mark.gpu

BASE = dict(U=0.5, Gamma=0.03, Gamma2=0.01, ed=-0.25, Lambda=2.5, model=1)
CASES = [
    # FDM, truncated from N = 1 (ncut 200 of 1024 states), rho_Nmax from D_Mtemp below the last shell
    dict(BASE, nsitesmax=5, ncut=200, Mtemp=3, betabar=0.779, method=2, bbroad=0.6),
    # FDM off particle-hole symmetry with equal channels, Mtemp = Nmax - 1
    dict(BASE, U=0.3, Gamma=0.05, Gamma2=0.05, ed=-0.1, nsitesmax=4, ncut=300, Mtemp=3, betabar=1.0, method=2),
    # CFS: the chain is cut at Mtemp + 1
    dict(BASE, nsitesmax=5, ncut=200, Mtemp=3, betabar=0.779, method=1, bbroad=0.6),
    # FDM with the one-channel environment count 4 of the reference's weights
    dict(BASE, nsitesmax=4, ncut=200, Mtemp=3, betabar=0.779, method=2, bbroad=0.6, fdm_site_dim=4.0),
    # the exact full density matrix (every FDM quirk of the reference switched off), truncated,
    # Mtemp below Nmax - 1, environment count 16 and 4
    dict(BASE, nsitesmax=5, ncut=200, Mtemp=2, betabar=0.779, method=2, bbroad=0.6, fdm_fix=63),
    dict(BASE, nsitesmax=5, ncut=200, Mtemp=2, betabar=0.779, method=2, bbroad=0.6, fdm_fix=63,
         fdm_site_dim=4.0),
    # one quirk switched off alone: the all-discarded block pairs keep their terms
    dict(BASE, nsitesmax=5, ncut=200, Mtemp=2, betabar=0.779, method=2, bbroad=0.6, fdm_fix=8),
]


@pytest.fixture(scope="module")
def ctx():
    import twochnrg_amd as t
    c = t.Context(0)
    yield c
    c.close()


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


def _rel(got, ref):
    got, ref = np.asarray(got), np.asarray(ref)
    scale = np.maximum(np.abs(ref), 1e-12 * max(np.abs(ref).max(), 1e-300))
    return np.abs(got - ref) / scale


@pytest.fixture(scope="module")
def runs():
    return {}


def _oracle(runs, i):
    if i not in runs:
        runs[i] = po.DmRun(**CASES[i])
    return runs[i]


@pytest.mark.parametrize("i", range(len(CASES)))
def test_qsz_rho_and_weights(ctx, runs, i):
    case = CASES[i]
    print("twochqsz dm rho/weights", case, flush=True)
    o = _oracle(runs, i)
    with ctx.dm(**case) as d:
        assert (d.nshell_max, d.spec_nshell_max) == (o.nshell_max, o.spec_nshell_max)
        assert abs(d.temp - o.temp) <= 1e-15 * o.temp
        wn, ztot = d.weights()
        if case["method"] == 2:
            assert abs(ztot - o.ztot) <= 1e-12 * o.ztot
            wo = np.array([o.shell(N)["wn"] for N in range(o.nshell_max + 1)])
            np.testing.assert_allclose(wn, wo, rtol=1e-11, atol=1e-300)
            assert abs(wn.sum() - 1.0) <= 1e-12
        discarded = 0
        for N in range(o.nshell_max, -1, -1):
            sh = o.shell(N)
            discarded += int((sh["kept"] == 0).sum())
            sizes = [int(b1 - b0 + 1) for b0, b1 in sh["bbe"]]
            blocks, tr = d.rho(N, sizes)
            assert abs(tr - sh["trace"]) <= 1e-10 * abs(sh["trace"]), (N, tr, sh["trace"])
            for b, (R, Ro) in enumerate(zip(blocks, sh["rho"])):
                en = sh["en"][sh["bbe"][b, 0]:sh["bbe"][b, 1] + 1]
                scale = max(1e-300, np.abs(Ro).max())
                for c1 in _clusters(en):
                    for c2 in _clusters(en):
                        a = np.linalg.norm(R[np.ix_(c1, c2)])
                        x = np.linalg.norm(Ro[np.ix_(c1, c2)])
                        assert abs(a - x) <= 1e-9 * scale, (N, b, c1, c2, a, x)
        assert discarded > 0


@pytest.mark.parametrize("i", range(len(CASES)))
def test_qsz_shell_sums_and_norm(ctx, runs, i):
    case = CASES[i]
    print("twochqsz dm shell sums", case, flush=True)
    o = _oracle(runs, i)
    with ctx.dm(**case) as d:
        sh, w = [], []
        for N in range(o.spec_nshell_max):
            for x in (-4.3, -1.0, -0.2, -1e-3, 0.0, 1e-3, 0.05, 0.7, 5.0):
                sh.append(N)
                w.append(x)
        got = d.spec(sh, w)
        ref = np.array([o.spec_shell(N, x) for N, x in zip(sh, w)])
        err = _rel(got, ref)
        assert err.max() <= 1e-8, (err.max(), sh[int(err.argmax())], w[int(err.argmax())])
        ws = np.array([-1.3, -0.2, -3e-3, -2e-5, 4e-5, 1e-3, 0.06, 0.9])
        got = d.spec_omega(ws)
        ref = np.array([o.spec_omega(x) for x in ws])
        err = _rel(got, ref)
        assert err.max() <= 1e-8, (err.max(), ws[int(err.argmax())])
        n = d.norm()
        assert abs(n - o.norm()) <= 1e-10 * abs(o.norm())
        if case.get("fdm_fix") == 63:   # the sum rule of the full density matrix
            assert abs(n - 1.0) <= 1e-10, n


@pytest.mark.parametrize("Nw", [0, 1])
def test_qsz_spectrum(ctx, runs, Nw):
    case = dict(CASES[0], Nw=Nw)
    print("twochqsz dm spectrum", case, flush=True)
    o = po.DmRun(**case)
    wo, ro = o.spectrum()
    with ctx.dm(**case) as d:
        w, r = d.spectrum()
    np.testing.assert_allclose(w, wo, rtol=1e-14, atol=0)
    assert _rel(r, ro).max() <= 1e-8


# plain DM-NRG (UseCFS 0) over the same shells: rho_N from the kept states'
# children, the shell sums over every pair of each shell
DMNRG = [
    dict(BASE, nsitesmax=5, ncut=200, Mtemp=3, betabar=0.9, method=0),
    dict(BASE, U=0.3, Gamma=0.05, Gamma2=0.05, ed=-0.1, nsitesmax=4, ncut=300, Mtemp=3, betabar=1.1, method=0),
]


@pytest.mark.parametrize("i", range(len(DMNRG)))
def test_qsz_dmnrg(ctx, i):
    case = DMNRG[i]
    print("twochqsz plain dm-nrg", case, flush=True)
    o = po.DmRun(**case)
    with ctx.dm(**case) as d:
        assert (d.nshell_max, d.spec_nshell_max) == (o.nshell_max, o.spec_nshell_max)
        assert abs(d.temp - o.temp) <= 1e-15 * o.temp
        for N in range(o.nshell_max, -1, -1):
            sh = o.shell(N)
            sizes = [int(b1 - b0 + 1) for b0, b1 in sh["bbe"]]
            blocks, tr = d.rho(N, sizes)
            assert abs(tr - sh["trace"]) <= 1e-10 * abs(sh["trace"]), (N, tr, sh["trace"])
            for b, (R, Ro) in enumerate(zip(blocks, sh["rho"])):
                en = sh["en"][sh["bbe"][b, 0]:sh["bbe"][b, 1] + 1]
                scale = max(1e-300, np.abs(Ro).max())
                for c1 in _clusters(en):
                    for c2 in _clusters(en):
                        a = np.linalg.norm(R[np.ix_(c1, c2)])
                        x = np.linalg.norm(Ro[np.ix_(c1, c2)])
                        assert abs(a - x) <= 1e-9 * scale, (N, b, c1, c2, a, x)
        sh, w = [], []
        for N in range(o.spec_nshell_max):
            for x in (-4.3, -1.0, -0.2, -1e-3, 1e-3, 0.05, 0.7, 5.0):
                sh.append(N)
                w.append(x)
        got = d.spec(sh, w)
        ref = np.array([o.spec_shell(N, x) for N, x in zip(sh, w)])
        err = _rel(got, ref)
        assert err.max() <= 1e-8, (err.max(), sh[int(err.argmax())], w[int(err.argmax())])
    for Nw in (0, 1):
        c2 = dict(case, Nw=Nw)
        o2 = po.DmRun(**c2)
        wo, ro = o2.spectrum()
        with ctx.dm(**c2) as d:
            wg, rg = d.spectrum()
        np.testing.assert_allclose(wg, wo, rtol=1e-14, atol=0)
        assert _rel(rg, ro).max() <= 1e-8


def test_qsz_rejects_derivatives_and_bad_model(ctx):
    import twochnrg_amd as t
    with pytest.raises(t.NrgError):
        ctx.dm(**dict(CASES[0], method=2, model=2))
    with pytest.raises(t.NrgError):
        ctx.dm(**dict(DMNRG[0]), dpar=[[1.0, 0.0, 0.0]])
