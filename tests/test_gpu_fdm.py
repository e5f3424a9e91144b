"""CFS and FDM on the device (nrg_dm_* with method 1 / 2 on a chain run with
UpdateBefCut 1) against the CPU oracle (orc_dm_*) on the same inputs: the FDM
weights w_N and Z_tot, rho_N (kept block from the sweep, discarded diagonal)
traces and gauge-invariant block norms, the shell sums CFS_SpecDens_M_STL /
FDM_SpecDens_M_STL / CalcSpecCosti_T_N, the shell-window sums of
CalcSpecDM_NRG, CalcNorm and the spectral function on both omega grids.
Tolerance: 1e-8 relative for spectral values, 1e-10 for traces and weights."""
import numpy as np
import pytest

import pyoracle as po

pytestmark = pytest.mark.gpu

CASES = [
    # FDM, rho_Nmax exponent from D_Mtemp below the last shell (the reference's -M with -F)
    dict(U=0.5, Gamma=0.02, ed=-0.25, Lambda=2.5, nsitesmax=14, ncut=120, Mtemp=10, betabar=1.0, method=2,
         update_bef_cut=1),
    # FDM off particle-hole symmetry, Mtemp = Nmax-1, log-Gaussian width 0.6
    dict(U=0.6, Gamma=0.04, ed=-0.2, Lambda=2.5, nsitesmax=12, ncut=150, Mtemp=10, betabar=0.8, method=2,
         update_bef_cut=1, bbroad=0.6),
    # CFS: the chain is cut at Mtemp+1, rho from the plain sweep over kept + discarded states
    dict(U=0.5, Gamma=0.02, ed=-0.25, Lambda=2.5, nsitesmax=14, ncut=120, Mtemp=10, betabar=1.0, method=1,
         update_bef_cut=1),
    # the exact full density matrix (every FDM quirk of the reference switched off), off particle-hole
    # symmetry, Mtemp below Nmax - 1
    dict(U=0.6, Gamma=0.04, ed=-0.2, Lambda=2.5, nsitesmax=12, ncut=150, Mtemp=8, betabar=0.8, method=2,
         update_bef_cut=1, bbroad=0.6, fdm_fix=63),
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


@pytest.mark.parametrize("case", CASES)
def test_fdm_rho_and_weights(ctx, case):
    print("fdm rho/weights", case, flush=True)
    o = po.DmRun(**case)
    with ctx.dm(**case) as d:
        assert (d.nshell_max, d.spec_nshell_max) == (o.nshell_max, o.spec_nshell_max)
        assert abs(d.temp - o.temp) <= 1e-15 * o.temp
        wn, ztot = d.weights()
        if case["method"] == 2:
            assert abs(ztot - o.ztot) <= 1e-12 * o.ztot
            wo = np.array([o.shell(N)["wn"] for N in range(o.nshell_max + 1)])
            np.testing.assert_allclose(wn, wo, rtol=1e-12, atol=1e-300)
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


@pytest.mark.parametrize("case", CASES)
def test_fdm_shell_sums_and_norm(ctx, case):
    print("fdm shell sums", case, flush=True)
    o = po.DmRun(**case)
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


@pytest.mark.parametrize("Nw", [0, 2])
@pytest.mark.parametrize("method", [1, 2])
def test_fdm_spectrum(ctx, method, Nw):
    case = dict(CASES[0], method=method, Nw=Nw)
    print("fdm spectrum", case, flush=True)
    o = po.DmRun(**case)
    wo, ro = o.spectrum()
    with ctx.dm(**case) as d:
        w, r = d.spectrum()
    np.testing.assert_allclose(w, wo, rtol=1e-14, atol=0)
    assert _rel(r, ro).max() <= 1e-8
