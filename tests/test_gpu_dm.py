"""DM-NRG on the device (nrg_dm_*) against the CPU oracle (orc_dm_*) on the same
inputs: rho_N traces and gauge-invariant block norms, the shell spectral sums
DMNRG_SpecDens_M_STL / CalcSpecCosti_T_N, and the spectral function on both
omega grids.  Tolerance: 1e-8 relative for A(omega) (the parity bar for
spectral results), 1e-10 for the density-matrix traces."""
import numpy as np
import pytest

import pyoracle as po

pytestmark = pytest.mark.gpu

CASES = [
    dict(U=0.5, Gamma=0.02, ed=-0.25, Lambda=2.5, nsitesmax=12, ncut=200, Mtemp=10, betabar=1.0),
    dict(U=0.6, Gamma=0.04, ed=-0.2, Lambda=2.5, nsitesmax=11, ncut=150, Mtemp=8, betabar=0.8),
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


@pytest.mark.parametrize("case", CASES)
def test_dm_rho_sweep(ctx, case):
    print("dm rho sweep", case, flush=True)
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


@pytest.mark.parametrize("case", CASES)
def test_dm_shell_sums(ctx, case):
    print("dm shell sums", case, flush=True)
    o = po.DmRun(**case)
    with ctx.dm(**case) as d:
        sh, w = [], []
        for N in range(o.spec_nshell_max):
            for x in (-4.3, -3.1, -1.0, -0.2, 0.05, 0.7, 3.1, 5.0):
                sh.append(N)
                w.append(x)
        got = d.spec(sh, w)
        ref = np.array([o.spec_shell(N, x) for N, x in zip(sh, w)])
        scale = np.maximum(np.abs(ref), 1e-12 * np.abs(ref).max())
        err = np.abs(got - ref) / scale
        assert err.max() <= 1e-8, (err.max(), sh[int(err.argmax())], w[int(err.argmax())])


@pytest.mark.parametrize("Nw", [0, 3])
def test_dm_spectrum(ctx, Nw):
    case = dict(CASES[0], Nw=Nw)
    print("dm spectrum", case, flush=True)
    o = po.DmRun(**case)
    wo, ro = o.spectrum()
    with ctx.dm(**case) as d:
        w, r = d.spectrum()
    # the grid points are products of the same factors in a different order
    np.testing.assert_allclose(w, wo, rtol=1e-14, atol=0)
    scale = np.maximum(np.abs(ro), 1e-12 * np.abs(ro).max())
    assert np.max(np.abs(r - ro) / scale) <= 1e-8
