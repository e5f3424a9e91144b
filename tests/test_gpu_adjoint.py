"""Derivative pass of the fixed-truncation OneChQS loop on the device
(nrg_onechqs_jvp through the C-ABI) against the oracle's derivative, against
central differences of the device loop with frozen Kept sets, and through
the torch.autograd harness."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle as po  # noqa: E402

pytestmark = pytest.mark.gpu

KEYS = ("U", "Gamma", "ed")


@pytest.fixture(scope="module")
def ctx():
    import twochnrg_amd as t
    if t.device_count() < 1:
        pytest.fail("no GPU visible")
    c = t.Context(0)
    yield c
    c.close()


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


def _rel(a, b, floor):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), floor)


@pytest.mark.parametrize("case", [
    dict(U=0.5, Gamma=0.02, ed=-0.2, Lambda=2.5, nsitesmax=9, ncut=120),
    dict(U=0.5, Gamma=0.03, ed=-0.25, Lambda=2.5, nsitesmax=8, ncut=100),
    dict(U=0.3, Gamma=0.05, ed=-0.1, Lambda=3.0, nsitesmax=4, ncut=10000, disc_scheme=1),
])
def test_jvp_matches_oracle(ctx, case):
    kw = dict(case, calcdens=2)
    print("jvp vs oracle", kw, flush=True)
    base = po.run_onechqs(**kw)
    kept = po.kept_arrays(base)
    os_, ot = po.run_onechqs_jvp(kept=kept, dpar=np.eye(3), **kw)
    gs, gt = ctx.onechqs_jvp(kept=kept, dpar=np.eye(3), **kw)
    assert len(gs) == len(os_)
    for g, o, a, b in zip(gs, os_, gt, ot):
        assert np.array_equal(g["qn"], o["qn"]) and np.array_equal(g["kept"], o["kept"])
        assert np.max(np.abs(g["en"] - o["en"]) / np.maximum(np.abs(o["en"]), 1.0)) <= 1e-10
        for f in ("dsuscep", "dentropy", "dopavg"):
            assert _rel(a[f], b[f], 1e-3) <= 1e-8, (g["N"], f, a[f], b[f])
        for t in range(3):
            assert _rel(_cluster_sums(o, a["den"][t]), _cluster_sums(o, b["den"][t]), 1.0) <= 1e-8, (g["N"], t)


def test_plain_and_frozen_runs_agree(ctx):
    """ntan = 0 with the Kept sets of a plain device run reproduces that run."""
    kw = dict(U=0.5, Gamma=0.02, ed=-0.2, Lambda=2.5, nsitesmax=12, ncut=200, calcdens=2)
    import twochnrg_amd as t
    plain, _ = ctx.run_onechqs(**kw)
    fz, tz = ctx.onechqs_jvp(kept=t.kept_arrays(plain), **kw)
    for a, b in zip(plain, fz):
        assert np.array_equal(a["kept"], b["kept"]) and np.array_equal(a["en"], b["en"])
        assert np.array_equal(a["opavg"], b["opavg"]) and np.array_equal(a["suscep"], b["suscep"])


def test_c4_gradients_vs_central_fd(ctx):
    """Single-channel Anderson, Lambda 2.5, Ncut 1000 fixed truncation, T = 1e-6
    at N = 31 (betabar = D_31 / T): d<n_d> and d(T chi) with respect to
    (U, Gamma, ed) against central differences h = 1e-5 of the device loop with
    the base point's Kept sets, at the particle-hole symmetric C1 point and
    off it (ed = -0.2)."""
    import twochnrg_amd as t
    L, N = 2.5, 31
    hlf = 0.5 * (1 + 1 / L)
    bb = hlf * L ** (-(N - 1) / 2.0) / 1e-6
    for ed in (-0.25, -0.2):
        kw = dict(U=0.5, Gamma=0.02, ed=ed, Lambda=L, nsitesmax=N + 1, ncut=1000, calcdens=2, betabar=bb)
        print("c4 point", kw, flush=True)
        plain, _ = ctx.run_onechqs(**kw)
        kept = t.kept_arrays(plain)
        st, tn = ctx.onechqs_jvp(kept=kept, dpar=np.eye(3), **kw)
        s, d = st[-1], tn[-1]
        assert s["N"] == N and abs(s["T_avg"] - 1e-6) < 1e-15
        an = np.stack([d["dopavg"][:, 0], d["dsuscep"][:, 3]])
        fd = np.zeros((2, 3))
        for q, key in enumerate(KEYS):
            h = 1e-5
            kp, km = dict(kw), dict(kw)
            kp[key] += h
            km[key] -= h
            sp, _ = ctx.onechqs_jvp(kept=kept, **kp)
            sm, _ = ctx.onechqs_jvp(kept=kept, **km)
            fd[0, q] = (sp[-1]["opavg"][0] - sm[-1]["opavg"][0]) / (2 * h)
            fd[1, q] = (sp[-1]["suscep"][3] - sm[-1]["suscep"][3]) / (2 * h)
        print("c4 jacobian", an.tolist(), "fd", fd.tolist(), flush=True)
        # central differences with h = 1e-5 carry O(h^2) truncation; relative to each row's scale
        for r in range(2):
            assert np.max(np.abs(an[r] - fd[r])) <= 1e-5 * max(np.max(np.abs(fd[r])), 1e-2), (ed, r, an, fd)
        if ed == -0.25:   # <n_d> = 1 on the particle-hole line: d/dGamma = 0, (d/dU - d/ded / 2) = 0
            assert abs(an[0, 1]) < 1e-9 and abs(an[0, 0] - 0.5 * an[0, 2]) < 1e-9


def test_autograd_harness(ctx):
    """torch.autograd.gradcheck of (<n_d>, T chi) at N = 7 through NRGObservables."""
    import torch
    from twochnrg_amd import autograd as ag
    run = ag.make_run(ctx, 7, (0.5, 0.03, -0.2), Lambda=2.5, ncut=100, calcdens=2)
    x = torch.tensor([0.5, 0.03, -0.2], dtype=torch.float64, requires_grad=True)
    y = ag.observables(x, run)
    y.sum().backward()
    assert x.grad is not None and torch.all(torch.isfinite(x.grad))
    assert torch.autograd.gradcheck(lambda p: ag.observables(p, run), (x,), eps=1e-6, atol=1e-6, rtol=1e-4)


def test_dm_jvp_matches_oracle(ctx):
    """Derivative of the DM-NRG shell sums on the device against the oracle's,
    every shell (the last in the Costi form), 7 omegas per shell."""
    import twochnrg_amd as t
    kw = dict(U=0.5, Gamma=0.03, ed=-0.2, Lambda=2.5, nsitesmax=11, ncut=150, Mtemp=8, betabar=0.9)
    print("dm jvp vs oracle", kw, flush=True)
    base = po.run_onechqs(U=0.5, Gamma=0.03, ed=-0.2, Lambda=2.5, nsitesmax=11, ncut=150, calcdens=1)
    kept = po.kept_arrays(base)
    D = po.DmRun(kept=kept, dpar=np.eye(3), **kw)
    ws = [0.3, -0.7, 1.5, -2.2, 3.0, 0.05, -0.02]
    shells = [N for N in range(D.spec_nshell_max) for _ in ws]
    wb = [w for _ in range(D.spec_nshell_max) for w in ws]
    ref = np.array([D.spec_shell_jvp(N, w) for N, w in zip(shells, wb)])
    with ctx.dm(kept=kept, dpar=np.eye(3), **kw) as d:
        assert d.spec_nshell_max == D.spec_nshell_max
        got = d.spec_jvp(shells, wb)
        val = d.spec(shells, wb)
    assert np.allclose(val, [D.spec_shell(N, w) for N, w in zip(shells, wb)], rtol=1e-10, atol=1e-12)
    for q in range(3):
        err = np.max(np.abs(got[:, q] - ref[:, q])) / np.max(np.abs(ref[:, q]))
        assert err <= 1e-8, (q, err)


def _dm_fd(ctx, x0, kept, kw, sh, wb, dn, q, h):
    xp, xm = list(x0), list(x0)
    xp[q] += h
    xm[q] -= h
    with ctx.dm(U=xp[0], Gamma=xp[1], ed=xp[2], kept=kept, **kw) as d:
        ap = d.spec(sh, wb) / dn
    with ctx.dm(U=xm[0], Gamma=xm[1], ed=xm[2], kept=kept, **kw) as d:
        am = d.spec(sh, wb) / dn
    return (ap - am) / (2 * h)


def _window_crossing_shells(ctx, x0, kept, kw, q, h, temp):
    """Shells N where a pole E_j - E_i of c_d (kept states of blocks (Q, S) and
    (Q+1, S +- 1/2)) lies on different sides of the BDeltaTemp window edge
    twindow T / D_N at x0 - h and x0 + h (frozen Kept sets)."""
    from twochnrg_amd import autograd as ag
    tw = kw["twindow"] if kw["twindow"] > 0 else 2.0
    side = []
    for sgn in (1.0, -1.0):
        x = list(x0)
        x[q] += sgn * h
        steps, _ = ctx.onechqs_jvp(U=x[0], Gamma=x[1], ed=x[2], Lambda=kw["Lambda"], nsitesmax=kw["nsitesmax"],
                                   ncut=kw["ncut"], calcdens=1, kept=kept)
        per = {}
        for st in steps:
            N = st["N"]
            if N < 0:
                continue
            edge = tw * temp / ag.calc_dn(kw["Lambda"], N)
            blk = []
            for (Q, S), (b0, b1) in zip(st["qn"][:, :2], st["bbe"]):
                e = st["en"][b0:b1 + 1][st["kept"][b0:b1 + 1]]
                blk.append((Q, S, e))
            inside = []
            for Q1, S1, e1 in blk:
                for Q2, S2, e2 in blk:
                    if abs(Q2 - Q1 - 1) < 1e-9 and abs(abs(S2 - S1) - 0.5) < 1e-9 and len(e1) and len(e2):
                        inside.append((np.abs(e2[None, :] - e1[:, None]) < edge).ravel())
            per[N] = np.concatenate(inside) if inside else np.zeros(0, bool)
        side.append(per)
    return {N for N in side[0] if N in side[1] and np.any(side[0][N] != side[1][N])}


def test_c4_spectral_gradients_vs_central_fd(ctx):
    """dA(omega)/d(U, Gamma, ed) on omega = +-logspace(-8, 0, 100) at T = 1e-6
    (Mtemp 31, betabar = D_31 / T), single-channel Anderson Lambda 2.5, Ncut
    1000 fixed truncation, against central differences of the device DM-NRG
    with the base point's Kept sets; plus the autograd harness.

    The reference broadens a pole with a Lorentzian (BDeltaTemp) when
    |E_j - E_i| < twindow T / D_N and with BroadDelta otherwise, so A(omega)
    jumps where a pole crosses that edge; the derivative is that of the
    piecewise function.  The tight check therefore uses a vanishing window
    (twindow 1e-12, A smooth): Richardson differences of relative steps 1e-5
    and 5e-6.  With the default window (2) at the relative step 1e-5 every
    omega whose shell has no pole crossing the window edge between x0 - h and
    x0 + h agrees to 1e-6, and the omegas that disagree by more than 1e-4 all
    sit on such shells."""
    import torch
    from twochnrg_amd import autograd as ag
    L, M = 2.5, 31
    bb = ag.calc_dn(L, M) / 1e-6
    om = np.concatenate([-np.logspace(-8, 0, 100)[::-1], np.logspace(-8, 0, 100)])
    x0 = (0.5, 0.02, -0.2)
    for tw in (1e-12, 0.0):
        kw = dict(Lambda=L, nsitesmax=40, ncut=1000, Mtemp=M, betabar=bb, twindow=tw)
        print("c4 spectral", kw, flush=True)
        run = ag.make_spectral_run(ctx, x0, om, **kw)
        x = torch.tensor(x0, dtype=torch.float64, requires_grad=True)
        A = ag.spectral(x, run)
        assert A.shape == (200,) and torch.all(torch.isfinite(A))
        v = torch.linspace(-1.0, 1.0, 200, dtype=torch.float64)
        (A * v).sum().backward()
        grid = run["grid"]
        sh, wb = [g[0] for g in grid], [g[1] for g in grid]
        dn = np.array([g[2] for g in grid])
        with ctx.dm(U=x0[0], Gamma=x0[1], ed=x0[2], kept=run["kept"], dpar=np.eye(3), **kw) as d:
            J = d.spec_jvp(sh, wb) / dn[:, None]
        assert np.allclose(x.grad.numpy(), v.numpy() @ J, rtol=1e-12, atol=0)
        for q in range(3):
            h = 1e-5 * abs(x0[q])
            d1 = _dm_fd(ctx, x0, run["kept"], kw, sh, wb, dn, q, h)
            scale = np.max(np.abs(J[:, q]))
            if tw > 0:
                d2 = _dm_fd(ctx, x0, run["kept"], kw, sh, wb, dn, q, h / 2)
                err = np.abs((4 * d2 - d1) / 3 - J[:, q]) / scale
                print("c4 dA/d%s (smooth window): max |dA| %.4g, max rel err %.3g" % (KEYS[q], scale, err.max()),
                      flush=True)
                assert err.max() <= 1e-6, (KEYS[q], err.max())
            else:
                err = np.abs(d1 - J[:, q]) / scale
                cross = _window_crossing_shells(ctx, x0, run["kept"], kw, q, h, 1e-6)
                hit = np.array([g in cross for g in sh])
                bad = err > 1e-4
                print("c4 dA/d%s (default window): max |dA| %.4g, median rel err %.3g, %d of 200 above 1e-4; "
                      "shells with a pole crossing the window edge %s (%d omegas), worst elsewhere %.3g" % (
                          KEYS[q], scale, np.median(err), int(bad.sum()), sorted(cross), int(hit.sum()),
                          err[~hit].max() if (~hit).any() else 0.0), flush=True)
                # the omegas of a shell with a pole that the step carries across the
                # window edge see the jump; every other omega agrees
                assert not np.any(bad & ~hit), (KEYS[q], np.flatnonzero(bad & ~hit))
                assert np.all(err[~hit] <= 1e-6), (KEYS[q], err[~hit].max())
