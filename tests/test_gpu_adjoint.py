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
