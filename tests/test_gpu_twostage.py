"""Two-stage reduction (dense -> band of width 16 -> tridiagonal) of the
batched eigensolver against LAPACK (numpy) and against the one-stage path:
eigenvalues, residuals and orthogonality at the eigen tests' bar; uneven
batches that mix two-stage, one-stage and small blocks; rank-deficient panels
(block-sparse matrices with exact zeros, as H_N has them); the kept-row
back-transformation through NRG chains compared with the oracle (kept sets
bit for bit, energies 1e-10)."""
import numpy as np
import pytest

import pyoracle as po
import twochnrg_amd as t

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = t.Context(0)
    yield c
    c.close()


def _sym(rng, n, kind):
    if kind == "rand":
        a = rng.standard_normal((n, n))
        return (a + a.T) / 2
    if kind == "degen":  # repeated eigenvalues (multiplets)
        q, _ = np.linalg.qr(rng.standard_normal((n, n)))
        lam = np.repeat(rng.standard_normal((n + 2) // 3), 3)[:n]
        a = (q * lam) @ q.T
        return (a + a.T) / 2
    if kind == "diag":
        return np.diag(rng.standard_normal(n))
    if kind == "sparse" and n > 48:  # diagonal plus dense off-diagonal patches: rank-deficient panels
        a = np.diag(rng.standard_normal(n))
        for _ in range(8):
            i0, j0 = rng.integers(0, n - 24, 2)
            a[i0:i0 + 24, j0:j0 + 7] += rng.standard_normal((24, 7))
        return (a + a.T) / 2
    if kind == "sparse":
        return np.diag(rng.standard_normal(n))
    raise ValueError(kind)


def _check(a, w, z, tol=1e-12):
    n = a.shape[0]
    nrm = max(1.0, np.linalg.norm(a, 2))
    wl = np.linalg.eigvalsh(a)
    assert np.all(np.diff(w) >= 0)
    ew = np.max(np.abs(w - wl))
    assert ew <= tol * nrm * max(1, n) ** 0.5, ew
    r = np.max(np.abs(a @ z.T - z.T * w))
    assert r <= tol * nrm * n, r
    o = np.max(np.abs(z @ z.T - np.eye(n)))
    assert o <= tol * n, o
    return ew, r, o


@pytest.mark.parametrize("kind", ["rand", "degen", "diag", "sparse"])
def test_two_stage_uneven_batch(ctx, kind):
    rng = np.random.default_rng(21)
    sizes = [3, 40, 100, 211, 300, 517, 1000]
    blocks = [_sym(rng, n, kind) for n in sizes]
    ctx.set_two_stage(34)
    try:
        print("two-stage batch", kind, sizes, flush=True)
        W, Z, info = ctx.diag_blocks(blocks)
    finally:
        ctx.set_two_stage(0)
    assert not info.any()
    for a, w, z in zip(blocks, W, Z):
        ew, r, o = _check(a, w, z)
        print("  n=%d eig %.2e res %.2e orth %.2e" % (a.shape[0], ew, r, o))


def test_two_stage_matches_one_stage(ctx):
    rng = np.random.default_rng(5)
    sizes = [700, 1500, 2100]
    blocks = [_sym(rng, n, "rand") for n in sizes]
    print("one-stage batch", sizes, flush=True)
    W1, _, _ = ctx.diag_blocks(blocks)
    ctx.set_two_stage(512)
    try:
        print("two-stage batch", sizes, flush=True)
        W2, Z2, _ = ctx.diag_blocks(blocks)
    finally:
        ctx.set_two_stage(0)
    for a, w1, w2, z2 in zip(blocks, W1, W2, Z2):
        assert np.max(np.abs(w1 - w2)) <= 1e-12 * max(1.0, np.abs(w1).max()) * a.shape[0] ** 0.5
        _check(a, w2, z2)


def _compare(gs, os_, tol=1e-10):
    assert len(gs) == len(os_)
    for g, o in zip(gs, os_):
        assert np.array_equal(g["qn"], o["qn"])
        assert np.array_equal(g["kept"], o["kept"])
        err = np.max(np.abs(g["en"] - o["en"]) / np.maximum(np.abs(o["en"]), 1.0))
        assert err <= tol, (g["N"], err)


def test_two_stage_twochqs_chain(ctx):
    # blocks above 150 rows through the two-stage path; only the kept rows are
    # back-transformed (the chain's deferred eigenvectors)
    kw = dict(U=0.4, Gamma1=0.03, Gamma2=0.03, ed=-0.2, Lambda=3.0, nsitesmax=4, ncut=400)
    ctx.set_two_stage(150)
    try:
        print("two-stage twochqs", kw, flush=True)
        gs, _ = ctx.run_twochqs(**kw)
    finally:
        ctx.set_two_stage(0)
    _compare(gs, po.run_twochqs(**kw))


def test_two_stage_onechqs_chain(ctx):
    kw = dict(U=0.5, Gamma=0.02, ed=-0.25, Lambda=2.5, nsitesmax=10, ncut=400, calcdens=2)
    ctx.set_two_stage(100)
    try:
        print("two-stage onechqs", kw, flush=True)
        gs, _ = ctx.run_onechqs(**kw)
    finally:
        ctx.set_two_stage(0)
    _compare(gs, po.run_onechqs(**kw))
