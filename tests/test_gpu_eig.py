"""Batched FP64 eigensolver (DiagBlock replacement) on the GPU against the
oracle's Householder+QL and LAPACK (numpy): eigenvalues within 1e-12 of the
block norm, residuals and orthogonality at the same level, ascending order,
row i = eigenvector i."""
import sys

import numpy as np
import pytest

import pyoracle as po
import twochnrg_amd as t

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=[2, 1], ids=["dc", "ql"])
def ctx(request):
    c = t.Context(0)
    c.set_eig_impl(request.param)
    c.impl = request.param
    yield c
    c.close()


def _sym(rng, n, kind):
    if kind == "rand":
        a = rng.standard_normal((n, n))
        return (a + a.T) / 2
    if kind == "degen":  # repeated eigenvalues (multiplets)
        q, _ = np.linalg.qr(rng.standard_normal((n, n)))
        lam = np.repeat(rng.standard_normal((n + 2) // 3), 3)[:n]
        return (q * lam) @ q.T
    if kind == "diag":
        return np.diag(rng.standard_normal(n))
    if kind == "tri":
        d = rng.standard_normal(n)
        e = rng.standard_normal(n - 1)
        return np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    raise ValueError(kind)


def _check(a, w, z, tol=1e-12):
    n = a.shape[0]
    nrm = max(1.0, np.linalg.norm(a, 2))
    wl = np.linalg.eigvalsh(a)
    assert np.all(np.diff(w) >= 0)
    assert np.max(np.abs(w - wl)) <= tol * nrm * max(1, n) ** 0.5
    r = a @ z.T - z.T * w
    assert np.max(np.abs(r)) <= tol * nrm * n
    assert np.max(np.abs(z @ z.T - np.eye(n))) <= tol * n


@pytest.mark.parametrize("kind", ["rand", "degen", "diag", "tri"])
def test_uneven_batch(ctx, kind):
    rng = np.random.default_rng(11)
    sizes = [1, 2, 3, 5, 17, 64, 65, 100, 211, 300]
    blocks = [_sym(rng, n, kind) for n in sizes]
    print("eig batch", kind, sizes, flush=True)
    W, Z, info = ctx.diag_blocks(blocks)
    assert not info.any()
    for a, w, z in zip(blocks, W, Z):
        _check(a, w, z)


def test_packed_upper_matches_oracle(ctx):
    rng = np.random.default_rng(3)
    sizes = [4, 33, 128]
    mats = [_sym(rng, n, "rand") for n in sizes]
    packed = [a[np.triu_indices(a.shape[0])] for a in mats]  # row-wise upper = ij2rUpTr order
    print("eig packed", sizes, flush=True)
    W, Z, info = ctx.diag_blocks(packed, packed=True)
    for a, p, w, z in zip(mats, packed, W, Z):
        wo, zo = po.eigh_packed(a.shape[0], p)
        assert np.max(np.abs(w - wo)) <= 1e-12 * max(1, np.abs(wo).max())
        # same eigenvectors up to sign (spectra here are non-degenerate)
        s = np.sign(np.sum(z * zo, axis=1))
        assert np.max(np.abs(z - s[:, None] * zo)) <= 1e-9
        _check(a, w, z)


def test_large_block(ctx):
    rng = np.random.default_rng(5)
    a = _sym(rng, 700, "degen")
    print("eig large n=700 impl", ctx.impl, flush=True)
    W, Z, info = ctx.diag_blocks([a])
    _check(a, W[0], Z[0], tol=1e-11)


def test_multislab_batch(ctx):
    # several blocks above the single-workgroup limit, uneven, with multiplets
    if ctx.impl != 2:
        pytest.skip("multi-slab path belongs to the batched solver")
    rng = np.random.default_rng(9)
    sizes = [257, 300, 513, 1500, 40, 3]
    blocks = [_sym(rng, n, "degen" if n % 2 else "rand") for n in sizes]
    print("eig multislab", sizes, flush=True)
    W, Z, info = ctx.diag_blocks(blocks)
    for a, w, z in zip(blocks, W, Z):
        _check(a, w, z, tol=1e-11)


def test_nrg_like_graded(ctx):
    # diagonal spanning decades plus weak couplings: the shape of H_N blocks
    rng = np.random.default_rng(4)
    n = 400
    d = np.sort(rng.random(n) * 20.0)
    a = np.diag(d) + 1e-3 * _sym(rng, n, "rand")
    a[0, 1] = a[1, 0] = 0.0
    print("eig graded n=400", flush=True)
    W, Z, info = ctx.diag_blocks([a])
    _check(a, W[0], Z[0], tol=1e-12)


@pytest.mark.parametrize("kind", ["rand", "degen", "tri"])
def test_resident_hand_off_boundaries(ctx, kind):
    # blocks above 256 rows leave their column chain at the panel boundary
    # 32 ceil((n - 208) / 32) and the register-resident kernel reduces the
    # trailing matrix from its lower triangle: n = 257 hands over 193 rows (the
    # 16 x 16 corner variant), 272 exactly 208, 273 and 305 177 (12 row slots),
    # 288 192, 465 177 after 9 panels; 208 is resident from the start, 209 and
    # 241 take the streamed small-block path
    if ctx.impl != 2:
        pytest.skip("the hand-off belongs to the batched solver")
    rng = np.random.default_rng(77)
    sizes = [257, 272, 273, 288, 305, 465, 241, 208, 209]
    blocks = [_sym(rng, n, kind) for n in sizes]
    print("eig hand-off", kind, sizes, flush=True)
    W, Z, info = ctx.diag_blocks(blocks)
    assert not info.any()
    for a, w, z in zip(blocks, W, Z):
        _check(a, w, z)


def test_reduction_streams_give_identical_results(ctx):
    # the large blocks split over 1..4 reduction streams: every block runs the
    # same kernels on the same data, so the results are bit-identical
    if ctx.impl != 2:
        pytest.skip("the reduction streams belong to the batched solver")
    rng = np.random.default_rng(21)
    sizes = [257, 700, 300, 1100, 520, 90, 400, 33]
    blocks = [_sym(rng, n, "degen" if n % 2 else "rand") for n in sizes]
    ref = None
    try:
        for ns in (1, 2, 3, 4):
            ctx.set_red_streams(ns)
            print("eig streams", ns, sizes, flush=True)
            W, Z, info = ctx.diag_blocks(blocks)
            assert not info.any()
            if ref is None:
                ref = (W, Z)
                for a, w, z in zip(blocks, W, Z):
                    _check(a, w, z, tol=1e-11)
            else:
                for w0, z0, w, z in zip(ref[0], ref[1], W, Z):
                    assert np.array_equal(w0, w) and np.array_equal(z0, z)
    finally:
        ctx.set_red_streams(3)
    with pytest.raises(t.NrgError):
        ctx.set_red_streams(5)
