"""Batched operator rotation Zi F Zj^T (UpdateMatrices' dgemm pair) on the GPU
against the oracle's dense FP64 products."""
import numpy as np
import pytest

import pyoracle as po
import twochnrg_amd as t

pytestmark = pytest.mark.gpu


def test_rotate_uneven():
    rng = np.random.default_rng(2)
    probs = []
    for ni, nj, ki, kj in [(1, 1, 1, 1), (5, 7, 3, 2), (64, 64, 64, 64), (130, 70, 33, 70), (300, 257, 100, 90)]:
        probs.append((rng.standard_normal((ki, ni)), rng.standard_normal((ni, nj)), rng.standard_normal((kj, nj))))
    print("rotate", [(p[0].shape, p[2].shape) for p in probs], flush=True)
    with t.Context(0) as c:
        out = c.rotate(probs)
    for (zi, f, zj), o in zip(probs, out):
        ref = po.rotate(zi, f, zj)
        assert np.max(np.abs(o - ref)) <= 1e-12 * max(1, np.abs(ref).max())
