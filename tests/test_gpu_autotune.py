"""Autotuner of the eigensolver's launch configuration (nrg_autotune_eig):
every candidate is timed and reported, the fastest is applied to the
context, and eigensolves under the chosen configuration stay correct
(LAPACK eigenvalues, residuals, orthogonality)."""
import numpy as np
import pytest

import twochnrg_amd as t

pytestmark = pytest.mark.gpu


def test_autotune_applies_fastest_and_stays_correct():
    with t.Context(0) as ctx:
        sizes = [7, 64, 300, 700, 1200, 2100]
        print("autotune", sizes, flush=True)
        best, rows = ctx.autotune(sizes, krow_frac=0.1, reps=1)
        print(best, rows)
        # streams 1..3 (4 blocks above 256 rows), two-stage 2048 and 1024, graph on
        assert len(rows) == 3 + 2 + 1
        assert all(r["ms"] > 0 for r in rows)
        assert {r["red_streams"] for r in rows[:3]} == {1, 2, 3}
        assert best in rows
        # a candidate displaces the running choice only when > 2 % faster
        assert all(r["ms"] >= 0.98 * best["ms"] for r in rows)
        rng = np.random.default_rng(8)
        blocks = []
        for n in sizes:
            a = rng.standard_normal((n, n))
            blocks.append((a + a.T) / 2)
        print("diag_blocks under", best, flush=True)
        W, Z, _ = ctx.diag_blocks(blocks)
        for a, w, z in zip(blocks, W, Z):
            n = a.shape[0]
            nrm = max(1.0, np.linalg.norm(a, 2))
            assert np.max(np.abs(w - np.linalg.eigvalsh(a))) <= 1e-12 * nrm * n ** 0.5
            assert np.max(np.abs(a @ z.T - z.T * w)) <= 1e-12 * nrm * n
            assert np.max(np.abs(z @ z.T - np.eye(n))) <= 1e-12 * n


def test_autotune_rejects_bad_arguments():
    with t.Context(0) as ctx:
        with pytest.raises(t.NrgError):
            ctx.autotune([10, 20], krow_frac=1.5)
