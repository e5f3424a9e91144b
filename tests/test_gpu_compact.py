"""nrg_compact (CutStates' eigenvector compaction, src/CutStates.cpp:130-215,
CNRGarray::RemoveStatesFromBlock src/NRGarrayClass.cpp:1467-1548) through the
C-ABI on device arrays: the first keep[b] rows of every block, in block order,
blocks with nothing kept dropped -- bit for bit against the same slicing in
numpy, with keep counts from nrg_truncate on energy-ordered blocks."""
import numpy as np
import pytest

from hip_util import DeviceBuffers
import twochnrg_amd as t

pytestmark = pytest.mark.gpu


def _blocks(rng, sizes, gap=3):
    """Eigenvector-row blocks n x n at offsets with gaps between them"""
    z_off, parts, off = [], [], 0
    for n in sizes:
        off += gap
        z_off.append(off)
        parts.append((off, rng.standard_normal((n, n))))
        off += n * n
    Z = np.full(off + gap, np.nan)
    for o, a in parts:
        Z[o:o + a.size] = a.ravel()
    return Z, np.array(z_off, dtype=np.int64), [a for _, a in parts]


def _run(ctx, sizes, keep, Z, z_off):
    hip = DeviceBuffers()
    off, tot = ctx.compact(sizes, z_off, keep)   # sizes only
    dZ, dC = hip.upload(Z), hip.upload(np.full(max(tot, 1), np.nan))
    try:
        off2, tot2 = ctx.compact(sizes, z_off, keep, dZ.value, dC.value)
        assert np.array_equal(off, off2) and tot == tot2
        return off, hip.download(dC, tot)
    finally:
        hip.free(dZ, dC)


def test_compact_matches_slicing():
    rng = np.random.default_rng(3)
    sizes = [5, 1, 17, 0, 40, 3, 64]
    keep = [2, 0, 17, 0, 13, 1, 64]
    Z, z_off, mats = _blocks(rng, sizes)
    with t.Context(0) as ctx:
        print("compact", sizes, keep, flush=True)
        off, out = _run(ctx, sizes, keep, Z, z_off)
    want = np.concatenate([a[:k].ravel() for a, k in zip(mats, keep) if k > 0])
    assert np.array_equal(out, want)
    exp_off, o = [], 0
    for n, k in zip(sizes, keep):
        exp_off.append(o if k > 0 else -1)
        o += k * n
    assert list(off) == exp_off


def test_compact_with_truncation():
    # kept sets from nrg_truncate on energy-ordered blocks are prefixes of each
    # block; their sizes are the keep counts
    rng = np.random.default_rng(8)
    sizes = [30, 12, 7, 25, 1, 50]
    en = [np.sort(rng.random(n) * (1 + b)) for b, n in enumerate(sizes)]
    flat = np.concatenate(en)
    _, kept, _, nkept = t.truncate(flat, 40, 1e-8)
    assert nkept >= 40
    keep, s = [], 0
    for n in sizes:
        k = int(kept[s:s + n].sum())
        assert np.all(kept[s:s + k]) and not np.any(kept[s + k:s + n])   # a prefix
        keep.append(k)
        s += n
    Z, z_off, mats = _blocks(rng, sizes)
    with t.Context(0) as ctx:
        print("compact truncated", sizes, keep, flush=True)
        _, out = _run(ctx, sizes, keep, Z, z_off)
    assert np.array_equal(out, np.concatenate([a[:k].ravel() for a, k in zip(mats, keep) if k > 0]))


def test_compact_refuses_bad_sizes():
    with t.Context(0) as ctx:
        with pytest.raises(t.NrgError):
            ctx.compact([4, 3], [0, 16], [5, 1])   # keep > n
        with pytest.raises(t.NrgError):
            ctx.compact([4], [0], [-1])
        off, tot = ctx.compact([], [], [])
        assert tot == 0
