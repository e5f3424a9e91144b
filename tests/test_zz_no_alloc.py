"""nrg_rotate_ops_noalloc (the rotations of nrg_rotate_ops on caller-owned
device buffers, its scratch in a fixed per-device array) against
nrg_rotate_ops on the same inputs, through the C-ABI: bit for bit, on two calls
in a row, the second with a smaller input (the scratch keeps the first call's
contents), plus a request larger than the scratch area, which fails before any
work."""
import ctypes

import numpy as np
import pytest

from hip_util import DeviceBuffers

pytestmark = pytest.mark.gpu


def _problems(rng, shapes):
    """NRG-like block pairs: (ki, ni, kj, nj) with kept rows of eigenvectors"""
    out = []
    for ki, ni, kj, nj in shapes:
        out.append((rng.standard_normal((ki, ni)), rng.standard_normal((ni, nj)), rng.standard_normal((kj, nj))))
    return out


def _pack(t, problems):
    zs, fs, descs = [], [], []
    zoff = foff = ooff = 0
    for zi, f, zj in problems:
        ki, ni = zi.shape
        kj, nj = zj.shape
        descs.append(t.PairDesc(ni, nj, ki, kj, zoff, zoff + zi.size, foff, ooff + 3))   # 3 doubles gap first
        zs += [zi.ravel(), zj.ravel()]
        fs.append(f.ravel())
        zoff += zi.size + zj.size
        foff += f.size
        ooff += ki * kj
    Z = np.concatenate(zs)
    F = np.concatenate(fs)
    return descs, Z, F, ooff + 3 + 5   # 5 doubles no pair writes at the end


def _host(t, ctx, descs, Z, F, outlen):
    arr = (t.PairDesc * len(descs))(*descs)
    out = np.full(outlen, np.nan)
    t._check(t.lib().nrg_rotate_ops(ctx.h, len(descs), arr, t._ptr(Z), Z.size, t._ptr(F), F.size, t._ptr(out),
                                    outlen, 0))
    return out


_Hip = DeviceBuffers


def _noalloc(t, descs, Z, F, outlen):
    hip = _Hip()
    arr = (t.PairDesc * len(descs))(*descs)
    dZ, dF, dO = hip.upload(Z), hip.upload(F), hip.upload(np.full(outlen, np.nan))
    try:
        t._check(t.lib().nrg_rotate_ops_noalloc(None, len(descs), arr, dZ, Z.size, dF, F.size, dO, outlen, 0))
        return hip.download(dO, outlen)
    finally:
        hip.free(dZ, dF, dO)


def test_noalloc_matches_rotate_ops_twice():
    import twochnrg_amd as t
    rng = np.random.default_rng(11)
    big = [(40, 300, 70, 260), (0, 120, 30, 90), (130, 130, 128, 129), (64, 200, 1, 77), (200, 512, 190, 480)]
    small = [(12, 30, 7, 26), (5, 9, 5, 9)]
    with t.Context(0) as ctx:
        for shapes in (big, small):
            print("noalloc rotations", shapes, flush=True)
            descs, Z, F, outlen = _pack(t, _problems(rng, shapes))
            h = _host(t, ctx, descs, Z, F, outlen)
            d = _noalloc(t, descs, Z, F, outlen)
            assert np.array_equal(h, d)
            # the products themselves, and zeros where no pair writes
            assert np.all(d[:3] == 0.0) and np.all(d[-5:] == 0.0)
            for ds in descs:
                zi = Z[ds.zi_off:ds.zi_off + ds.ki * ds.ni].reshape(ds.ki, ds.ni)
                zj = Z[ds.zj_off:ds.zj_off + ds.kj * ds.nj].reshape(ds.kj, ds.nj)
                f = F[ds.f_off:ds.f_off + ds.ni * ds.nj].reshape(ds.ni, ds.nj)
                ref = zi @ f @ zj.T
                got = d[ds.out_off:ds.out_off + ds.ki * ds.kj].reshape(ds.ki, ds.kj)
                scale = max(1.0, np.abs(ref).max()) if ref.size else 1.0
                assert np.allclose(got, ref, rtol=1e-12, atol=1e-12 * scale)


def test_noalloc_rejects_more_than_its_scratch():
    import twochnrg_amd as t
    print("noalloc oversize request", flush=True)
    # one pair whose intermediate product (kj * ni doubles) exceeds the scratch area;
    # refused by the host check before any device access
    n = 3000
    d = t.PairDesc(n, n, n, n, 0, 0, 0, 0)
    arr = (t.PairDesc * 1)(d)
    hip = _Hip()
    buf = hip.upload(np.zeros(8))
    try:
        L = t.lib()
        rc = L.nrg_rotate_ops_noalloc(None, 1, arr, buf, n * n, buf, n * n, buf, n * n, 0)
        assert rc == -1, rc
        assert b"scratch" in L.nrg_last_error()
        assert np.all(hip.download(buf, 8) == 0.0)
    finally:
        hip.free(buf)
