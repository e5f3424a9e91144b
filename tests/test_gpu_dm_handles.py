"""DM-NRG handles on one context own their device buffers: two handles with
different chains live side by side and each keeps returning its own results
(bit-identical to its first answers), a rejected create leaves the live
handles untouched, and a destroyed handle's buffers are gone from the context
while the other's stay."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FDM = dict(U=0.5, Gamma=0.02, ed=-0.25, Lambda=2.5, nsitesmax=14, ncut=120, Mtemp=10, betabar=1.0, method=2,
           update_bef_cut=1)
DMNRG = dict(U=0.6, Gamma=0.04, ed=-0.2, Lambda=2.5, nsitesmax=12, ncut=150, Mtemp=10, betabar=0.8, method=0)
QSZ = dict(U=0.5, Gamma=0.03, Gamma2=0.01, ed=-0.25, Lambda=2.5, model=1, nsitesmax=4, ncut=200, Mtemp=3,
           betabar=0.779, method=2, bbroad=0.6)


@pytest.fixture(scope="module")
def ctx():
    import twochnrg_amd as t
    c = t.Context(0)
    yield c
    c.close()


def test_handles_side_by_side(ctx):
    import twochnrg_amd as t
    print("dm handles side by side", flush=True)
    ws = np.array([-0.3, -1e-3, 2e-4, 0.05, 0.8])
    a = ctx.dm(**FDM)
    ra = a.spec_omega(ws)
    b = ctx.dm(**DMNRG)
    sh = [0, 3, 7]
    rb = b.spec(sh, [0.1, -0.2, 0.3])
    q = ctx.dm(**QSZ)
    rq = q.spec_omega(ws)
    with pytest.raises(t.NrgError):
        ctx.dm(**dict(QSZ, model=2))   # no such chain
    # every live handle still answers as before, to the bit
    assert np.array_equal(a.spec_omega(ws), ra)
    assert np.array_equal(b.spec(sh, [0.1, -0.2, 0.3]), rb)
    assert np.array_equal(q.spec_omega(ws), rq)
    mq = q.memory()
    assert mq["before_sweep"] > 0 and mq["now"] > 0
    held_all = a.memory()["now"]
    q.close()
    b.close()
    assert np.array_equal(a.spec_omega(ws), ra)
    assert a.memory()["now"] < held_all   # the closed handles' buffers left the context
    a.close()


def test_context_closed_before_its_handles():
    """A context destroyed while DM handles are open: through the binding the
    handles are closed first; through the C-ABI the handles are detached, so
    their calls fail with NRG_ERR_STATE and nrg_dm_destroy frees nothing twice."""
    import ctypes
    import twochnrg_amd as t
    print("dm handles outliving their context", flush=True)
    c = t.Context(0)
    d = c.dm(**QSZ)
    c.close()
    assert d.h is None and c.h is None
    # C-ABI order: context first, handle second
    L = t.lib()
    ch, dh = ctypes.c_void_p(), ctypes.c_void_p()
    assert L.nrg_ctx_create(0, ctypes.byref(ch)) == 0
    p = t.DmParams(**{k: v for k, v in QSZ.items()})
    assert L.nrg_dm_create(ch, ctypes.byref(p), ctypes.byref(dh)) == 0
    L.nrg_ctx_destroy(ch)
    out = np.zeros(1)
    rc = L.nrg_dm_norm(dh, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc == -4, rc   # NRG_ERR_STATE
    assert b"context" in L.nrg_last_error()
    L.nrg_dm_destroy(dh)
