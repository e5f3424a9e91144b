"""hipGraph mode: every eigensolve's launch sequences captured into a graph and
launched once.  Each chain below replays many graphs (one or two per step, with
the chain's own eager work and a host synchronisation between them); every
step's spectrum and Kept set must equal the eager run's bit for bit, and both
the oracle's (energies 1e-10)."""
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


def _same(ga, gb):
    assert len(ga) == len(gb)
    for a, b in zip(ga, gb):
        assert np.array_equal(a["qn"], b["qn"]) and np.array_equal(a["kept"], b["kept"])
        assert np.array_equal(a["en"], b["en"]), a["N"]


def _oracle(gs, os_):
    for g, o in zip(gs, os_):
        assert np.array_equal(g["kept"], o["kept"])
        err = np.max(np.abs(g["en"] - o["en"]) / np.maximum(np.abs(o["en"]), 1.0))
        assert err <= 1e-10, (g["N"], err)


def test_graph_onechqs_chain(ctx):
    kw = dict(U=0.5, Gamma=0.02, ed=-0.25, Lambda=2.5, nsitesmax=12, ncut=400, calcdens=2)
    print("eager onechqs", kw, flush=True)
    ge, _ = ctx.run_onechqs(**kw)
    ctx.set_graph(True)
    try:
        print("graph onechqs", flush=True)
        gg, _ = ctx.run_onechqs(**kw)
        ms, nodes, graphs = ctx.graph_stats()
    finally:
        ctx.set_graph(False)
    print("graphs %d, nodes %d, host ms %.2f" % (graphs, nodes, ms))
    assert graphs >= len(gg)
    _same(ge, gg)
    _oracle(gg, po.run_onechqs(**kw))


def test_graph_twochqs_chain(ctx):
    # blocks above 256 rows: the blocked reduction's per-column launches on
    # three auxiliary streams inside the capture
    kw = dict(U=0.4, Gamma1=0.03, Gamma2=0.03, ed=-0.2, Lambda=3.0, nsitesmax=4, ncut=400)
    print("eager twochqs", kw, flush=True)
    ge, _ = ctx.run_twochqs(**kw)
    ctx.set_graph(True)
    try:
        print("graph twochqs", flush=True)
        gg, _ = ctx.run_twochqs(**kw)
        ms, nodes, graphs = ctx.graph_stats()
    finally:
        ctx.set_graph(False)
    print("graphs %d, nodes %d, host ms %.2f" % (graphs, nodes, ms))
    _same(ge, gg)
    _oracle(gg, po.run_twochqs(**kw))


def test_graph_diag_blocks_repeated(ctx):
    # the same uneven batch through eager and graph launches, alternately, three times
    rng = np.random.default_rng(4)
    sizes = [5, 64, 211, 300, 700]
    blocks = []
    for n in sizes:
        a = rng.standard_normal((n, n))
        blocks.append((a + a.T) / 2)
    ref, refz, _ = ctx.diag_blocks(blocks)
    for rep in range(3):
        ctx.set_graph(True)
        try:
            print("graph diag_blocks rep", rep, flush=True)
            W, Z, _ = ctx.diag_blocks(blocks)
        finally:
            ctx.set_graph(False)
        for w, w0, z, z0 in zip(W, ref, Z, refz):
            assert np.array_equal(w, w0) and np.array_equal(z, z0)
        W2, _, _ = ctx.diag_blocks(blocks)   # eager work between the replays
        for w, w0 in zip(W2, ref):
            assert np.array_equal(w, w0)
