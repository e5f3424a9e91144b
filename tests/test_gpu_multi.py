"""Lockstep batching of independent OneChQS instances (nrg_run_onechqs_multi):
every instance's records must equal its own single-instance run bit for bit in
structure and to 1e-12 in energies (the blocks go through the same kernels,
only batched with other instances' blocks), and one instance is also checked
against the CPU oracle."""
import json
import os

import numpy as np
import pytest

import pyoracle as po

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def ctx():
    import twochnrg_amd as t
    c = t.Context(0)
    yield c
    c.close()


def _instances(k, ncut_scale=0.2):
    d = json.load(open(os.path.join(HERE, "golden", "c2_instances.json")))
    out = []
    for x in d["instances"][:k]:
        y = dict(x)
        y["ncut"] = max(40, int(x["ncut"] * ncut_scale))   # small versions of the seeded instances
        out.append(y)
    return out


def _same(a, b, tol):
    assert len(a) == len(b)
    for g, o in zip(a, b):
        assert g["N"] == o["N"]
        assert np.array_equal(g["qn"], o["qn"]) and np.array_equal(g["bbe"], o["bbe"])
        assert np.array_equal(g["kept"], o["kept"])
        assert np.array_equal(g["cut_qn"], o["cut_qn"]) and np.array_equal(g["cut_bbe"], o["cut_bbe"])
        err = np.max(np.abs(g["en"] - o["en"]) / np.maximum(np.abs(o["en"]), 1.0))
        assert err <= tol, (g["N"], err)
        assert np.allclose(g["opavg"], o["opavg"], rtol=1e-8, atol=1e-14)


def test_multi_equals_single(ctx):
    inst = _instances(6)
    nsites = 10
    print("multi", len(inst), "instances", nsites, "sites", flush=True)
    multi, _ = ctx.run_onechqs_multi(inst, nsitesmax=nsites, calcdens=2)
    for i, d in enumerate(inst):
        single, _ = ctx.run_onechqs(U=d["U"], Gamma=d["Gamma"], ed=d["ed"], Lambda=d["Lambda"], nsitesmax=nsites,
                                    ncut=d["ncut"], calcdens=2)
        _same(multi[i], single, 1e-12)


def test_multi_vs_oracle(ctx):
    inst = _instances(4)
    nsites = 8
    multi, _ = ctx.run_onechqs_multi(inst, nsitesmax=nsites, calcdens=2)
    d = inst[2]
    o = po.run_onechqs(U=d["U"], Gamma=d["Gamma"], ed=d["ed"], Lambda=d["Lambda"], nsitesmax=nsites, ncut=d["ncut"],
                       calcdens=2)
    _same(multi[2], o, 1e-10)
