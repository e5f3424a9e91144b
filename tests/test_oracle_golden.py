"""Pin the CPU oracle against the reference's golden run Runs/run1.

The log (6 significant digits) was produced before NRG_main multiplied the
first hopping by sqrt(A_Lambda): chi_{-1} = 0.121207 is the disc_scheme=1
("NRG_main -D 1") convention, so that is what is run here.
"""
import json
import os

import numpy as np
import pytest

import pyoracle as po

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "run1_golden.json")))


def _rel6(a, b):
    # value printed with 6 significant digits (cout default precision):
    # rounding error at most half a unit of the 6th digit
    return abs(a - b) <= 5.0001e-6 * abs(b) + 1e-9


@pytest.fixture(scope="module")
def steps():
    p = GOLD["params"]
    return po.run_onechqs(U=p["U"], Gamma=p["Gamma"], ed=p["ed"], Lambda=p["Lambda"],
                          nsitesmax=p["nsitesmax"], ncut=p["ncut"], calcdens=0, disc_scheme=1)


def test_levels_labels_and_counts(steps):
    assert [s["N"] for s in steps] == [-1, 0, 1, 2, 3]
    total = 0
    for st in steps:
        gl = GOLD["levels"][str(st["N"])]
        ge = np.array([x[0] for x in gl])
        gq = np.array([x[1] for x in gl])
        gk = np.array([x[2] for x in gl])
        sizes = st["bbe"][:, 1] - st["bbe"][:, 0] + 1
        qs = np.repeat(st["qn"], sizes, axis=0)
        assert np.array_equal(qs, gq), "block labels differ at N=%d" % st["N"]
        for a, b in zip(st["en"], ge):
            assert _rel6(a, b), (st["N"], a, b)
        if st["N"] >= 0:
            assert np.array_equal(st["kept"], gk)
            assert [st["nblocks"], len(st["en"])] == GOLD["counts"][str(st["N"])]
        total += len(ge)
    assert total == 636


def test_emin_ecut(steps):
    for st in steps[1:]:
        assert _rel6(st["emin"], GOLD["emin"][str(st["N"])])
        assert _rel6(st["ecut"], GOLD["ecut"][str(st["N"])])


def test_op_averages(steps):
    for st in steps:
        key = "%.6g" % st["T_avg"]
        g = GOLD["ops"][key]
        assert _rel6(st["opavg"][0], g["Ndot"])
        assert _rel6(st["opavg"][1], g["NdSq"])
        assert _rel6(st["opavg"][2], g["SdSq"])


def test_chain_closed_form():
    # printed chi_N, N = 0..2 (src/NRGchainClass.cpp:140-147)
    for n, v in [(0, 0.800641), (1, 0.872723), (2, 0.941601)]:
        assert abs(po.chin(n, 2.5) - v) < 1e-6
    lam, gam = 2.5, 0.0282691
    hlf = 0.5 * (1 + 1 / lam)
    chi_m1 = np.sqrt(2 * gam / np.pi) / (np.sqrt(lam) * hlf)
    assert abs(chi_m1 - GOLD["chi_m1_printed"]) < 1e-6


def test_hm1_closed_form(steps):
    U, ed, lam = 0.5, -0.25, 2.5
    f = lam * 0.5 * (1 + 1 / lam)
    ref = np.array([0.5 * U, ed + 0.5 * U, 2 * ed + 1.5 * U]) / f
    assert np.allclose(steps[0]["en"], ref, rtol=0, atol=1e-15)
