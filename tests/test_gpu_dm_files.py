"""DM-NRG over shells the caller owns and the reference's .bin hand-off files.

nrg_save_arrays runs a chain and writes, per shell N, the Abasis_/Acut_/Mat0_
files CNRGCodeHandler::SaveArrays writes (src/NRGcodehandlerClass.cpp:887-933);
nrg_dm_create_from_files reads them back as DM_NRG does
(src/DM_NRG/DM_NRG.cpp:177-298, :312-434) and hands them to nrg_dm_create_from.
The spectral sums, the window sums, CalcNorm and rho_N must equal those of
nrg_dm_create on the same chain to 1e-12.  The byte layout is checked by a
parser written here from the reference's SaveBin field order
(CNRGarray::SaveBin src/NRGarrayClass.cpp:778-869, CNRGbasisarray::SaveBin
:1947-2000, CNRGmatrix::SaveBin src/NRGmatrixClass.cpp:957-1010): int32 counts,
doubles, one-byte bools; the parsed energies and Kept flags must be the chain's
own records, and the rhoDM_ files the handle's rho_N."""
import struct

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ONE = dict(U=0.5, Gamma=0.02, ed=-0.25, Lambda=2.5, nsitesmax=14, ncut=120)
CASES = [
    dict(ONE, Mtemp=10, betabar=1.0, method=0),                                      # DM-NRG, UpdateBefCut 0
    dict(ONE, Mtemp=10, betabar=1.0, method=2, update_bef_cut=1, bbroad=0.6),        # FDM
    dict(ONE, Mtemp=9, betabar=0.8, method=2, update_bef_cut=1, bbroad=0.6, fdm_fix=63),
    dict(U=0.5, Gamma=0.03, Gamma2=0.01, ed=-0.25, Lambda=2.5, model=1, nsitesmax=4, ncut=200, Mtemp=3,
         betabar=0.779, method=2, bbroad=0.6),                                      # TwoChQSz FDM
]


@pytest.fixture(scope="module")
def ctx():
    import twochnrg_amd as t
    c = t.Context(0)
    yield c
    c.close()


class _Reader:
    """The SaveBin byte stream, field by field"""

    def __init__(self, path):
        self.b = open(path, "rb").read()
        self.o = 0

    def _take(self, fmt, n):
        v = struct.unpack_from("<%d%s" % (n, fmt), self.b, self.o)
        self.o += struct.calcsize("<%d%s" % (n, fmt))
        return v

    def i32(self):
        return self._take("i", 1)[0]

    def vec(self, fmt):
        """int32 count then the elements; "C": complex<double>, two doubles each"""
        n = self.i32()
        if fmt == "C":
            return np.array(self._take("d", 2 * n))
        return np.array(self._take(fmt, n))

    def array(self):
        a = dict(nshell=self.i32(), nq=self.i32(), qn=self.vec("d"), en=self.vec("d"), evec=self.vec("d"),
                 cevec=self.vec("C"), bbe=self.vec("i"), idegen=self.vec("i"), kept=self.vec("?"))
        a["totalS"] = self._take("?", 1)[0]
        a["sqn"] = self.vec("i")
        return a

    def basis(self):
        a = self.array()
        a.update(itype=self.vec("i"), stcf=self.vec("i"), bbe_bc=self.vec("i"), bbe_ev=self.vec("i"))
        a["nq_stcf"] = self.i32()
        a["stcf_qn"] = self.vec("d")
        return a

    def matrix(self):
        a = self.array()
        a.update(matel=self.vec("d"), matel_cplx=self.vec("C"), map=self.vec("i"), mbe=self.vec("i"))
        a["need_old"], a["upper"], a["is_complex"] = self._take("?", 3)
        a["wigner_eckart_l"] = self._take("d", 1)[0]
        return a

    def done(self):
        return self.o == len(self.b)


def _rel(got, ref):
    got, ref = np.asarray(got), np.asarray(ref)
    return np.abs(got - ref) / np.maximum(np.abs(ref), 1e-12 * max(np.abs(ref).max(), 1e-300))


@pytest.mark.parametrize("i", range(len(CASES)))
def test_files_round_trip(ctx, tmp_path, i):
    case = CASES[i]
    print("dm files", case, flush=True)
    nsh = ctx.save_arrays(tmp_path, "t%d" % i, **case)
    qsz = case.get("model", 0) == 1
    assert nsh == case["nsitesmax"] + (1 if qsz else 0)
    with ctx.dm(**case) as d0, ctx.dm_from_files(tmp_path, "t%d" % i, nsh, **case) as d1:
        assert (d1.nshell_max, d1.spec_nshell_max) == (d0.nshell_max, d0.spec_nshell_max)
        assert d1.temp == d0.temp
        sh, w = [], []
        for N in range(d0.spec_nshell_max):
            for x in (-1.0, -0.2, -1e-3, 0.0, 1e-3, 0.05, 0.7):
                sh.append(N)
                w.append(x)
        assert _rel(d1.spec(sh, w), d0.spec(sh, w)).max() <= 1e-12
        for N in range(d0.nshell_max + 1):
            a = _Reader(tmp_path / ("Acut_t%d_N%d.bin" % (i, N))).basis()
            sizes = [int(a["bbe"][2 * b + 1] - a["bbe"][2 * b] + 1) for b in range(len(a["bbe"]) // 2)]
            b0, t0 = d0.rho(N, sizes)
            b1, t1 = d1.rho(N, sizes)
            assert abs(t1 - t0) <= 1e-12 * abs(t0)
            for x, y in zip(b0, b1):
                assert np.allclose(y, x, rtol=0, atol=1e-12 * max(1e-300, np.abs(x).max()))
        if case["method"] != 0:
            ws = np.array([-0.9, -3e-3, 2e-5, 1e-3, 0.4])
            assert _rel(d1.spec_omega(ws), d0.spec_omega(ws)).max() <= 1e-12
            assert abs(d1.norm() - d0.norm()) <= 1e-12 * abs(d0.norm())
            if case.get("fdm_fix") == 63:
                assert abs(d1.norm() - 1.0) <= 1e-10
        w0, r0 = d0.spectrum()
        w1, r1 = d1.spectrum()
        assert np.array_equal(w0, w1) and _rel(r1, r0).max() <= 1e-12


def test_file_layout_is_the_references(ctx, tmp_path):
    """The bytes of Acut_/Abasis_/Mat0_/rhoDM_ parsed in SaveBin's field order:
    energies, Kept flags and block structure are the chain's records, the
    eigenvector rows are unit vectors, rhoDM_ holds the handle's rho_N."""
    case = CASES[1]
    print("dm file layout", case, flush=True)
    nsh = ctx.save_arrays(tmp_path, "lay", **case)
    steps, _ = ctx.run_onechqs(U=case["U"], Gamma=case["Gamma"], ed=case["ed"], Lambda=case["Lambda"],
                               nsitesmax=case["nsitesmax"], ncut=case["ncut"], calcdens=1)
    recs = {st["N"]: st for st in steps}
    for N in range(nsh):
        r = _Reader(tmp_path / ("Acut_lay_N%d.bin" % N))
        a = r.basis()
        assert r.done()
        assert a["nshell"] == N and a["nq"] == 2 and a["totalS"] and list(a["sqn"]) == [1]
        st = recs[N]
        # UpdateBefCut 1: Acut keeps every state of the step with its Kept flags
        assert np.array_equal(a["en"], st["en"]) and np.array_equal(a["kept"], st["kept"])
        assert np.array_equal(a["qn"], st["qn"].ravel()) and np.array_equal(a["bbe"], st["bbe"].ravel())
        for b in range(len(a["bbe"]) // 2):
            k = a["bbe"][2 * b + 1] - a["bbe"][2 * b] + 1
            n = a["bbe_bc"][2 * b + 1] - a["bbe_bc"][2 * b] + 1
            Z = a["evec"][a["bbe_ev"][2 * b]:a["bbe_ev"][2 * b + 1] + 1].reshape(k, n)
            assert np.allclose(Z @ Z.T, np.eye(k), atol=1e-12)
        rb = _Reader(tmp_path / ("Abasis_lay_N%d.bin" % N))
        ab = rb.basis()
        assert rb.done() and len(ab["itype"]) == len(ab["stcf"]) == ab["bbe"][-1] + 1
        rm = _Reader(tmp_path / ("Mat0_lay_N%d.bin" % N))
        m = rm.matrix()
        assert rm.done() and len(m["map"]) == len(m["mbe"]) and m["mbe"][-1] + 1 == len(m["matel"])
    with ctx.dm(**case) as d:
        d.save_rho(tmp_path, "lay")
        for N in range(d.nshell_max + 1):
            rr = _Reader(tmp_path / ("rhoDM_lay_N%d.bin" % N))
            m = rr.matrix()
            assert rr.done() and m["upper"]
            sizes = [int(m["bbe"][2 * b + 1] - m["bbe"][2 * b] + 1) for b in range(len(m["bbe"]) // 2)]
            blocks, _ = d.rho(N, sizes)
            for q, R in enumerate(blocks):
                k = sizes[q]
                got = np.zeros((k, k))
                got[np.triu_indices(k)] = m["matel"][m["mbe"][2 * q]:m["mbe"][2 * q + 1] + 1]
                assert np.array_equal(got, np.triu(R))


def test_create_from_rejects_inconsistent_shells(ctx, tmp_path):
    """A shell whose StCameFrom points outside the previous shell, or whose
    operator blocks do not fit the block sizes, is refused before any device work."""
    import twochnrg_amd as t
    case = CASES[0]
    print("dm create_from refusals", flush=True)
    nsh = ctx.save_arrays(tmp_path, "bad", **case)
    shells = []
    for N in range(nsh):
        a = _Reader(tmp_path / ("Acut_bad_N%d.bin" % N)).basis()
        b = _Reader(tmp_path / ("Abasis_bad_N%d.bin" % N)).basis()
        m = _Reader(tmp_path / ("Mat0_bad_N%d.bin" % N)).matrix()
        shells.append(dict(qn=a["qn"], bbe=a["bbe"], bbe_bc=a["bbe_bc"], bbe_ev=a["bbe_ev"], en=a["en"],
                           kept=a["kept"].astype(np.uint8), evec=a["evec"], basis_qn=b["qn"], basis_bbe=b["bbe"],
                           basis_itype=b["itype"], basis_stcf=b["stcf"], op_map=m["map"], op_mbe=m["mbe"],
                           op=m["matel"]))
    with ctx.dm(**case) as d0, ctx.dm_from_shells(shells, **case) as d1:
        w = [0.1, -0.3]
        assert _rel(d1.spec([3, 5], w), d0.spec([3, 5], w)).max() <= 1e-12
    bad = [dict(s) for s in shells]
    bad[4] = dict(bad[4], basis_stcf=np.array(bad[4]["basis_stcf"]) + len(shells[3]["en"]))
    with pytest.raises(t.NrgError):
        ctx.dm_from_shells(bad, **case)
    bad = [dict(s) for s in shells]
    mbe = np.array(bad[2]["op_mbe"]).copy()
    mbe[1] += 1
    bad[2] = dict(bad[2], op_mbe=mbe)
    with pytest.raises(t.NrgError):
        ctx.dm_from_shells(bad, **case)
