"""Correctness at the sizes the configurations are timed at.

Two-channel Anderson (TwoChQS, Lambda = 3, Gamma1 = Gamma2 = 0.03, U = 0.4,
ed = -0.2) keeping 3000 multiplets, through its first two truncated steps
(N = 3, 4: ~45k multiplets in ~53 (Q, S) blocks of up to ~4400 rows), with the
device-assembled H_N blocks recorded:

* every block's eigenvalues against LAPACK (numpy eigvalsh of the recorded H,
  minus the step's Emin) within 1e-10 relative to max(|E|, 1);
* eigen-residual |H z - E z| and orthogonality |Z Z^T - 1| of every block above
  4000 rows, within 1e-10 of max|H| and 1e-10;
* the (Q, S) labels and block sizes against the SU(2) coupling series of the
  previous step's kept multiplets with the 16-state two-channel site
  (old multiplet S_o x site multiplet S_s -> |S_o - S_s| .. S_o + S_s), and the
  state count sum (2S+1) n = 16 sum_kept (2S+1);
* the Kept set rebuilt from the LAPACK eigenvalues with the reference's Ecut
  rule (CNRGarray::Ecut src/NRGarrayClass.cpp:661-691, SetKept :461-491) is the
  device's bit for bit, apart from states within 1e-9 of the cutoff (reported).

The 64-instance batched run (OneChQS, Ncut 803-2000 from the seeded table) for
its first steps at full size: every instance's records equal its own single
run's (kept sets bit for bit, energies 1e-12).

The single-channel FDM at Ncut 400, T = 1e-8 and the two-channel FDM: the
exact full density matrix has CalcNorm 1 (1e-10), particle-hole symmetry
A(w) = A(-w), and the Friedel sum rule pi Gamma A(0) = 1 (with the A_Lambda
factor) at w << T << T_K: within 6 % on the single-channel chain (1.05: the
log-Gaussian broadening, b = 0.6, of a discrete spectrum), within 10 % on the
two-channel chain keeping 1500 states of its 16-state sites (0.94).
"""
import json
import math
import os

import numpy as np
import pytest

import pyoracle as po
import twochnrg_amd as t

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))

# the two-channel site: (Q, S) multiplets and their multiplicity
# (TwoChQS_SetSingleSite: 16 states in 10 (Q, S, Sz) blocks)
SITE = [(-2.0, 0.0, 1), (-1.0, 0.5, 2), (0.0, 0.0, 3), (0.0, 1.0, 1), (1.0, 0.5, 2), (2.0, 0.0, 1)]


@pytest.fixture(scope="module")
def ctx():
    c = t.Context(0)
    yield c
    c.close()


def _unpack(hp, n):
    """row-wise packed upper (ij2rUpTr) -> full symmetric"""
    A = np.zeros((n, n))
    A[np.triu_indices(n)] = hp
    return A + np.triu(A, 1).T


def _expected_blocks(prev):
    """(Q, S) -> multiplet count of the next basis from the previous step's cut blocks"""
    cnt = {}
    for (Q, S), (b0, b1) in zip(prev["cut_qn"][:, :2], prev["cut_bbe"]):
        k = int(b1 - b0 + 1)
        for Qs, Ss, m in SITE:
            Sn = abs(S - Ss)
            while Sn <= S + Ss + 1e-9:
                key = (round(Q + Qs, 6), round(Sn, 6))
                cnt[key] = cnt.get(key, 0) + k * m
                Sn += 1.0
    return cnt


def test_twochqs_ncut3000_first_truncated_steps(ctx):
    kw = dict(U=0.4, Gamma1=0.03, Gamma2=0.03, ed=-0.2, Lambda=3.0, ncut=3000)
    print("fullsize twochqs", kw, flush=True)
    recs = {}

    def on_step(st):
        d = t.step_to_dict(st)
        if d["N"] >= 2:
            recs[d["N"]] = d
        print("  step N=%d recorded, %d blocks, %d multiplets" % (d["N"], d["nblocks"], len(d["en"])), flush=True)

    ctx.run_twochqs(nsitesmax=4, record_blocks=True, keep_steps=False, on_step=on_step, **kw)
    assert sorted(recs) == [2, 3, 4]
    worst = {"eig": 0.0, "res": 0.0, "orth": 0.0}
    for N in (3, 4):
        r = recs[N]
        sizes = (r["bbe"][:, 1] - r["bbe"][:, 0] + 1).astype(int)
        assert sizes.max() > 4000, sizes.max()
        # labels and sizes from the SU(2) coupling series of the previous cut
        exp = _expected_blocks(recs[N - 1])
        got = {(round(q, 6), round(s, 6)): int(n) for (q, s), n in zip(r["qn"][:, :2], sizes)}
        assert got == exp, N
        prev = recs[N - 1]
        kept_mult = sum((2 * S + 1) * (b1 - b0 + 1) for (Q, S), (b0, b1) in zip(prev["cut_qn"][:, :2], prev["cut_bbe"]))
        assert abs(sum((2 * s + 1) * n for (q, s), n in got.items()) - 16 * kept_mult) < 1e-9
        # eigenvalues per block against LAPACK
        hoff = np.concatenate([[0], np.cumsum(sizes * (sizes + 1) // 2)])
        zoff = np.concatenate([[0], np.cumsum(sizes * sizes)])
        lap = np.empty(len(r["en"]))
        for b, n in enumerate(sizes):
            H = _unpack(r["hpacked"][hoff[b]:hoff[b + 1]], n)
            w = np.linalg.eigvalsh(H)
            i0 = r["bbe"][b, 0]
            lap[i0:i0 + n] = w - r["emin"]
            err = np.max(np.abs(r["en"][i0:i0 + n] - lap[i0:i0 + n]) / np.maximum(np.abs(lap[i0:i0 + n]), 1.0))
            worst["eig"] = max(worst["eig"], err)
            assert err <= 1e-10, (N, b, n, err)
            if n > 4000:
                Z = r["evec"][zoff[b]:zoff[b + 1]].reshape(n, n)   # rows = eigenvectors
                E = r["en"][i0:i0 + n] + r["emin"]
                res = np.max(np.abs(H @ Z.T - Z.T * E[None, :])) / np.abs(H).max()
                orth = np.max(np.abs(Z @ Z.T - np.eye(n)))
                worst["res"] = max(worst["res"], res)
                worst["orth"] = max(worst["orth"], orth)
                assert res <= 1e-10 and orth <= 1e-10, (N, b, n, res, orth)
                del Z
        # Kept set from the LAPACK spectrum with the reference's Ecut rule
        ec = po.ecut(lap, kw["ncut"])
        kept = lap <= ec
        near = np.abs(lap - ec) <= 1e-9 * max(1.0, abs(ec))
        diff = kept != r["kept"]
        print("  N=%d: Ecut %.12g (device %.12g), %d kept, %d states within 1e-9 of the cutoff, %d differ"
              % (N, ec, r["ecut"], int(kept.sum()), int(near.sum()), int(diff.sum())), flush=True)
        assert not np.any(diff & ~near), N
        assert abs(r["ecut"] - ec) <= 1e-10 * max(1.0, abs(ec))
    print("  worst: eigenvalue %.2e, residual %.2e, orthogonality %.2e" % (worst["eig"], worst["res"], worst["orth"]),
          flush=True)


def test_batch64_full_size_matches_single_runs(ctx):
    with open(os.path.join(HERE, "golden", "c2_instances.json")) as f:
        inst = json.load(f)["instances"]
    assert len(inst) == 64
    nsites = 8   # N = -1 .. 6: every instance truncates from N = 4 or 5 on
    params = [dict(U=i["U"], Gamma=i["Gamma"], ed=i["ed"], Lambda=2.5, ncut=i["ncut"]) for i in inst]
    print("fullsize batch64, %d sites" % nsites, flush=True)
    multi, _ = ctx.run_onechqs_multi(params, nsitesmax=nsites, calcdens=2)
    truncated = 0
    for k, p in enumerate(params):
        single, _ = ctx.run_onechqs(nsitesmax=nsites, calcdens=2, **p)
        assert len(single) == len(multi[k])
        for g, s in zip(multi[k], single):
            assert np.array_equal(g["qn"], s["qn"]) and np.array_equal(g["kept"], s["kept"]), (k, g["N"])
            err = np.max(np.abs(g["en"] - s["en"]) / np.maximum(np.abs(s["en"]), 1.0))
            assert err <= 1e-12, (k, g["N"], err)
            np.testing.assert_allclose(g["suscep"], s["suscep"], rtol=1e-10, atol=0)
            truncated += int(not g["kept"].all())
    print("  %d truncated instance-steps compared" % truncated, flush=True)
    assert truncated >= 128


def _dm_betabar(T, L, M):
    return 0.5 * (1.0 + 1.0 / L) * L ** (-(M - 1) / 2.0) / T


def _friedel_checks(h, gamma, a_lambda, tol, w0):
    """w0: a frequency below T whose shell window (CalcSpecDM_NRG's N_w-8 .. N_w+13)
    reaches the chain's last shells"""
    om = np.array([-w0, w0, -1e-3, 1e-3, -0.05, 0.05])
    a = h.spec_omega(om)
    assert abs(a[0] - a[1]) <= 1e-8 * a[1] and abs(a[2] - a[3]) <= 1e-8 * a[3] and abs(a[4] - a[5]) <= 1e-8 * a[5]
    ratio = math.pi * gamma * 0.5 * (a[0] + a[1]) / a_lambda
    n = h.norm()
    print("  CalcNorm %.14f, pi Gamma A(0) / A_Lambda = %.5f" % (n, ratio), flush=True)
    assert abs(n - 1.0) <= 1e-10, n
    assert abs(ratio - 1.0) <= tol, ratio


def test_onechqs_fdm_c1_exact_sum_rules(ctx):
    # NRG_main chain with the A_Lambda factor: pi Gamma A(0) = 1
    kw = dict(U=0.5, Gamma=0.02, ed=-0.25, Lambda=2.5, nsitesmax=40, ncut=400, Mtemp=38,
              betabar=_dm_betabar(1e-8, 2.5, 38), method=2, update_bef_cut=1, bbroad=0.6, fdm_fix=63)
    print("fullsize fdm c1 exact", kw, flush=True)
    with ctx.dm(**kw) as h:
        _friedel_checks(h, kw["Gamma"], 1.0, 0.06, 1e-9)


def test_twochqsz_fdm_exact_sum_rules(ctx):
    # the legacy two-channel chain has no A_Lambda factor: pi Gamma_tot A(0) = A_Lambda;
    # shells N = 0 .. 33 keeping 1500 states reach T = 1e-6 << T_K ~ 1e-3 (D_30 ~ T)
    L = 2.5
    kw = dict(U=0.5, Gamma=0.03, Gamma2=0.01, ed=-0.25, Lambda=L, nsitesmax=33, ncut=1500, model=1, Mtemp=30,
              betabar=_dm_betabar(1e-6, L, 30), method=2, bbroad=0.6, fdm_fix=63)
    print("fullsize fdm twochqsz exact", kw, flush=True)
    a_lam = 0.5 * math.log(L) * (L + 1) / (L - 1)
    with ctx.dm(**kw) as h:
        _friedel_checks(h, kw["Gamma"] + kw["Gamma2"], a_lam, 0.10, 1e-7)
