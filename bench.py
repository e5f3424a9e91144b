"""Benchmark of the MI355X NRG iterative-diagonalisation path.

One "step" = one NRG iteration (one Wilson site) of one impurity instance:
BuildBasis -> H_N assembly -> batched FP64 eigensolve -> SetKept/thermo ->
CutStates -> UpdateMatrices rotations.  The chain first runs the steps in which
the basis is still growing (no truncation yet), then W untimed warm-up steps,
then K timed steps at full truncated size.

Prints ONE JSON line: value = NRG steps/s on one GPU, with the eigensolver's
roofline (HIP-event time on the library's stream) and the CPU oracle timed on
a bounded sample of the same workload.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "NRG steps/s at Ncut=3000, batched vs plain GPU (1 GPU); DM-NRG A(ω) wall time"
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # two-channel Anderson (TwoChQS), the Ncut=3000 configuration of the metric
    "twochqs_c3": dict(model="twochqs", U=0.4, Gamma1=0.03, Gamma2=0.03, ed=-0.2, Lambda=3.0, ncut=3000, ramp=3),
    "twochqs_ncut1000": dict(model="twochqs", U=0.4, Gamma1=0.03, Gamma2=0.03, ed=-0.2, Lambda=3.0, ncut=1000,
                             ramp=3),
    # single-channel Anderson (OneChQS, NRG_main), flat band
    "onechqs_c1": dict(model="onechqs", U=0.5, Gamma=0.02, ed=-0.25, Lambda=2.5, ncut=400, calcdens=2, ramp=5),
    "onechqs_ncut3000": dict(model="onechqs", U=0.5, Gamma=0.02, ed=-0.25, Lambda=2.5, ncut=3000, calcdens=2,
                             ramp=7),
}


def _params(w):
    w = dict(w)
    model = w.pop("model")
    ramp = w.pop("ramp")
    return model, ramp, w


def gpu_run(name, steps, warmup, eig_impl=2, plain=False, stats_all=False, red_streams=None, two_stage=None,
            graph=False, info=None):
    import twochnrg_amd as t
    model, ramp, w = _params(WORKLOADS[name])
    first = -1 if model == "onechqs" else 0   # first record emitted
    total = ramp + warmup + steps             # NRG steps after the first record
    stamps = {}
    sizes = {}
    ctx = t.Context(0)
    ctx.set_eig_impl(eig_impl)
    ctx.set_plain(plain)
    if red_streams is not None:
        ctx.set_red_streams(red_streams)
    if two_stage is not None:
        ctx.set_two_stage(two_stage)
    if graph:
        ctx.set_graph(True)
    ctx.set_timing(True)
    t_start_N = first + ramp + warmup         # stats reset after this step's record

    def on_step(st):
        s = st.contents
        stamps[s.nshell] = time.perf_counter()
        sizes[s.nshell] = [s.bbe[2 * i + 1] - s.bbe[2 * i] + 1 for i in range(s.nblocks)]
        print("bench%s: N=%d done, %d blocks, max block %d, %d states" %
              (" (plain)" if plain else "", s.nshell, s.nblocks, max(sizes[s.nshell]), sum(sizes[s.nshell])),
              file=sys.stderr, flush=True)
        if s.nshell == t_start_N and not stats_all:
            ctx.reset_stats()
            if graph:
                ctx.graph_stats()   # reset

    if model == "onechqs":
        ctx.run_onechqs(nsitesmax=total, keep_steps=False, on_step=on_step, **w)
    else:
        ctx.run_twochqs(nsitesmax=total, keep_steps=False, on_step=on_step, **w)
    st = ctx.stats()
    if graph and info is not None:
        ms, nodes, graphs = ctx.graph_stats()
        info.update(graph_host_ms=ms, graph_nodes=nodes, graphs=graphs)
    ctx.close()
    last = max(stamps)
    dt = stamps[last] - stamps[t_start_N]
    timed = [sizes[N] for N in range(t_start_N + 1, last + 1)]
    return dt, st, timed, model, w, last - t_start_N


# DM-NRG A(omega) of the single-channel Anderson chain at T = 1e-8
# (40 sites, 400 kept multiplets; T = D_M / betabar with M = 38, the fixed-omega grid)
DM_C1 = dict(U=0.5, Gamma=0.02, ed=-0.25, Lambda=2.5, nsitesmax=40, ncut=400, Mtemp=38)


def dm_betabar(T=1e-8, Lambda=2.5, M=38):
    hlf = 0.5 * (1.0 + 1.0 / Lambda)
    return hlf * Lambda ** (-(M - 1) / 2.0) / T


def dm_run(cpu=True):
    import twochnrg_amd as t
    kw = dict(DM_C1, betabar=dm_betabar())
    out = {"workload": "onechqs_dmnrg_c1", "T": 1e-8, "betabar": kw["betabar"], "params": DM_C1}
    with t.Context(0) as ctx:
        t0 = time.perf_counter()
        with ctx.dm(**kw) as d:
            w, r = d.spectrum()
            t1 = time.perf_counter()
            tm = d.timing()
    out.update({"wall_ms": 1e3 * (t1 - t0), "breakdown_ms": tm, "nomega": len(w),
                "post_ms": tm["sweep"] + tm["poles"] + tm["broaden"]})
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as po
        t0 = time.perf_counter()
        o = po.DmRun(**kw)
        wo, ro = o.spectrum()
        t1 = time.perf_counter()
        import numpy as np
        out["cpu_oracle_ms"] = 1e3 * (t1 - t0)
        out["max_rel_err_vs_oracle"] = float(np.max(np.abs(r - ro) / np.maximum(np.abs(ro), 1e-12 * np.abs(ro).max())))
    return out


def fdm_omegas():
    """omega = +-logspace(-10, 0, 1000): 2000 points, negative first (ascending)."""
    import numpy as np
    pos = np.logspace(-10.0, 0.0, 1000)
    return np.concatenate([-pos[::-1], pos])


def fdm_run(cpu=True, cpu_omegas=12):
    """FDM A(omega) (DM_NRG -F -e 0.6) at T = 1e-8 on the single-channel chain of
    the Ncut=400 workload run with UpdateBefCut 1 (every state and its operator
    elements kept per shell), evaluated through CalcSpecDM_NRG's shell window on
    2000 omegas; wall = forward chain + FDM sweep + pole weights + all omegas.
    The CPU oracle runs the same chain and sweep, and a sample of the omegas
    (every len/cpu_omegas-th), extrapolated to the 2000."""
    import numpy as np
    import twochnrg_amd as t
    kw = dict(DM_C1, betabar=dm_betabar(), method=2, update_bef_cut=1, bbroad=0.6)
    om = fdm_omegas()
    out = {"workload": "onechqs_fdm_c1", "T": 1e-8, "betabar": kw["betabar"], "bbroad": 0.6, "nomega": len(om),
           "params": DM_C1}
    with t.Context(0) as ctx:
        t0 = time.perf_counter()
        with ctx.dm(**kw) as d:
            r = d.spec_omega(om)
            t1 = time.perf_counter()
            tm = d.timing()
            norm = d.norm()
            wn, ztot = d.weights()
    # A(0) read at the grid points nearest +-1e-9: the 40-site chain ends near
    # D_N ~ 2e-8, so the shell window of |omega| = 1e-10 holds no shell and A is 0 there
    h = len(om) // 2
    ip = h + int(np.argmin(np.abs(np.log10(om[h:]) + 9.0)))
    a0 = 0.5 * (float(r[ip]) + float(r[len(om) - 1 - ip]))
    out.update({"wall_ms": 1e3 * (t1 - t0), "breakdown_ms": tm, "post_ms": tm["sweep"] + tm["poles"] + tm["broaden"],
                "norm": norm, "ztot": ztot, "shell_weight_max": int(np.argmax(wn)),
                # Friedel sum rule at particle-hole symmetry: pi Gamma A(0) = 1 (the chain carries A_Lambda)
                "friedel_pi_gamma_A0": math.pi * DM_C1["Gamma"] * a0, "friedel_omega": float(om[ip])})
    print("bench fdm: device done (%.0f ms)" % out["wall_ms"], file=sys.stderr, flush=True)
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as po
        t0 = time.perf_counter()
        o = po.DmRun(**kw)
        t1 = time.perf_counter()
        print("bench fdm: oracle chain + sweep %.1f s" % (t1 - t0), file=sys.stderr, flush=True)
        idx = np.linspace(0, len(om) - 1, cpu_omegas).astype(int)
        ro = np.array([o.spec_omega(x) for x in om[idx]])
        t2 = time.perf_counter()
        out["cpu_oracle_ms"] = 1e3 * ((t1 - t0) + (t2 - t1) * len(om) / len(idx))
        out["cpu_sample"] = "chain + sweep timed, %d of %d omegas timed and scaled" % (len(idx), len(om))
        out["max_rel_err_vs_oracle"] = float(np.max(np.abs(r[idx] - ro) / np.maximum(np.abs(ro), 1e-12 * np.abs(ro).max())))
        out["norm_oracle"] = o.norm()
    return out


C5 = dict(U=0.5, Gamma=0.03, Gamma2=0.01, ed=-0.25, Lambda=2.5, nsitesmax=49, ncut=4000, model=1)


def c5_run(cpu=True, nsitesmax=None, ncut=None, modes=((16.0, 0), (4.0, 0), (16.0, 63))):
    """FDM A(omega) of d_up (DM_NRG -F -e 0.6) at T = 1e-9 (Mtemp 46) on the
    TwoChQSz two-channel chain, 50 sites (N = 0 .. 49) keeping 4000 states, on
    omega = +-logspace(-10, 0, 1000); wall = forward chain with every shell's
    states and d_up kept + FDM sweep + pole weights + the 2000 omegas.  Runs
    (d, fdm_fix): the reference's FDM with the two-channel environment count
    d = 16 per later site and with its one-channel d = 4, and the exact full
    density matrix (fdm_fix 63, every quirk of the reference's FDM switched off)
    at d = 16.  Each run reports CalcNorm and the Friedel ratio
    pi Gamma_tot A(0) / A_Lambda (the legacy two-channel chain has no A_Lambda;
    1 at particle-hole symmetry for T << T_K).  The CPU oracle runs the same
    model at a size it finishes in seconds (5 sites, 200 states), with 8 omegas."""
    import numpy as np
    import twochnrg_amd as t
    base = dict(C5)
    if nsitesmax is not None:
        base["nsitesmax"] = nsitesmax
    if ncut is not None:
        base["ncut"] = ncut
    M = min(46, base["nsitesmax"] - 1)   # shorter probe chains keep betabar = D_46 / 1e-9
    T = 1e-9 if M == 46 else dm_betabar(1.0, 2.5, M) / dm_betabar(1e-9, 2.5, 46)
    kw = dict(base, Mtemp=M, betabar=dm_betabar(T, 2.5, M), method=2, bbroad=0.6)
    om = fdm_omegas()
    out = {"workload": "twochqsz_fdm", "T": T, "betabar": kw["betabar"], "bbroad": 0.6, "nomega": len(om),
           "params": base, "runs": {}}
    L = base["Lambda"]
    a_lam = 0.5 * math.log(L) * (L + 1.0) / (L - 1.0)
    g_tot = base["Gamma"] + base["Gamma2"]
    with t.Context(0) as ctx:
        for d, fix in modes:
            t0 = time.perf_counter()
            with ctx.dm(**kw, fdm_site_dim=d, fdm_fix=fix) as h:
                r = h.spec_omega(om)
                t1 = time.perf_counter()
                tm = h.timing()
                norm = h.norm()
                wn, ztot = h.weights()
                mem = h.memory()
            rec = {"wall_ms": 1e3 * (t1 - t0), "breakdown_ms": tm,
                   "device_GB": {"before_sweep": mem["before_sweep"] / 1e9, "with_poles": mem["now"] / 1e9},
                   "post_ms": tm["sweep"] + tm["poles"] + tm["broaden"], "norm": norm, "ztot": ztot,
                   "shell_weight_max": int(np.argmax(wn)), "A_at_omega_pm1e-10": [float(r[999]), float(r[1000])],
                   "friedel_ratio": math.pi * g_tot * 0.5 * (float(r[999]) + float(r[1000])) / a_lam,
                   "A_max": float(r.max()), "fdm_fix": fix}
            out["runs"]["d%d%s" % (int(d), "_exact" if fix == 63 else ("" if fix == 0 else "_fix%d" % fix))] = rec
            print("bench c5: site dim %d fdm_fix %d done (%.0f ms, forward %.0f ms, norm %.12f, friedel %.4f)"
                  % (d, fix, rec["wall_ms"], tm["forward"], norm, rec["friedel_ratio"]), file=sys.stderr, flush=True)
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as po
        small = dict(kw, nsitesmax=5, ncut=200, Mtemp=3, betabar=0.779)
        idx = np.linspace(0, len(om) - 1, 8).astype(int)
        t0 = time.perf_counter()
        o = po.DmRun(**small)
        t1 = time.perf_counter()
        ro = np.array([o.spec_omega(x) for x in om[idx]])
        t2 = time.perf_counter()
        with t.Context(0) as ctx:
            t3 = time.perf_counter()
            with ctx.dm(**small) as h:
                rg = h.spec_omega(om)
            t4 = time.perf_counter()
        out["cpu_oracle_small"] = {
            "sample": "5 sites, 200 states, Mtemp 3: chain + sweep, 8 of the 2000 omegas timed and scaled",
            "cpu_ms": 1e3 * ((t1 - t0) + (t2 - t1) * len(om) / len(idx)), "gpu_ms": 1e3 * (t4 - t3),
            "max_rel_err_vs_oracle": float(np.max(np.abs(rg[idx] - ro) / np.maximum(np.abs(ro),
                                                                                   1e-12 * np.abs(ro).max())))}
        print("bench c5: oracle sample %.1f s" % (t2 - t0), file=sys.stderr, flush=True)
    return out


def c2_run(steps=4, ramp=6, plain_steps=1):
    """64 seeded single-channel Anderson instances (splitmix64 seed 20261203,
    tests/golden/c2_instances.json), Lambda 2.5, Ncut 800..2000, stepped in
    lockstep with one batched eigensolve per step; value = instances x steps / s.
    The plain leg runs the same instances with one block per launch."""
    import json
    import twochnrg_amd as t
    d = json.load(open(os.path.join(ROOT, "tests", "golden", "c2_instances.json")))
    inst = d["instances"]
    ni = len(inst)
    out = {"workload": "onechqs_batch64", "instances": ni, "seed": d["seed"], "ncut_range": [min(x["ncut"] for x in inst),
                                                                                                max(x["ncut"] for x in inst)]}
    for plain in (False, True):
        nst = plain_steps if plain else steps
        total = ramp + nst
        stamps, sizes = {}, {}
        with t.Context(0) as ctx:
            ctx.set_plain(plain)

            def on_step(i, st):
                s = st.contents
                if i == ni - 1:
                    stamps[s.nshell] = time.perf_counter()
                    print("bench c2%s: N=%d done" % (" (plain)" if plain else "", s.nshell), file=sys.stderr,
                          flush=True)
                sizes.setdefault(s.nshell, 0)
                sizes[s.nshell] += s.nstates
            ctx.run_onechqs_multi(inst, nsitesmax=total, calcdens=2, keep_steps=False, on_step=on_step)
        t0 = stamps[ramp - 1]
        t1 = stamps[total - 1]
        rate = ni * nst / (t1 - t0)
        key = "plain_gpu" if plain else "batched"
        out[key] = {"value": rate, "unit": "instance-steps/s", "steps": nst, "ms_per_step": 1e3 * (t1 - t0) / nst,
                    "states_per_step": sizes[total - 1]}
    out["batched_over_plain"] = out["batched"]["value"] / out["plain_gpu"]["value"]
    return out


def adjoint_run(cpu=True):
    """Derivative pass (forward mode, 3 tangents U/Gamma/ed) of the single-channel
    Anderson chain, Lambda 2.5, Ncut 1000 with the Kept sets frozen, to N = 31
    (T = 1e-6 with betabar = D_31/T): d<n_d>, d(T chi) per step; then the DM-NRG
    A(omega) and its Jacobian on omega = +-logspace(-8, 0, 100).  The CPU leg is
    the oracle's derivative pass on the same chain to N = 6, timed per
    full-size step."""
    import numpy as np
    import twochnrg_amd as t
    from twochnrg_amd import autograd as ag
    L, Nout = 2.5, 31
    bb = ag.calc_dn(L, Nout) / 1e-6
    kw = dict(U=0.5, Gamma=0.02, ed=-0.2, Lambda=L, nsitesmax=Nout + 1, ncut=1000, calcdens=2, betabar=bb)
    out = {"workload": "onechqs_adjoint_c4", "params": kw, "T": 1e-6, "ntan": 3}
    with t.Context(0) as ctx:
        ctx.run_onechqs(**dict(kw, nsitesmax=4))   # first-touch of the buffers
        t0 = time.perf_counter()
        steps, _ = ctx.run_onechqs(**kw)
        t1 = time.perf_counter()
        kept = t.kept_arrays(steps)
        st, tn = ctx.onechqs_jvp(kept=kept, dpar=np.eye(3), **kw)
        t2 = time.perf_counter()
        nst = len(steps) - 1
        out["forward_ms"] = 1e3 * (t1 - t0)
        out["jvp_ms"] = 1e3 * (t2 - t1)
        out["forward_steps_per_s"] = nst / (t1 - t0)
        out["jvp_steps_per_s"] = nst / (t2 - t1)
        out["jvp_over_forward"] = (t2 - t1) / (t1 - t0)
        out["jacobian_nd_Tchi"] = [tn[-1]["dopavg"][:, 0].tolist(), tn[-1]["dsuscep"][:, 3].tolist()]
        print("bench adjoint: chain forward %.0f ms, derivative pass %.0f ms" % (out["forward_ms"], out["jvp_ms"]),
              file=sys.stderr, flush=True)
        om = np.concatenate([-np.logspace(-8, 0, 100)[::-1], np.logspace(-8, 0, 100)])
        dkw = dict(Lambda=L, nsitesmax=40, ncut=1000, Mtemp=Nout, betabar=bb)
        run = ag.make_spectral_run(ctx, (0.5, 0.02, -0.2), om, **dkw)
        grid = run["grid"]
        sh, wb = [g[0] for g in grid], [g[1] for g in grid]
        t3 = time.perf_counter()
        with ctx.dm(U=0.5, Gamma=0.02, ed=-0.2, kept=run["kept"], **dkw) as d:
            d.spec(sh, wb)
        t4 = time.perf_counter()
        with ctx.dm(U=0.5, Gamma=0.02, ed=-0.2, kept=run["kept"], dpar=np.eye(3), **dkw) as d:
            d.spec(sh, wb)
            d.spec_jvp(sh, wb)
            tm = d.timing()
        t5 = time.perf_counter()
        out["dm_A_ms"] = 1e3 * (t4 - t3)
        out["dm_A_and_jacobian_ms"] = 1e3 * (t5 - t4)
        out["dm_breakdown_ms"] = tm
        out["nomega"] = len(om)
        print("bench adjoint: A(omega) %.0f ms, with its Jacobian %.0f ms" % (out["dm_A_ms"], out["dm_A_and_jacobian_ms"]),
              file=sys.stderr, flush=True)
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import ctypes
        import pyoracle as po
        okw = dict(kw, nsitesmax=6)
        base = po.run_onechqs(**{k: v for k, v in okw.items()})
        okept = po.kept_arrays(base)
        lib = po.load()
        stamps = []

        def cb(_u, st_, dt_):
            stamps.append((st_.contents.nshell, time.perf_counter(), st_.contents.nstates))
        p = po.OneChQSParams(okw["U"], okw["Gamma"], okw["ed"], L, okw["nsitesmax"], okw["ncut"], 2, 0, 1e-18, bb, 0)
        dp = np.ascontiguousarray(np.eye(3))
        kf, ko = np.ascontiguousarray(okept[0]), np.ascontiguousarray(okept[1])
        fn = po.TAN_CB(cb)
        t0 = time.perf_counter()
        lib.orc_run_onechqs_jvp(ctypes.byref(p), kf.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)),
                                ko.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)), 3,
                                dp.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), fn, None)
        dt = stamps[-1][1] - stamps[-2][1]
        out["cpu_oracle"] = {"value": 1.0 / dt, "unit": "derivative-pass steps/s", "cores": 1, "kind": "port",
                             "sample": "oracle/nrg_oracle.c derivative pass (3 tangents) on the same chain to N = %d "
                                       "(%d states at the last step), last step timed: %.2f s, run %.1f s" % (
                                           stamps[-1][0], stamps[-1][2], dt, time.perf_counter() - t0)}
        # the device's rate over its full-size steps, for the same unit
        out["device_vs_cpu_oracle"] = out["jvp_steps_per_s"] * dt
        print("bench adjoint: oracle step %.2f s" % dt, file=sys.stderr, flush=True)
    return out


def autotune_run(name):
    """The eigensolver's launch configuration for this workload, picked by the
    library's autotuner (nrg_autotune_eig) on random blocks of the sizes of a
    truncated step of the workload (recorded from a short chain), with the
    step's kept fraction of eigenvector rows.  Outside the timed region."""
    import twochnrg_amd as t
    model, ramp, w = _params(WORKLOADS[name])
    sizes = {}

    def on_step(st):
        s = st.contents
        sizes[s.nshell] = [s.bbe[2 * i + 1] - s.bbe[2 * i] + 1 for i in range(s.nblocks)]

    t0 = time.perf_counter()
    with t.Context(0) as ctx:
        if model == "onechqs":
            ctx.run_onechqs(nsitesmax=ramp + 1, keep_steps=False, on_step=on_step, **w)
        else:
            ctx.run_twochqs(nsitesmax=ramp + 1, keep_steps=False, on_step=on_step, **w)
        last = max(sizes)
        sz = sizes[last]
        frac = min(1.0, w["ncut"] / float(sum(sz)))
        best, rows = ctx.autotune(sz, krow_frac=frac, reps=1)
    return {"sizes_from_step": last, "blocks": len(sz), "max_block": max(sz), "krow_frac": frac,
            "chosen": best, "table": rows, "wall_s": time.perf_counter() - t0}


def hbm_probe_child():
    """hbm_probe in a child process: torch's HIP runtime cannot open the device
    in a process where the library's runtime already holds it."""
    import subprocess
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--hbm-probe"], stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=300)
    if r.returncode != 0:
        return {"error": r.stderr.strip().splitlines()[-1] if r.stderr.strip() else "rc %d" % r.returncode}
    return json.loads(r.stdout.strip().splitlines()[-1])


def symv_probe(n=4419, reps=2, copy_bytes=1 << 30):
    """For hardware-counter runs of the reduction's y = A v kernel: one random
    n x n symmetric eigensolve through the library (reduction on one stream),
    then a device-to-device copy of copy_bytes through the HIP runtime (the
    counters' calibration: its kernel reads and writes copy_bytes).  Prints the
    algorithmic bytes of the average symv launch of the solve: the lower
    triangle of the trailing matrix, 4 m (m + 1) bytes for m = n - k - 1, over
    the columns k = 0 .. n - 3."""
    import ctypes
    import numpy as np
    import twochnrg_amd as t
    rng = np.random.default_rng(0)
    a = rng.standard_normal((n, n))
    a = (a + a.T) / 2
    with t.Context(0) as ctx:
        ctx.set_red_streams(1)
        for _ in range(reps):
            ctx.diag_blocks([a])
    hip = ctypes.CDLL("libamdhip64.so.7")
    x, y = ctypes.c_void_p(), ctypes.c_void_p()
    hip.hipMalloc(ctypes.byref(x), ctypes.c_size_t(copy_bytes))
    hip.hipMalloc(ctypes.byref(y), ctypes.c_size_t(copy_bytes))
    hip.hipMemset(x, 0, ctypes.c_size_t(copy_bytes))
    hip.hipMemcpy(y, x, ctypes.c_size_t(copy_bytes), 3)   # device to device
    hip.hipDeviceSynchronize()
    hip.hipFree(x)
    hip.hipFree(y)
    cols = range(0, n - 2)
    alg = sum(4.0 * (n - k - 1) * (n - k) for k in cols) / len(cols)
    return {"n": n, "solves": reps, "symv_launches_per_solve": len(cols), "alg_bytes_per_launch": alg,
            "copy_bytes": copy_bytes}


def hbm_probe(nbytes=2 << 30, reps=10):
    """STREAM-like calibration of the nominal HBM peak on this device: a 2 GiB
    FP64 copy (torch's copy kernel, read + write) and a 2 GiB read (torch.sum),
    timed with events on torch's stream."""
    import torch
    n = nbytes // 8
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    out = {}
    for name, fn, moved in (("copy", lambda: y.copy_(x), 2 * nbytes), ("read", lambda: x.sum(), nbytes)):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        out[name + "_GBs"] = moved * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del x, y
    torch.cuda.empty_cache()
    # the FP64 matrix-core ceiling as the vendor library reaches it: torch's
    # DGEMM (hipBLASLt / rocBLAS) on 8192^3, the rate the rotation and
    # back-transformation GEMMs are compared with
    m = 8192
    a = torch.randn(m, m, dtype=torch.float64, device="cuda")
    b = torch.randn(m, m, dtype=torch.float64, device="cuda")
    c = a @ b
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        c = a @ b
    e1.record()
    torch.cuda.synchronize()
    out["dgemm_TFLOPs"] = 5 * 2.0 * m ** 3 / (e0.elapsed_time(e1) * 1e-3) / 1e12
    out["dgemm_shape"] = [m, m, m]
    del a, b, c
    torch.cuda.empty_cache()
    return out


def twochqs_error(name):
    """Errors of the device eigensolve at the timed size, outside the timed
    region: the workload's first truncated step (N = 3: 44-45k multiplets in
    ~53 blocks of up to ~4300 rows, the sizes of the timed steps) run again
    with its H_N blocks recorded.  Per block the residual max|H z - E z| /
    max|H| over every eigenvector (for symmetric H it bounds the eigenvalue
    error), the orthogonality max|Z Z^T - 1| of the largest block, and the
    eigenvalues against numpy's eigvalsh for blocks up to 1500 rows."""
    import numpy as np
    import twochnrg_amd as t
    model, ramp, w = _params(WORKLOADS[name])
    if model != "twochqs":
        return None
    rec = {}

    def on_step(st):
        d = t.step_to_dict(st)
        if d["N"] == 3:
            rec.update(d)

    t0 = time.perf_counter()
    with t.Context(0) as ctx:
        ctx.run_twochqs(nsitesmax=4, record_blocks=True, keep_steps=False, on_step=on_step, **w)
    sizes = (rec["bbe"][:, 1] - rec["bbe"][:, 0] + 1).astype(int)
    hoff = np.concatenate([[0], np.cumsum(sizes * (sizes + 1) // 2)])
    zoff = np.concatenate([[0], np.cumsum(sizes * sizes)])
    res = eig = 0.0
    orth = None
    for b in np.argsort(-sizes):
        n = int(sizes[b])
        A = np.zeros((n, n))
        A[np.triu_indices(n)] = rec["hpacked"][hoff[b]:hoff[b + 1]]
        H = A + np.triu(A, 1).T
        Z = rec["evec"][zoff[b]:zoff[b + 1]].reshape(n, n)
        i0 = int(rec["bbe"][b, 0])
        E = rec["en"][i0:i0 + n] + rec["emin"]
        scale = max(np.abs(H).max(), 1e-300)
        res = max(res, float(np.abs(H @ Z.T - Z.T * E[None, :]).max() / scale))
        if orth is None:
            orth = float(np.abs(Z @ Z.T - np.eye(n)).max())
        if n <= 1500:
            wl = np.linalg.eigvalsh(H)
            eig = max(eig, float(np.max(np.abs(np.sort(E) - wl) / np.maximum(np.abs(wl), 1.0))))
    out = {"step": 3, "blocks": int(len(sizes)), "max_block": int(sizes.max()), "states": int(sizes.sum()),
           "max_residual_rel": res, "orthogonality_largest_block": orth,
           "max_eigenvalue_rel_err_vs_lapack_n_le_1500": eig, "wall_s": time.perf_counter() - t0}
    print("bench: error check at N=3: residual %.2e, orthogonality %.2e, eigenvalues %.2e (%.1f s)"
          % (res, orth, eig, out["wall_s"]), file=sys.stderr, flush=True)
    return out


def _host_cpu():
    """The host's CPU model (/proc/cpuinfo, as lscpu reports it), the logical
    CPU count and the CPUs this process may run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 1
    return {"model": model, "nproc": os.cpu_count() or 1, "usable": usable}


def cpu_baseline(name, timed_sizes, rot_flops_per_step, budget_s=24.0):
    """The CPU oracle (gcc -O2 -funroll-loops, FP64) on a bounded sample of one
    timed step of the same workload.

    * Eigensolves: the oracle's dspev-equivalent (Householder + QL) on blocks of
      that step's sizes, the largest that fit the budget down to the smallest,
      each run three times (median), fitted as a*n^3 and summed over the step's
      full block list.
    * Rotations: the oracle's naive GEMM rotation Zi F Zj^T timed on one pair
      (median of three) and scaled by the step's rotation flops.
    * Single core, as the reference runs.  A second figure, labelled multi-core
      CPU, runs the same oracle eigensolves on all usable host CPUs at once (one
      thread per block, ctypes releases the GIL): its parallel efficiency is
      measured on P concurrent blocks, and the step's blocks are scheduled
      largest-first over P cores with that efficiency.
    Basis build and H_N assembly, O(n^2) per block, are left out on both sides
    of the comparison's CPU figure (they are below 1 % of a step here)."""
    import concurrent.futures as cf
    import statistics
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import pyoracle as po
    host = _host_cpu()
    sizes = sorted(timed_sizes[-1], reverse=True)
    rng = np.random.default_rng(0)
    t_all = time.perf_counter()

    def sym(n):
        a = rng.standard_normal((n, n))
        return (a + a.T) / 2

    def timed(fn, reps=3):
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return statistics.median(ts)

    # eigensolve coefficient from distinct sizes of the step, three runs each;
    # a 256-row block first sets the budget estimate (hosts differ several-fold)
    cal = sym(256)
    a3_est = timed(lambda: po.eigh_full(cal), reps=1) / 256.0 ** 3
    done, coef = [], []
    for n in sorted(set(sizes), reverse=True):
        if n < 64:
            continue
        a3 = max(float(np.median(coef)) if coef else a3_est, a3_est)
        left = 0.55 * budget_s - (time.perf_counter() - t_all)
        if 3.0 * a3 * float(n) ** 3 > left:
            continue
        a = sym(n)
        dt = timed(lambda: po.eigh_full(a))
        done.append(n)
        coef.append(dt / float(n) ** 3)
        print("bench: cpu sample n=%d %.2f s (median of 3)" % (n, dt), file=sys.stderr, flush=True)
        if len(done) >= 6:
            break
    if not coef:
        return None
    a3 = float(np.median(coef))
    eig_s = a3 * float(sum(float(n) ** 3 for n in sizes))
    # rotation rate of the oracle's GEMM
    nr, kr = 900, 300
    zi, f, zj = rng.standard_normal((kr, nr)), sym(nr), rng.standard_normal((kr, nr))
    rt = timed(lambda: po.rotate(zi, f, zj))
    rot_rate = (2.0 * kr * nr * nr + 2.0 * kr * kr * nr) / rt
    rot_s = rot_flops_per_step / rot_rate if rot_flops_per_step else 0.0
    step_s = eig_s + rot_s
    # multi-core: P concurrent eigensolves of one size against one alone
    P = max(1, min(host["usable"], 16))
    nm = 700
    mats = [sym(nm) for _ in range(P)]
    t1 = timed(lambda: po.eigh_full(mats[0]), reps=1)
    with cf.ThreadPoolExecutor(max_workers=P) as ex:
        t0 = time.perf_counter()
        list(ex.map(po.eigh_full, mats))
        tP = time.perf_counter() - t0
    eff = min(1.0, t1 / tP)   # each of the P blocks took tP instead of t1
    loads = [0.0] * P
    for n in sizes:   # largest first onto the least-loaded core
        i = loads.index(min(loads))
        loads[i] += a3 * float(n) ** 3 / eff
    mc_s = max(loads) + rot_s / (P * eff)
    print("bench: cpu baseline %.1f s per step (1 core), %.1f s (%d cores, efficiency %.2f); sample %.1f s"
          % (step_s, mc_s, P, eff, time.perf_counter() - t_all), file=sys.stderr, flush=True)
    return {"value": 1.0 / step_s, "unit": "NRG steps/s", "cores": 1, "kind": "port",
            "host": host,
            "sample": "oracle/nrg_oracle.c on one timed step of this workload (%d blocks, sum n = %d): eigensolve "
                      "(Householder+QL, as dspev) timed on %d block sizes (n=%s), median of 3 runs each, fitted as "
                      "%.3g*n^3 s = %.1f s per step; rotations (naive GEMM, as netlib cblas_dgemm) %.2f GFLOP/s "
                      "over the step's %.3g flops = %.1f s; sample %.1f s of CPU time"
                      % (len(sizes), sum(sizes), len(done), ",".join(map(str, done)), a3, eig_s, rot_rate / 1e9,
                         rot_flops_per_step, rot_s, time.perf_counter() - t_all),
            "multi_core": {"label": "multi-core CPU", "value": 1.0 / mc_s, "unit": "NRG steps/s", "cores": P,
                           "parallel_efficiency": eff,
                           "sample": "the same oracle eigensolves on %d threads: %d concurrent n=%d blocks took "
                                     "%.2f s against %.2f s for one alone; the step's blocks scheduled largest-first "
                                     "over %d cores" % (P, P, nm, tP, t1, P)}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="twochqs_c3", choices=sorted(WORKLOADS))
    ap.add_argument("--eig-impl", type=int, default=2)
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip the eigensolver launch-configuration autotuner (library defaults)")
    ap.add_argument("--graph", action="store_true",
                    help="capture each eigensolve's launch sequences into a hipGraph (launched once)")
    ap.add_argument("--two-stage", type=int, default=None,
                    help="blocks with n >= this use the two-stage reduction (0: none)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-dm", action="store_true", help="skip the DM-NRG A(omega) leg")
    ap.add_argument("--no-c2", action="store_true", help="skip the 64-instance batched leg")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the derivative-pass leg")
    ap.add_argument("--no-c5", action="store_true", help="skip the two-channel FDM leg")
    ap.add_argument("--c5-only", action="store_true",
                    help="run only the two-channel FDM leg and print its JSON object")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="run only the derivative-pass leg (device part) and print its JSON object")
    ap.add_argument("--no-error", action="store_true",
                    help="skip the eigensolve error check at the timed size (a recorded first truncated step)")
    ap.add_argument("--stats-all", action="store_true",
                    help="kernel statistics over the whole run (ramp-up included), for counter runs")
    ap.add_argument("--plain-steps", type=int, default=2,
                    help="steps of the plain one-block-per-launch comparison run (0 = skip)")
    ap.add_argument("--hbm-probe", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--symv-probe", action="store_true",
                    help="one 4419-row eigensolve and a 1 GiB copy, for rocprofv3 --pmc runs of tdb_symv_kernel")
    a = ap.parse_args()
    if a.hbm_probe:
        print(json.dumps(hbm_probe()))
        return
    if a.symv_probe:
        print(json.dumps(symv_probe()))
        return
    if a.gpus != 1:
        print("bench.py: this build measures one GPU only (the NRG step and its batched blocks run on one "
              "device); --gpus %d is not supported" % a.gpus, file=sys.stderr)
        sys.exit(2)
    if a.adjoint_only:
        print(json.dumps(adjoint_run(cpu=False)))
        return
    if a.c5_only:
        print(json.dumps(c5_run(cpu=not a.no_cpu)))
        return
    ginfo = {}
    tune = None
    red, ts, gr = None, a.two_stage, a.graph
    if not a.no_autotune and a.eig_impl == 2 and a.two_stage is None and not a.graph:
        tune = autotune_run(a.workload)
        red, ts, gr = tune["chosen"]["red_streams"], tune["chosen"]["ts_min"], bool(tune["chosen"]["graph"])
    dt, st, timed, model, w, nsteps = gpu_run(a.workload, a.steps, a.warmup, a.eig_impl, stats_all=a.stats_all,
                                              red_streams=red, two_stage=ts, graph=gr, info=ginfo)
    value = nsteps / dt
    eig = st["eig"]
    eig_rate = eig["flops"] / (eig["ms"] * 1e-3) / 1e12 if eig["ms"] > 0 else 0.0
    cfg = {"workload": a.workload,
           "model": "TwoChQS Anderson (legacy standalone)" if model == "twochqs" else "OneChQS Anderson (NRG_main)",
           "eig_impl": "tridiag+dc+wy" if a.eig_impl == 2 else "wg-per-block householder+ql",
           "blocks_per_step": round(sum(len(s) for s in timed) / len(timed), 1),
           "max_block": max(max(s) for s in timed),
           "states_per_step": round(sum(sum(s) for s in timed) / len(timed), 1)}
    cfg.update({k: v for k, v in w.items()})
    _, ramp, _ = _params(WORKLOADS[a.workload])
    # sites run: the timed steps sit at the truncated steady state (reached from N = 3),
    # so a longer chain repeats the same per-step work
    cfg["nsitesmax"] = ramp + a.warmup + a.steps
    sy = st["tdb_symv"]
    if sy["launches"] > 0:
        # dominant kernel of the large-block workloads: y = A v of the blocked reduction,
        # algorithmic bytes = the lower triangles of the trailing matrices it reads
        # (4 m (m+1) per active block).  The headline run overlaps the reduction of
        # three block groups on three streams, so its launches share the GPU; the
        # kernel's own rate is taken from a short run of the same workload with the
        # reduction on one stream.
        _, st1, _, _, _, n1 = gpu_run(a.workload, 2, 0, a.eig_impl, red_streams=1)
        sy = st1["tdb_symv"]
        avg_s = sy["ms"] * 1e-3 / sy["launches"]
        bpl = sy["bytes"] / sy["launches"]
        ach = bpl / avg_s / 1e9
        roof = {"bound": "hbm", "kernel": "tdb_symv_kernel", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": ach / HBM_PEAK_GBS, "traffic": None,
                "traffic_note": "rocprofv3 --pmc FETCH_SIZE/WRITE_SIZE aborted in the profiler at this kernel's "
                                "first dispatch; not collected",
                "bytes_per_launch": bpl,
                "avg_launch_us": avg_s * 1e6, "timed_launches_per_step": sy["launches"] / n1,
                "timing": "HIP events around every 8th column's launch, in a separate 2-step run of the "
                          "same workload with the reduction on one stream (the headline run uses three)"}
    else:
        roof = {"bound": "mfma", "kernel": "eigensolve (tridiag+dc+wy)", "achieved": eig_rate,
                "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": eig_rate / FP64_PEAK_TFLOPS, "traffic": None}
    out = {
        "metric": METRIC, "value": value, "unit": "NRG steps/s", "n_gpus": 1, "steps": nsteps,
        "warmup": a.warmup, "ms_per_step": 1e3 * dt / nsteps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic", "config": cfg,
        "roofline": roof,
        "eigensolve": {"tflops": eig_rate, "frac_fp64_peak": eig_rate / FP64_PEAK_TFLOPS,
                       "flops_per_step": eig["flops"] / nsteps, "ms_per_step": eig["ms"] / nsteps},
        "kernel_ms_per_step": {k: v["ms"] / nsteps for k, v in st.items()},
    }
    if tune is not None:
        out["autotune"] = tune
    if gr:
        out["graph"] = {"graphs_per_step": ginfo.get("graphs", 0) / nsteps,
                        "nodes_per_step": ginfo.get("graph_nodes", 0) / nsteps,
                        "host_ms_per_step": ginfo.get("graph_host_ms", 0.0) / nsteps}
    gm = st.get("gemm")
    if gm and gm["ms"] > 0:
        # operator rotations Z_i F Z_j^T (FP64 MFMA tile kernel): algorithmic flops / HIP-event time
        out["rotation"] = {"ms_per_step": gm["ms"] / nsteps, "tflops": gm["flops"] / (gm["ms"] * 1e-3) / 1e12,
                           "frac_fp64_peak": gm["flops"] / (gm["ms"] * 1e-3) / 1e12 / FP64_PEAK_TFLOPS}
    if model == "twochqs" and not a.no_error:
        out["error_vs_fp64"] = twochqs_error(a.workload)
    cal = hbm_probe_child()
    roof["measured_hbm"] = {k: v for k, v in cal.items() if not k.startswith("dgemm")}
    if roof["unit"] == "GB/s" and "read_GBs" in cal:
        roof["frac_of_measured_read"] = roof["achieved"] / cal["read_GBs"]
    if "dgemm_TFLOPs" in cal:
        # measured FP64 matrix-core ceiling (vendor DGEMM on the box) beside the nominal 78.6
        out["measured_dgemm"] = {"TFLOPs": cal["dgemm_TFLOPs"], "shape": cal.get("dgemm_shape"),
                                 "frac_of_nominal": cal["dgemm_TFLOPs"] / FP64_PEAK_TFLOPS}
        if "rotation" in out:
            out["rotation"]["frac_of_measured_dgemm"] = out["rotation"]["tflops"] / cal["dgemm_TFLOPs"]
        if roof["unit"] == "TFLOP/s":
            roof["frac_of_measured_dgemm"] = roof["achieved"] / cal["dgemm_TFLOPs"]
    if a.plain_steps > 0:
        # the same kernels, one symmetry block (eigensolve) or one block pair
        # (rotation) per launch, host-synchronised, no batching
        pdt, pst, _, _, _, pn = gpu_run(a.workload, a.plain_steps, 0, a.eig_impl, plain=True)
        out["plain_gpu"] = {"value": pn / pdt, "unit": "NRG steps/s", "steps": pn,
                            "ms_per_step": 1e3 * pdt / pn, "batched_over_plain": value / (pn / pdt)}
    if not a.no_cpu:
        cb = cpu_baseline(a.workload, timed, gm["flops"] / nsteps if gm else 0.0)
        if cb:
            out["cpu_baseline"] = cb
    if not a.no_dm:
        out["dmnrg"] = dm_run(cpu=not a.no_cpu)
        out["fdm"] = fdm_run(cpu=not a.no_cpu)
    if not a.no_c2:
        out["batched_instances"] = c2_run()
    if not a.no_adjoint:
        out["adjoint"] = adjoint_run(cpu=not a.no_cpu)
    if not a.no_c5:
        out["twochqsz_fdm"] = c5_run(cpu=not a.no_cpu)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
