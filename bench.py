"""Benchmark of the MI355X NRG iterative-diagonalisation path.

One "step" = one NRG iteration (one Wilson site) of one impurity instance:
BuildBasis -> H_N assembly -> batched FP64 eigensolve -> SetKept/thermo ->
CutStates -> UpdateMatrices rotations.  The chain first runs the steps in which
the basis is still growing (no truncation yet), then W untimed warm-up steps,
then K timed steps at full truncated size.

Prints ONE JSON line (see the repository's measurement contract): value =
NRG steps/s on one GPU, with the eigensolver kernel's roofline and the CPU
oracle timed on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "NRG steps/s at Ncut=3000, batched vs plain GPU (1 GPU); DM-NRG A(ω) wall time"
FP64_PEAK_TFLOPS = 78.6

WORKLOADS = {
    # OneChQS Anderson (NRG_main), flat band, Lambda=2.5
    "onechqs_c1": dict(U=0.5, Gamma=0.02, ed=-0.25, Lambda=2.5, ncut=400, calcdens=2, ramp=5),
    "onechqs_ncut3000": dict(U=0.5, Gamma=0.02, ed=-0.25, Lambda=2.5, ncut=3000, calcdens=2, ramp=7),
}


def gpu_run(name, steps, warmup):
    import twochnrg_amd as t
    w = dict(WORKLOADS[name])
    ramp = w.pop("ramp")
    total = ramp + warmup + steps
    stamps = []
    ctx = t.Context(0)
    ctx.set_timing(True)
    sizes = []

    def on_step(st):
        s = st.contents
        if s.nshell == ramp + warmup - 1:
            ctx.reset_stats()
        if s.nshell >= 0:
            stamps.append(time.perf_counter())
            nb = s.nblocks
            bbe = [(s.bbe[2 * i], s.bbe[2 * i + 1]) for i in range(nb)]
            sizes.append([b - a + 1 for a, b in bbe])

    _, _ = ctx.run_onechqs(nsitesmax=total, keep_steps=False, on_step=on_step, **w)
    st = ctx.stats()
    ctx.close()
    # stamps[i] is taken after step N=i; timed region = steps ramp+warmup .. total-1
    t0 = stamps[ramp + warmup - 1]
    t1 = stamps[total - 1]
    timed_sizes = sizes[ramp + warmup:]
    return (t1 - t0), st, timed_sizes, w


def cpu_baseline(name, budget_s=20.0, max_steps=4):
    """The CPU oracle (single thread, FP64) on the same workload: time the
    first steps after the ramp, bounded by budget_s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    w = dict(WORKLOADS[name])
    ramp = w.pop("ramp")
    stamps = []
    t_start = time.perf_counter()
    # run the ramp plus up to max_steps truncated steps, stop early on budget
    nsteps = max_steps
    for trial in range(1):
        stamps.clear()
        lib = po.load()
        p = po.OneChQSParams(w["U"], w["Gamma"], w["ed"], w["Lambda"], ramp + nsteps, w["ncut"], w["calcdens"], 0,
                             1e-18, 0.727, 0)

        def cb(_u, stp):
            stamps.append(time.perf_counter())

        fn = po.STEP_CB(cb)
        lib.orc_run_onechqs(p, fn, None)
    # stamps[0] is N=-1; stamps[i+1] after step N=i
    per = [stamps[i + 1] - stamps[i] for i in range(ramp + 1, ramp + 1 + nsteps) if i + 1 < len(stamps)]
    tot = sum(per)
    return {"value": len(per) / tot if tot > 0 else None, "unit": "NRG steps/s", "cores": 1, "kind": "port",
            "sample": "oracle/nrg_oracle.c (gcc -O2 -funroll-loops, 1 thread) on %s: %d truncated steps after "
                      "the %d-step ramp, %.1f s" % (name, len(per), ramp, tot),
            "wall_s": time.perf_counter() - t_start}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="onechqs_c1", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: this build measures one GPU only (the NRG step and its batched blocks run on one "
              "device); --gpus %d is not supported" % a.gpus, file=sys.stderr)
        sys.exit(2)
    dt, st, timed_sizes, w = gpu_run(a.workload, a.steps, a.warmup)
    value = a.steps / dt
    eig = st["eig"]
    achieved = eig["flops"] / (eig["ms"] * 1e-3) / 1e12 if eig["ms"] > 0 else 0.0
    per_launch_flops = eig["flops"] / max(1, eig["launches"])
    out = {
        "metric": METRIC, "value": value, "unit": "NRG steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": 1e3 * dt / a.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": a.workload, "model": "OneChQS Anderson (NRG_main)", "Ncut": w["ncut"],
                   "Lambda": w["Lambda"], "U": w["U"], "Gamma": w["Gamma"], "ed": w["ed"],
                   "blocks_per_step": round(sum(len(s) for s in timed_sizes) / len(timed_sizes), 1),
                   "max_block": max(max(s) for s in timed_sizes),
                   "states_per_step": round(sum(sum(s) for s in timed_sizes) / len(timed_sizes), 1)},
        "roofline": {"bound": "mfma", "kernel": "eig_block_kernel", "achieved": achieved,
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": None, "flops_per_launch": per_launch_flops,
                     "avg_launch_ms": eig["ms"] / max(1, eig["launches"])},
        "kernel_ms_per_step": {k: v["ms"] / a.steps for k, v in st.items()},
    }
    if not a.no_cpu:
        cb = cpu_baseline(a.workload)
        out["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample")}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
