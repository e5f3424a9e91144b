"""Extract the reference's golden run (Runs/run1) into a JSON fixture.

Reads /root/reference/Runs/run1/{input_nrg.dat,output_NRG_test.txt} as data
(no reference code is run) and writes run1_golden.json beside this script:
per N the printed levels with their (Q,S) labels and kept flags, block/state
counts, Emin, Ecut, chi_N and the "T= ... Ndot=/NdSq=/SdSq=" lines.
"""
import json
import os
import re
import sys

REF = "/root/reference/Runs/run1"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    with open(os.path.join(REF, "input_nrg.dat")) as f:
        toks = f.read().split()
    params = {
        "nsitesmax": int(toks[0]), "ncut": int(toks[1]), "U": float(toks[2]),
        "Gamma": float(toks[3]), "ed": float(toks[4]), "Lambda": float(toks[5]),
        "calcdens": int(toks[9]),
    }
    lines = open(os.path.join(REF, "output_NRG_test.txt")).read().splitlines()
    levels = {}
    counts = {}
    emin = {}
    ecut = {}
    chin = {}
    temps = {}
    cur_blocks = None
    cur_states = None
    pending_emin = None
    pending_ecut = None
    pending_chi = None
    t_vals = []
    re_en = re.compile(r"^En = (\S+) ---\| (.*?)\s*(\(K\))?>_N= (-?\d+)$")
    for ln in lines:
        m = re_en.match(ln.strip())
        if m:
            n = int(m.group(4))
            qn = [float(x) for x in m.group(2).split()]
            levels.setdefault(n, []).append([float(m.group(1)), qn, bool(m.group(3))])
            if pending_emin is not None:
                emin[n] = pending_emin
                pending_emin = None
            if pending_ecut is not None:
                ecut[n] = pending_ecut
                pending_ecut = None
            if pending_chi is not None:
                chin[n] = pending_chi
                pending_chi = None
            if cur_blocks is not None:
                counts[n] = [cur_blocks, cur_states]
                cur_blocks = cur_states = None
            continue
        s = ln.strip()
        if s.startswith("No blocks ="):
            cur_blocks = int(s.split("=")[1])
        elif s.startswith("No states ="):
            cur_states = int(s.split("=")[1])
        elif s.startswith("Emin ="):
            pending_emin = float(s.split("=")[1])
        elif s.startswith("SetKept: Ncutoff ="):
            pending_ecut = float(s.split("Ecut =")[1])
        elif s.startswith("chi_N ="):
            pending_chi = float(s.split("=")[1])
        elif s.startswith("T= "):
            parts = s.split()
            t_vals.append([float(parts[1]), parts[2].rstrip("="), float(parts[3])])
    for T, name, val in t_vals:
        temps.setdefault("%.6g" % T, {})[name] = val
    out = {
        "source": "Runs/run1/output_NRG_test.txt (reference golden run, 6 significant digits)",
        "params": params,
        "levels": {str(k): v for k, v in sorted(levels.items())},
        "counts": {str(k): v for k, v in sorted(counts.items())},
        "emin": {str(k): v for k, v in sorted(emin.items())},
        "ecut": {str(k): v for k, v in sorted(ecut.items())},
        "chi_N": {str(k): v for k, v in sorted(chin.items())},
        "ops": temps,
        "chi_m1_printed": 0.121207,
    }
    with open(os.path.join(HERE, "run1_golden.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("levels per N:", {k: len(v) for k, v in levels.items()}, file=sys.stderr)


if __name__ == "__main__":
    main()
