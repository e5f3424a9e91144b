"""C-ABI checks that need no GPU: the library loads, exports every symbol the
public header declares, the structs the binding passes have the header's
layout, and its host-side truncation (SetDegen/Ecut/SetKept) matches the
oracle bit-exactly."""
import os
import re

import numpy as np
import pytest

import pyoracle as po
import twochnrg_amd as t

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_exports_match_header():
    hdr = open(os.path.join(ROOT, "include", "twochnrg_amd.h")).read()
    declared = set(re.findall(r"^\s*(?:int|void|const char \*)\s*\*?\s*(nrg_\w+)\s*\(", hdr, re.M))
    assert declared, "no declarations parsed"
    assert declared == set(t.EXPORTS)
    L = t.lib()
    for name in declared:
        assert hasattr(L, name), name


def test_truncate_matches_oracle_random():
    rng = np.random.default_rng(7)
    for _ in range(300):
        n = int(rng.integers(2, 80))
        scale = rng.choice([1.0, 1e-17, 1e-3])
        e = np.round(rng.random(n) * scale, int(rng.choice([3, 20])))
        e[rng.integers(0, n)] = 0.0
        ncut = int(rng.integers(1, n + 3))
        eps = float(rng.choice([1e-18, 1e-13, 1e-3]))
        e2, kept, ec, nk = t.truncate(e, ncut, eps)
        o = po.setdegen(e, eps)
        assert np.array_equal(e2, o)
        assert ec == po.ecut(o, ncut)
        assert np.array_equal(kept, o <= ec) and nk == int(np.sum(o <= ec))


def test_truncate_matches_oracle_large():
    # tens of thousands of states: the library sorts on several host threads
    # there (chunk sorts and stable merges); the result must still be the oracle's
    rng = np.random.default_rng(11)
    for n, eps in ((40000, 1e-4), (40000, 1.5e-3), (20000, 1e-13)):
        e = np.round(rng.random(n) * 400.0, 3)
        e[rng.integers(0, n)] = 0.0
        ncut = n // 7
        e2, kept, ec, nk = t.truncate(e, ncut, eps)
        o = po.setdegen(e, eps)
        assert np.array_equal(e2, o)
        assert ec == po.ecut(o, ncut) == t.ecut(o, ncut)
        assert np.array_equal(kept, o <= ec) and nk == int(np.sum(o <= ec))


def test_ecut_window_and_edges():
    # 1% window extends the cut over near-degenerate levels
    e = np.array([0.0, 1.0, 1.001, 1.005, 1.2, 2.0])
    assert t.ecut(e, 2) == po.ecut(e, 2) == 1.005
    # Ncut >= n-1 keeps everything
    assert t.ecut(e, 5) == 2.0 and t.ecut(e, 9) == 2.0
    # window running to the end of the array keeps every state
    e = np.array([0.0, 1.0, 1.001, 1.002])
    assert t.ecut(e, 2) == po.ecut(e, 2) == 1.002
    # Ecut level 0 (Ncut=1): relative test divides by zero -> no extension
    e = np.array([0.0, 0.0, 0.5])
    assert t.ecut(e, 1) == po.ecut(e, 1) == 0.0


def test_setdegen_order_dependence():
    # chains closer than eps snap in index order (src/NRGarrayClass.cpp:244-268)
    e = np.array([3e-19, 0.0, 6e-19, 1.1e-18, 5.0])
    for eps in (5e-19, 1e-18):
        e2, _, _, _ = t.truncate(e, 5, eps)
        assert np.array_equal(e2, po.setdegen(e, eps))


def test_bad_arguments_raise():
    with pytest.raises(t.NrgError):
        t.truncate(np.zeros(0), 1)
    with pytest.raises(t.NrgError):
        t.ecut(np.ones(3), 0)


def test_struct_layouts_match_header(tmp_path):
    """Every struct the Python binding passes through the C-ABI has the size and
    field offsets the public header gives it (compiled with the host C compiler)."""
    import ctypes
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no host C compiler")
    pairs = [("nrg_step", t.NrgStep), ("nrg_tangent_step", t.NrgTangentStep),
             ("nrg_onechqs_params", t.OneChQSParams), ("nrg_twochqs_params", t.TwoChQSParams),
             ("nrg_dmnrg_params", t.DmParams), ("nrg_pair_desc", t.PairDesc), ("nrg_stats", t.NrgStats),
             ("nrg_dm_shell", t.DmShellView)]
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "twochnrg_amd.h"', "int main(void) {"]
    expect = []
    for cname, cls in pairs:
        lines.append('  printf("%%zu\\n", sizeof(%s));' % cname)
        expect.append((cname, "sizeof", ctypes.sizeof(cls)))
        for f in cls._fields_:
            lines.append('  printf("%%zu\\n", offsetof(%s, %s));' % (cname, f[0]))
            expect.append((cname, f[0], getattr(cls, f[0]).offset))
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.check_call([cc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    got = [int(x) for x in subprocess.check_output([str(exe)]).split()]
    assert len(got) == len(expect)
    bad = [(e[0], e[1], e[2], g) for e, g in zip(expect, got) if e[2] != g]
    assert not bad, bad
