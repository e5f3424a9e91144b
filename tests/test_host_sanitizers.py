"""The library's plain C++ host units (truncation: SetDegen, Ecut, SetKept;
the reference's .bin reader and writer) built with g++ under AddressSanitizer
and UBSan and run on random spectra and on written, truncated and corrupted
files (tests/host_asan/host_asan_main.cpp)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_host_units_under_asan_ubsan(tmp_path):
    exe = str(tmp_path / "host_asan")
    subprocess.run(["make", "-s", "-B", "-C", os.path.join(HERE, "host_asan"), "OUT=" + exe], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe, str(tmp_path)], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                       timeout=600, env=env)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "host checks passed" in r.stdout
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
