"""The host side of the library (the sources that need no GPU: .bbtc and text
I/O, generators, partitioners, layouts, the CPU count, .bcsr files and their
C-ABI) built with g++ under AddressSanitizer (with its leak check) and
UndefinedBehaviorSanitizer, and driven through the C-ABI by
tests/host/sanitize_driver.cpp. Any report fails the test."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
HOST = os.path.join(ROOT, "bbtc_amd", "csrc", "host")
SOURCES = ["csr_io", "ingest", "bcsr", "generate", "partition", "layout", "cpu_count", "host_capi"]


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not installed")
def test_host_code_is_clean_under_asan_and_ubsan(tmp_path):
    exe = tmp_path / "drv"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fopenmp", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
           "-fno-sanitize-recover=undefined", "-I", os.path.join(ROOT, "include"), "-I", HOST]
    cmd += [os.path.join(HOST, s + ".cpp") for s in SOURCES]
    cmd += [os.path.join(ROOT, "tests", "host", "sanitize_driver.cpp"), "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1", OMP_NUM_THREADS="4")
    r = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True, timeout=900, env=env)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
    assert r.stdout.strip().endswith("done")
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
