"""Register budget of the counting kernels: count_probe must keep 4 waves per
SIMD (at most 128 VGPRs: 1,024 resident blocks of 256 threads) and its
lane-form instantiation 8, as must the row form's count_rows, since all are
latency-bound and their speed falls with occupancy; none may spill. Read from the compiler's resource
remarks for gfx950, with the library's own flags; no GPU needed."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC) and not shutil.which("hipcc"), reason="hipcc not installed")
def test_counting_kernels_keep_their_occupancy(tmp_path):
    src = os.path.join(ROOT, "bbtc_amd", "csrc", "device", "device.hip")
    cmd = [HIPCC, "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950", "-I", os.path.join(ROOT, "include"),
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "--offload-device-only",
           "-Rpass-analysis=kernel-resource-usage", "-c", src, "-o", str(tmp_path / "d.o")]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    report, name = {}, None
    for line in r.stderr.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            name = m.group(1)
            continue
        m = re.search(r"(VGPRs|Occupancy \[waves/SIMD\]|VGPRs Spill): (\d+)", line)
        if m and name:
            report.setdefault(name, {})[m.group(1)] = int(m.group(2))
    probe = report["_ZN4bbtc11count_probeENS_9CountArgsE"]
    lanes = report["_ZN4bbtc17count_probe_lanesENS_9CountArgsE"]
    rows = report["_ZN4bbtc10count_rowsENS_7RowArgsE"]
    assert probe["Occupancy [waves/SIMD]"] >= 4 and probe["VGPRs Spill"] == 0, probe
    assert lanes["Occupancy [waves/SIMD]"] >= 8 and lanes["VGPRs Spill"] == 0, lanes
    assert rows["Occupancy [waves/SIMD]"] >= 8 and rows["VGPRs Spill"] == 0, rows
