"""Live cross-check against the reference's own CPU binary where it has been
built from its sources (oracle/_ref/bbtc_cpu, `make -C oracle ref`, only in
the container that holds the reference tree): the same .bbtc file and
partitioner give the same triangle count as the product's CLI on its CPU path
and its device-free C-ABI count, and the reference reads the .bcsr files the
product writes. Skipped where the binary is absent."""
import os
import re
import subprocess

import pytest

import bbtc_amd as B

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
REF = os.path.join(ROOT, "oracle", "_ref", "bbtc_cpu")


@pytest.mark.skipif(not os.path.exists(REF), reason="reference CPU binary not built here")
@pytest.mark.parametrize("spec,part,ncut", [(("rmat", 14, 8, 5), "pal", 8), (("rmat", 13, 16, 6), "n1d", 5),
                                            (("gnm", 30000, 240000, 7), "uni", 4),
                                            (("rmat", 12, 24, 8), "rac", 6)])
def test_reference_binary_agrees(tmp_path, spec, part, ncut):
    kind, a, b, seed = spec
    g = B.Graph.rmat(a, b, seed=seed) if kind == "rmat" else B.Graph.gnm(a, b, seed)
    path = str(tmp_path / "g.bbtc")
    g.save(path)
    ref = subprocess.run([REF, "--graph", path, "--partitioner", part, "--ncut", str(ncut), "--repeat", "1"],
                         capture_output=True, text=True, timeout=300)
    assert ref.returncode == 0, ref.stderr
    want = int(re.search(r"Number of Triangles: (\d+)", ref.stdout).group(1))
    ours = subprocess.run([B.CLI_PATH, "--graph", path, "--partitioner", part, "--ncut", str(ncut), "--ngpu", "0",
                           "--repeat", "1"], capture_output=True, text=True, timeout=300)
    assert ours.returncode == 0, ours.stderr
    assert f"Number of Triangles: {want}" in ours.stdout
    L = B.Layout(g, g.partition(part, ncut))
    assert L.count_cpu(L.tasks("desc"))[1] == want
    # the report's header lines agree too
    for key in ("Number of Vertices", "Number of Edges", "Number of Cuts", "Number of Tasks"):
        line = re.search(rf"{key}: (\d+)", ref.stdout).group(0)
        assert line in ours.stdout, key


@pytest.mark.skipif(not os.path.exists(REF), reason="reference CPU binary not built here")
@pytest.mark.parametrize("spec,part,ncut", [(("rmat", 14, 8, 5), "pal", 8), (("gnm", 30000, 240000, 7), "uni", 4),
                                            (("rmat", 12, 24, 8), "rac", 6)])
def test_reference_reads_written_bcsr(tmp_path, spec, part, ncut):
    # BlockCSR::DeSerialize (block_csr.cpp:215-255) on a file this library
    # wrote: the reference counts the same triangles with the same cuts and
    # tasks, and reports the header's N (the vertices plus the end slot)
    kind, a, b, seed = spec
    g = B.Graph.rmat(a, b, seed=seed) if kind == "rmat" else B.Graph.gnm(a, b, seed)
    L = B.Layout(g, g.partition(part, ncut))
    path = str(tmp_path / "g.bcsr")
    L.save_bcsr(path)
    ref = subprocess.run([REF, "--graph", path, "--repeat", "1"], capture_output=True, text=True, timeout=300)
    assert ref.returncode == 0, ref.stderr
    want = L.count_cpu(L.tasks("desc"))[1]
    assert f"Number of Triangles: {want}" in ref.stdout
    assert f"Number of Cuts: {ncut}" in ref.stdout and f"Number of Edges: {g.m}" in ref.stdout
    assert f"Number of Vertices: {g.n + 1}" in ref.stdout
