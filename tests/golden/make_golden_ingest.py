"""Write golden vectors for text ingest from the reference's own code (run
once, offline).

Each case's text file comes from tests/golden/textgen.py. The reference's
CPU build (oracle/_ref/bbtc_cpu, built from its sources by oracle/Makefile)
is run on it; its driver converts the text and writes <file>.bbtc
(utils.hpp:171-268) and prints the triangle count. Only the SHA-256 of that
.bbtc, its n and m, and the count are committed (tests/golden/ingest.json).

Usage: python tests/golden/make_golden_ingest.py
Needs /root/reference (for `make -C oracle ref`).
"""
from __future__ import annotations

import hashlib
import json
import os
import re
import struct
import subprocess
import sys
import tempfile

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
sys.path.insert(0, ROOT)

from tests.golden import textgen  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(ROOT, "oracle", "_ref", "bbtc_cpu")


def main():
    if not os.path.exists(REF):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    out = []
    with tempfile.TemporaryDirectory() as tmp:
        for case in textgen.CASES:
            path = os.path.join(tmp, textgen.filename(case))
            with open(path, "wb") as f:
                f.write(textgen.make(case))
            r = subprocess.run([REF, "--graph", path, "--partitioner", "uni", "--ncut", "2", "--repeat", "1"],
                               capture_output=True, text=True, timeout=600)
            if r.returncode != 0:
                raise RuntimeError(f"{case}: reference failed: {r.stderr}")
            tri = int(re.search(r"Number of Triangles: (\d+)", r.stdout).group(1))
            data = open(path + ".bbtc", "rb").read()
            n, m = struct.unpack_from("<II", data, 13)
            rec = dict(case=case, file=textgen.filename(case), sha256=hashlib.sha256(data).hexdigest(),
                       bytes=len(data), n=n, m=m, triangles_reference=tri,
                       simple=case in textgen.SIMPLE)
            print(rec, flush=True)
            out.append(rec)
    with open(os.path.join(HERE, "ingest.json"), "w") as f:
        json.dump(dict(source="oracle/_ref/bbtc_cpu (reference CPU build from its sources): "
                              "text -> .bbtc conversion of utils::ReadGraph",
                       cases=out), f, indent=1)


if __name__ == "__main__":
    main()
