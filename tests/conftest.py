import glob
import json
import os
import sys

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN_DIR = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")
    config.addinivalue_line("markers", "slow: larger CPU cases")


def load_golden(name=None):
    """Per-graph fixtures (cuts and per-task counts); other fixture files,
    such as the text-ingest hashes, have no "runs" and are skipped. Fixtures
    of full-size graphs ("full_size": RMAT scale 26) are returned only when
    asked for by name."""
    out = []
    for path in sorted(glob.glob(os.path.join(GOLDEN_DIR, "*.json"))):
        with open(path) as f:
            d = json.load(f)
        if "runs" not in d:
            continue
        if name is None and not d.get("full_size") or d["name"] == name:
            out.append(d)
    return out


def graph_from_spec(spec):
    import bbtc_amd as B
    if spec["kind"] == "rmat":
        return B.Graph.rmat(spec["scale"], spec["ef"], spec["a"], spec["b"], spec["c"], spec["seed"])
    if spec["kind"] == "gnm":
        return B.Graph.gnm(spec["n"], spec["m"], spec["seed"])
    raise ValueError(spec)


@pytest.fixture(scope="session")
def golden():
    return load_golden()


def golden_ids():
    return [d["name"] for d in load_golden()]
