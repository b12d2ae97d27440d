"""Deterministic text graph files for the ingest fixtures: edge lists and
Matrix Market files that exercise the reader's grammar (comments, tabs, CRLF,
weights, signs, a last line without newline, duplicates, self loops,
1-based ids)."""
import numpy as np


def _pairs(rng, n, m, dup_frac=0.0, loops=0, reciprocal_frac=0.0):
    u = rng.integers(0, n, size=m)
    v = rng.integers(0, n, size=m)
    p = list(zip(u.tolist(), v.tolist()))
    if reciprocal_frac:
        k = int(reciprocal_frac * m)
        p += [(b, a) for a, b in p[:k]]
    if dup_frac:
        k = int(dup_frac * m)
        p += p[:k]
    p += [(int(x), int(x)) for x in rng.integers(0, n, size=loops)]
    rng.shuffle(p)
    return p


def _simple_pairs(rng, n, m):
    """Distinct undirected pairs, no loops, each listed once in a random direction."""
    seen = set()
    out = []
    while len(out) < m:
        a, b = (int(x) for x in rng.integers(0, n, size=2))
        if a == b or (min(a, b), max(a, b)) in seen:
            continue
        seen.add((min(a, b), max(a, b)))
        out.append((a, b) if rng.random() < 0.5 else (b, a))
    return out


def make(case: str) -> bytes:
    rng = np.random.default_rng(abs(hash_case(case)))
    if case == "snap_tabs":
        p = _pairs(rng, 300, 2000, dup_frac=0.05, loops=7, reciprocal_frac=0.2)
        s = "# Directed graph: synthetic\n# Nodes: 300 Edges: %d\n# FromNodeId\tToNodeId\n" % len(p)
        return (s + "".join(f"{a}\t{b}\n" for a, b in p)).encode()
    if case == "crlf_no_final_newline":
        p = _pairs(rng, 500, 3000, loops=3)
        body = "\r\n".join(f"  {a}   {b} " if i % 3 else f"{a} {b}" for i, (a, b) in enumerate(p))
        return body.encode()  # the last entry ends at end of file
    if case == "weighted_3col":
        p = _pairs(rng, 400, 2500, reciprocal_frac=0.3)
        w = rng.random(len(p))
        return "".join(f"{a} {b} {x:.6f}\n" for (a, b), x in zip(p, w)).encode()
    if case == "inline_comments_signs":
        p = _pairs(rng, 200, 1500, dup_frac=0.02)
        out = ["% leading percent comment\n"]
        for i, (a, b) in enumerate(p):
            if i % 50 == 0:
                out.append("# a comment line\n")
            if i % 17 == 0:
                out.append(f"-{a} +{b}\n")
            elif i % 13 == 0:
                out.append(f"{a},{b} # trailing note\n")
            else:
                out.append(f"{a} {b}\n")
        return "".join(out).encode()
    if case == "mtx_general_pattern":
        p = _pairs(rng, 350, 2200, reciprocal_frac=0.1, loops=2)
        s = "%%MatrixMarket matrix coordinate pattern general\n% generated\n%\n"
        s += f"350 350 {len(p)}\n"
        return (s + "".join(f"{a + 1} {b + 1}\n" for a, b in p)).encode()
    if case == "mtx_symmetric_real":
        p = [(max(a, b), min(a, b)) for a, b in _simple_pairs(rng, 600, 4000)]
        s = "%%MatrixMarket matrix coordinate real symmetric\n"
        s += f"600 600 {len(p)}\n"
        return (s + "".join(f"{a + 1} {b + 1} {rng.random():.3e}\n" for a, b in p)).encode()
    if case == "simple_20k":
        p = _simple_pairs(rng, 20000, 120000)
        return "".join(f"{a} {b}\n" for a, b in p).encode()
    if case == "simple_hubs":
        # a few hubs joined to many vertices: wide degree spread, many ties
        p = set()
        for h in range(12):
            for x in rng.choice(5000, size=600, replace=False).tolist():
                if x != h:
                    p.add((min(h, x), max(h, x)))
        p |= {(min(a, b), max(a, b)) for a, b in _simple_pairs(rng, 5000, 15000)}
        # smaller id first: the largest id appears as a destination (the
        # reference reads past its transpose when a source id exceeds every
        # destination id)
        return "".join(f"{a} {b}\n" for a, b in sorted(p)).encode()
    if case == "big_chunks":
        # over a megabyte: the reader splits it over threads at line boundaries
        p = _simple_pairs(rng, 100000, 150000)
        return "".join(f"{a}\t{b}\t1\n" for a, b in p).encode()
    raise KeyError(case)


def hash_case(case: str) -> int:
    h = 1469598103934665603
    for ch in case.encode():
        h = ((h ^ ch) * 1099511628211) % (1 << 61)
    return h


CASES = ["snap_tabs", "crlf_no_final_newline", "weighted_3col", "inline_comments_signs",
         "mtx_general_pattern", "mtx_symmetric_real", "simple_20k", "simple_hubs", "big_chunks"]
SIMPLE = {"mtx_symmetric_real", "simple_20k", "simple_hubs", "big_chunks"}


def filename(case: str) -> str:
    return case + (".mtx" if case.startswith("mtx_") else ".txt")
