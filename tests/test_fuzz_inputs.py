"""Damaged and random inputs: every reader either returns a valid graph (which
then counts without fault) or fails with BbtcError, never a crash.

Seeded byte mutations of a .bbtc and a .bcsr file (flipped bytes, truncation,
header bytes), and random text files for the edge-list and Matrix Market
reader (vertex ids kept below 10^5 so that a valid file stays small). A
damaged .bcsr header fails before anything is allocated for it."""
import re

import numpy as np
import pytest

import bbtc_amd as B


def padded_rows(b):
    """nrows and ncols of a .bbtc header with 4-byte labels (0 when short):
    a file may pad its row count beyond n, so a damaged nrows can make a
    valid graph of billions of empty rows."""
    at = 17 + (8 if len(b) > 12 and b[12] == 8 else 4)
    if len(b) < at + 8:
        return 0
    return max(int.from_bytes(b[at:at + 4], "little"), int.from_bytes(b[at + 4:at + 8], "little"))


def mutate(rng, base):
    b = bytearray(base)
    kind = rng.integers(0, 3)
    if kind == 0:
        for _ in range(int(rng.integers(1, 8))):
            b[int(rng.integers(0, len(b)))] = int(rng.integers(0, 256))
    elif kind == 1:
        b = b[: int(rng.integers(0, len(b)))]
    else:
        b[int(rng.integers(0, min(len(b), 64)))] = int(rng.integers(0, 256))
    return bytes(b)


@pytest.mark.parametrize("suffix", [".bbtc", ".bcsr"])
def test_binary_readers(tmp_path, suffix):
    g = B.Graph.rmat(8, 8, seed=1)
    src = str(tmp_path / ("g" + suffix))
    if suffix == ".bbtc":
        g.save(src)
    else:
        B.Layout(g, g.partition("pal", 4)).save_bcsr(src)
    base = open(src, "rb").read()
    rng = np.random.default_rng(7)
    out = str(tmp_path / ("m" + suffix))
    failed = 0
    for _ in range(300):
        b = mutate(rng, base)
        if suffix == ".bbtc" and padded_rows(b) > (1 << 20):
            continue
        with open(out, "wb") as f:
            f.write(b)
        try:
            if suffix == ".bbtc":
                h = B.Graph.load(out)
                L = B.Layout(h, h.partition("uni", 2))
            else:
                L = B.Layout.load_bcsr(out)
            L.count_cpu(L.tasks("lex"))
        except B.BbtcError:
            failed += 1
    assert failed > 0  # the mutations do reach the checks


def test_text_reader(tmp_path):
    rng = np.random.default_rng(11)
    alpha = np.frombuffer(b"0123456789 \t\n\r%#-+.eE,MatrixMarket", dtype=np.uint8)
    for it in range(300):
        n = int(rng.integers(0, 400))
        if rng.integers(0, 2):
            b = bytes(alpha[rng.integers(0, len(alpha), n)])
        else:
            b = bytes(rng.integers(0, 256, n, dtype=np.uint8))
        b = re.sub(rb"[0-9]{6,}", lambda m: m.group(0)[:5] + b" ", b)
        name = str(tmp_path / ("t.mtx" if rng.integers(0, 4) == 0 else "t.txt"))
        if name.endswith(".mtx") and rng.integers(0, 2):
            b = b"%%MatrixMarket matrix coordinate pattern general\n" + b
        with open(name, "wb") as f:
            f.write(b)
        try:
            g = B.Graph.load_text(name)
        except B.BbtcError:
            continue
        L = B.Layout(g, g.partition("uni", 2))
        L.count_cpu(L.tasks("lex"))
