"""The .bbtc file format (PIGO-CSR-v2) and CSR validation.

Pinned against the byte layout of pigo::CSR::save / read_bin_
(PIGO csr.impl.hpp:440-512): written here with struct, independent of the
library, and cross-read by the oracle's reader.
"""
import os
import struct

import numpy as np
import pytest

import bbtc_amd as B
import oracle as O

MAGIC = b"PIGO-CSR-v2"


def write_raw(path, off, col, n=None, nrows=None, ncols=None, L=4, O_=4, m=None, magic=MAGIC, trunc=0):
    off = np.asarray(off, dtype=np.uint64)
    col = np.asarray(col, dtype=np.uint32)
    n = len(off) - 1 if n is None else n
    nrows = n if nrows is None else nrows
    ncols = n if ncols is None else ncols
    m = int(off[-1]) if m is None else m
    b = bytearray(magic)
    b += struct.pack("<BB", L, O_)
    b += struct.pack("<I", n)
    b += struct.pack("<I" if O_ == 4 else "<Q", m)
    b += struct.pack("<II", nrows, ncols)
    b += off.astype(np.uint32 if O_ == 4 else np.uint64).tobytes()
    b += col.tobytes()
    if trunc:
        b = b[:-trunc]
    with open(path, "wb") as f:
        f.write(bytes(b))


# triangle 0-1-2 plus pendant edge 2-3, degree-ordered upper form
TRI_OFF = [0, 2, 3, 3, 3]
TRI_COL = [1, 2, 2]


def test_header_layout_matches_pigo(tmp_path):
    g = B.Graph.from_csr(TRI_OFF, TRI_COL)
    p = str(tmp_path / "t.bbtc")
    g.save(p)
    raw = open(p, "rb").read()
    assert raw[:11] == MAGIC
    assert raw[11] == 4 and raw[12] == 4  # L_size, O_size
    n, m, nrows, ncols = struct.unpack_from("<IIII", raw, 13)
    assert (n, m, nrows, ncols) == (4, 3, 4, 4)
    off = np.frombuffer(raw, dtype=np.uint32, count=5, offset=29)
    col = np.frombuffer(raw, dtype=np.uint32, count=3, offset=29 + 20)
    assert off.tolist() == TRI_OFF and col.tolist() == TRI_COL
    assert len(raw) == 29 + 4 * 5 + 4 * 3


def test_round_trip_and_oracle_reader(tmp_path):
    g = B.Graph.rmat(10, 8, seed=4)
    p = str(tmp_path / "g.bbtc")
    g.save(p)
    h = B.Graph.load(p)
    assert np.array_equal(h.offsets(), g.offsets())
    assert np.array_equal(h.cols(), g.cols())
    n, ncols, off, col = O.load_bbtc(p)
    assert n == g.n and ncols == g.n
    assert np.array_equal(off, g.offsets()) and np.array_equal(col, g.cols())


def test_reads_64bit_offsets(tmp_path):
    p = str(tmp_path / "o8.bbtc")
    write_raw(p, TRI_OFF, TRI_COL, O_=8)
    g = B.Graph.load(p)
    assert g.offsets().tolist() == TRI_OFF and g.cols().tolist() == TRI_COL


def test_nrows_larger_than_n_pads_rows(tmp_path):
    # sarma::Matrix pads indptr up to nrows+1 (csr_matrix.hpp:139-145)
    p = str(tmp_path / "pad.bbtc")
    write_raw(p, TRI_OFF, TRI_COL, nrows=6, ncols=6)
    g = B.Graph.load(p)
    assert g.n == 6 and g.offsets().tolist() == TRI_OFF + [3, 3]


@pytest.mark.parametrize("case", ["magic", "L8", "O2", "trunc", "unsorted", "dup", "lower", "diag",
                                  "range", "offsets"])
def test_rejects_malformed(tmp_path, case):
    p = str(tmp_path / f"{case}.bbtc")
    if case == "magic":
        write_raw(p, TRI_OFF, TRI_COL, magic=b"PIGO-CSR-v1")
    elif case == "L8":
        write_raw(p, TRI_OFF, TRI_COL, L=8)
    elif case == "O2":
        write_raw(p, TRI_OFF, TRI_COL, O_=2)
    elif case == "trunc":
        write_raw(p, TRI_OFF, TRI_COL, trunc=3)
    elif case == "unsorted":
        write_raw(p, [0, 2, 3, 3, 3], [2, 1, 2])
    elif case == "dup":
        write_raw(p, [0, 2, 3, 3, 3], [1, 1, 2])
    elif case == "lower":
        write_raw(p, [0, 1, 2, 2, 2], [1, 0])
    elif case == "diag":
        write_raw(p, [0, 1, 1, 1, 1], [0])
    elif case == "range":
        write_raw(p, [0, 1, 1, 1, 1], [7])
    elif case == "offsets":
        write_raw(p, [0, 2, 1, 3, 3], [1, 2, 2])
    with pytest.raises(B.BbtcError) as e:
        B.Graph.load(p)
    assert e.value.status == 3  # BBTC_ERR_FORMAT


def test_missing_file_is_io_error(tmp_path):
    with pytest.raises(B.BbtcError) as e:
        B.Graph.load(str(tmp_path / "nope.bbtc"))
    assert e.value.status == 2


def test_from_csr_validates():
    with pytest.raises(B.BbtcError):
        B.Graph.from_csr([0, 1, 1], [0])  # diagonal entry
