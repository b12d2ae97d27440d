"""ctypes view of the CPU oracle (oracle/liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() to
check results, and by bench.py's cpu_baseline leg. The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborc.so")
REF_DIR = os.path.join(HERE, "_ref")
_lib = None


class _Csr(C.Structure):
    _fields_ = [("n", C.c_uint32), ("ncols", C.c_uint32), ("m", C.c_uint64),
                ("off", C.POINTER(C.c_uint64)), ("col", C.POINTER(C.c_uint32))]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} not built (make -C oracle)")
        L = C.CDLL(LIB_PATH)
        vp, u32, u64 = C.c_void_p, C.c_uint32, C.c_uint64
        L.orc_load_bbtc.argtypes = [C.c_char_p, C.POINTER(_Csr)]
        L.orc_load_bbtc.restype = C.c_int
        L.orc_free_csr.argtypes = [C.POINTER(_Csr)]
        L.orc_cuts_uni.argtypes = [u32, u32, vp]
        L.orc_cuts_n1d.argtypes = [vp, u32, u32, vp]
        L.orc_cuts_rac.argtypes = [vp, vp, u32, u32, vp]
        L.orc_cuts_pal.argtypes = [vp, vp, u32, u32, u64, u32, vp]
        L.orc_bcsr_build.argtypes = [vp, vp, u32, u64, vp, u32]
        L.orc_bcsr_build.restype = vp
        L.orc_bcsr_free.argtypes = [vp]
        L.orc_bcsr_tile_deg.argtypes = [vp, u32]
        L.orc_bcsr_tile_deg.restype = C.c_float
        L.orc_bcsr_tile_nnz.argtypes = [vp, u32]
        L.orc_bcsr_tile_nnz.restype = u32
        L.orc_bcsr_tile_height.argtypes = [vp, u32]
        L.orc_bcsr_tile_height.restype = u32
        L.orc_bcsr_row.argtypes = [vp, u32, u32, vp]
        L.orc_bcsr_row.restype = u32
        L.orc_tasks.argtypes = [vp, C.c_int, vp]
        L.orc_tasks.restype = u64
        L.orc_count.argtypes = [vp, vp, u64, C.c_int, vp]
        L.orc_count.restype = u64
        L.orc_count_rows.argtypes = [vp, vp, u64, C.c_int, vp, u32, u32]
        L.orc_count_rows.restype = u64
        L.orc_count_definition.argtypes = [vp, vp, u32, vp, u32, vp]
        L.orc_count_definition.restype = u64
        L.orc_brute_force.argtypes = [vp, vp, u32]
        L.orc_brute_force.restype = u64
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def load_bbtc(path):
    g = _Csr()
    rc = lib().orc_load_bbtc(path.encode(), C.byref(g))
    if rc != 0:
        lib().orc_free_csr(C.byref(g))
        raise ValueError(f"oracle could not read {path} (rc={rc})")
    off = np.ctypeslib.as_array(g.off, shape=(g.n + 1,)).copy()
    col = np.ctypeslib.as_array(g.col, shape=(max(g.m, 1),))[: g.m].copy()
    n, ncols = g.n, g.ncols
    lib().orc_free_csr(C.byref(g))
    return n, ncols, off, col


def cuts(kind, off, col, p, n=None, ncols=None):
    """Driver cuts (cuts[p] = n+1) for uni / n1d / pal / rac."""
    off = np.ascontiguousarray(off, dtype=np.uint64)
    col = np.ascontiguousarray(col, dtype=np.uint32)
    n = len(off) - 1 if n is None else n
    ncols = n if ncols is None else ncols
    out = np.zeros(p + 1, dtype=np.uint32)
    if kind == "uni":
        lib().orc_cuts_uni(n, p, _p(out))
    elif kind == "n1d":
        lib().orc_cuts_n1d(_p(off), n, p, _p(out))
    elif kind == "pal":
        lib().orc_cuts_pal(_p(off), _p(col), n, ncols, int(off[-1]), p, _p(out))
    elif kind == "rac":
        lib().orc_cuts_rac(_p(off), _p(col), n, p, _p(out))
    else:
        raise ValueError(kind)
    out[p] = n + 1
    return out


class RefLayout:
    """The reference's BlockCSR restated (block_csr.cpp:286-376)."""

    def __init__(self, off, col, cuts_):
        self.off = np.ascontiguousarray(off, dtype=np.uint64)
        self.col = np.ascontiguousarray(col, dtype=np.uint32)
        self.cuts = np.ascontiguousarray(cuts_, dtype=np.uint32)
        self.p = len(self.cuts) - 1
        self.n = len(self.off) - 1
        self._h = lib().orc_bcsr_build(_p(self.off), _p(self.col), self.n, int(self.off[-1]),
                                       _p(self.cuts), self.p)

    def tile_deg(self, t):
        return lib().orc_bcsr_tile_deg(self._h, t)

    def tile_nnz(self, t):
        return lib().orc_bcsr_tile_nnz(self._h, t)

    def tile_height(self, t):
        return lib().orc_bcsr_tile_height(self._h, t)

    def row(self, t, r):
        buf = np.zeros(max(1, self.n + 1), dtype=np.uint32)
        k = lib().orc_bcsr_row(self._h, t, r, _p(buf))
        return buf[:k].copy()

    def tasks(self, order="lex"):
        p = self.p
        cap = p * (p + 1) * (p + 2) // 6
        out = np.zeros((cap, 3), dtype=np.uint32)
        n = lib().orc_tasks(self._h, {"lex": 0, "desc": 1, "asc": 2}[order], _p(out))
        return out[:n]

    def count(self, tasks, nthreads=0):
        t = np.ascontiguousarray(np.asarray(tasks, dtype=np.uint32).reshape(-1, 3))
        per = np.zeros(len(t), dtype=np.uint64)
        tot = lib().orc_count(self._h, _p(t), len(t), nthreads, _p(per))
        return per, tot

    def count_rows(self, tasks, stride, phase=0, nthreads=0):
        """Counts of only the 64-row blocks b % stride == phase of each
        task's band rows (a systematic sample of every task's rows, for
        timing)."""
        t = np.ascontiguousarray(np.asarray(tasks, dtype=np.uint32).reshape(-1, 3))
        per = np.zeros(len(t), dtype=np.uint64)
        tot = lib().orc_count_rows(self._h, _p(t), len(t), nthreads, _p(per), stride, phase)
        return per, tot

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.orc_bcsr_free(self._h)
            self._h = None


def count_definition(off, col, cuts_):
    """p^3 array of per-task counts from the triangle definition."""
    off = np.ascontiguousarray(off, dtype=np.uint64)
    col = np.ascontiguousarray(col, dtype=np.uint32)
    c = np.ascontiguousarray(cuts_, dtype=np.uint32)
    p = len(c) - 1
    cube = np.zeros(p * p * p, dtype=np.uint64)
    tot = lib().orc_count_definition(_p(off), _p(col), len(off) - 1, _p(c), p, _p(cube))
    return cube.reshape(p, p, p), tot


def brute_force(off, col):
    off = np.ascontiguousarray(off, dtype=np.uint64)
    col = np.ascontiguousarray(col, dtype=np.uint32)
    return lib().orc_brute_force(_p(off), _p(col), len(off) - 1)
