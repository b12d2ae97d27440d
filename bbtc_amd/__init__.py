"""Python view of the MI355X block-based triangle counter (libbbtc.so).

A thin ctypes layer over the C-ABI in include/bbtc.h. It mirrors the
reference's objects for this path (paths relative to the GT-TDAlab/bbTC tree):

    Graph.load / Graph.rmat / Graph.gnm   sarma::Matrix(path) (.bbtc input)
    Graph.partition(kind, p)              utils.hpp:277-288 (uni | n1d | pal)
    Layout(graph, cuts)                   BlockCSR(cuts, x_adj, adj)
    Layout.tasks(order)                   tasks::TaskList + tasks::SortTasks
    Layout.count_cpu(tasks)               BBTc::NoOfTriangles, CPU build
    Device / DeviceGraph.count(tasks)     BBTc::NoOfTriangles, GPU build

All counting happens in the native library; if libbbtc.so is missing every
entry point raises instead of computing anything in Python.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbbtc.so")
CLI_PATH = os.path.join(_HERE, "bbtc")

PART = {"uni": 0, "n1d": 1, "pal": 2, "rac": 3}
ORDER = {"lex": 0, "desc": 1, "asc": 2}
KERNEL = {"auto": 0, "merge": 1, "bitmap": 2, "dense": 3, "rdense": 4, "lane": 5, "rows": 6}
BUILD_HOST_ARRAYS = 1  # bbtc_graph_build: copy the layout's arrays back to the host


class BbtcError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"bbtc status {status}: {msg}")
        self.status = status


class bbtc_task(C.Structure):
    _fields_ = [("i", C.c_uint32), ("j", C.c_uint32), ("k", C.c_uint32)]


class bbtc_tile_info(C.Structure):
    _fields_ = [("nnz", C.c_uint64), ("height", C.c_uint32), ("width", C.c_uint32),
                ("nz_rows", C.c_uint32), ("avg_degree", C.c_float), ("max_row", C.c_uint32)]


class bbtc_opts(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64)]


class bbtc_server_opts(C.Structure):
    _fields_ = [("heap_bytes", C.c_uint64), ("quantum_items", C.c_uint64), ("max_inflight", C.c_uint32),
                ("max_segments", C.c_uint32), ("sequential", C.c_uint32), ("lane_form", C.c_uint32),
                ("launch_us", C.c_uint32), ("oldest_pct", C.c_uint32), ("list_chunks", C.c_uint32),
                ("defer_steps", C.c_uint32)]


class bbtc_server_result(C.Structure):
    _fields_ = [("id", C.c_uint64), ("total", C.c_uint64), ("status", C.c_int32), ("launches", C.c_uint32),
                ("bytes", C.c_uint64), ("arrival_s", C.c_double), ("admit_s", C.c_double),
                ("ready_s", C.c_double), ("start_s", C.c_double), ("done_s", C.c_double)]


class bbtc_server_stats(C.Structure):
    _fields_ = [("launches", C.c_uint64), ("segments", C.c_uint64), ("items", C.c_uint64),
                ("peak_resident_bytes", C.c_uint64), ("heap_bytes", C.c_uint64),
                ("peak_active", C.c_uint32), ("peak_segments", C.c_uint32),
                ("busy_s", C.c_double), ("upload_s", C.c_double), ("row_form_requests", C.c_uint64)]


TASK_DTYPE = np.dtype([("i", np.uint32), ("j", np.uint32), ("k", np.uint32)])

_lib = None


def _declare(lib):
    vp, u32, u64, i32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int
    pvp = C.POINTER(C.c_void_p)
    st = C.c_int
    sig = {
        "bbtc_strerror": (C.c_char_p, [i32]),
        "bbtc_last_error": (C.c_char_p, []),
        "bbtc_hgraph_load": (st, [C.c_char_p, pvp]),
        "bbtc_hgraph_save": (st, [vp, C.c_char_p]),
        "bbtc_convert_text": (st, [C.c_char_p, C.c_char_p, i32]),
        "bbtc_hgraph_load_text": (st, [C.c_char_p, pvp]),
        "bbtc_hgraph_from_csr": (st, [u32, u32, u64, vp, vp, pvp]),
        "bbtc_hgraph_from_edges": (st, [u32, u64, vp, vp, pvp]),
        "bbtc_hgraph_gen_rmat": (st, [u32, u32, C.c_double, C.c_double, C.c_double, u64, pvp]),
        "bbtc_hgraph_gen_gnm": (st, [u32, u64, u64, pvp]),
        "bbtc_hgraph_gen_rmat_device": (st, [vp, u32, u32, C.c_double, C.c_double, C.c_double, u64, pvp]),
        "bbtc_hgraph_gen_gnm_device": (st, [vp, u32, u64, u64, pvp]),
        "bbtc_hgraph_info": (None, [vp, C.POINTER(u32), C.POINTER(u32), C.POINTER(u64)]),
        "bbtc_hgraph_offsets": (vp, [vp]),
        "bbtc_hgraph_cols": (vp, [vp]),
        "bbtc_hgraph_free": (None, [vp]),
        "bbtc_partition": (st, [vp, i32, u32, vp]),
        "bbtc_layout_build": (st, [vp, vp, u32, pvp]),
        "bbtc_layout_ncut": (u32, [vp]),
        "bbtc_layout_cuts": (vp, [vp]),
        "bbtc_layout_tile": (st, [vp, u32, u32, C.POINTER(bbtc_tile_info)]),
        "bbtc_layout_device_bytes": (u64, [vp]),
        "bbtc_layout_arrays": (st, [vp, pvp, C.POINTER(u64), pvp, C.POINTER(u64), pvp, C.POINTER(u64)]),
        "bbtc_layout_save_bcsr": (st, [vp, C.c_char_p]),
        "bbtc_layout_load_bcsr": (st, [C.c_char_p, pvp, pvp]),
        "bbtc_layout_free": (None, [vp]),
        "bbtc_task_count": (u64, [u32]),
        "bbtc_tasks_build": (st, [vp, i32, vp, u64, C.POINTER(u64)]),
        "bbtc_tasks_bytes": (st, [vp, vp, u64, u32, vp, C.POINTER(u64), C.POINTER(u64)]),
        "bbtc_tasks_stats": (st, [vp, vp, u64, u32, vp]),
        "bbtc_count_cpu": (st, [vp, vp, u64, i32, vp, C.POINTER(u64)]),
        "bbtc_device_count": (st, [C.POINTER(i32)]),
        "bbtc_ctx_create": (st, [i32, C.POINTER(bbtc_opts), pvp]),
        "bbtc_ctx_destroy": (None, [vp]),
        "bbtc_ctx_bytes_in_use": (u64, [vp]),
        "bbtc_ctx_synchronize": (st, [vp]),
        "bbtc_ctx_read_bandwidth": (st, [vp, u64, i32, vp]),
        "bbtc_graph_upload": (st, [vp, vp, pvp]),
        "bbtc_layout_pin": (st, [vp, vp]),
        "bbtc_tasks_shard": (st, [vp, vp, u64, u32, vp, vp]),
        "bbtc_layout_pinned": (i32, [vp]),
        "bbtc_graph_build": (st, [vp, vp, vp, u32, u32, pvp, pvp]),
        "bbtc_graph_tasks_stats": (st, [vp, vp, u64, u32, vp]),
        "bbtc_graph_bytes": (u64, [vp]),
        "bbtc_graph_release": (None, [vp]),
        "bbtc_plan_create": (st, [vp, vp, vp, u64, i32, pvp]),
        "bbtc_plan_create_tuned": (st, [vp, vp, vp, u64, i32, vp, pvp]),
        "bbtc_plan_launch": (st, [vp, vp]),
        "bbtc_plan_results": (st, [vp, vp, C.POINTER(u64)]),
        "bbtc_plan_time": (st, [vp, i32, vp]),
        "bbtc_plan_bitmap_tasks": (u64, [vp]),
        "bbtc_plan_time_tasks": (st, [vp, vp]),
        "bbtc_plan_task_kinds": (st, [vp, vp]),
        "bbtc_batch_task_kinds": (st, [vp, vp]),
        "bbtc_plan_profile": (st, [vp, vp]),
        "bbtc_autotune": (st, [vp, vp, vp, u64, vp, vp]),
        "bbtc_plan_destroy": (None, [vp]),
        "bbtc_count": (st, [vp, vp, vp, u64, i32, vp, C.POINTER(u64), vp]),
        "bbtc_count_noalloc": (st, [vp, vp, vp, u64, i32, vp, C.POINTER(u64), vp]),
        "bbtc_count_hybrid": (st, [vp, vp, vp, vp, u64, u64, i32, i32, i32, vp, vp, C.POINTER(u64),
                                   C.POINTER(u64)]),
        "bbtc_layouts_build": (st, [vp, u64, vp, i32, vp]),
        "bbtc_tune_ncut": (st, [vp, vp, u64, vp, u32, i32, i32, vp, vp]),
        "bbtc_rule_ncut": (u32, [vp, u64]),
        "bbtc_batch_create": (st, [vp, vp, u64, i32, i32, pvp]),
        "bbtc_batch_ntasks": (u64, [vp]),
        "bbtc_batch_bytes": (u64, [vp]),
        "bbtc_batch_tasks": (st, [vp, vp, vp]),
        "bbtc_batch_autotune": (st, [vp, vp]),
        "bbtc_batch_set_forms": (st, [vp, i32, vp]),
        "bbtc_batch_capture": (st, [vp]),
        "bbtc_batch_launch": (st, [vp]),
        "bbtc_batch_results": (st, [vp, vp, vp, vp]),
        "bbtc_batch_destroy": (None, [vp]),
        "bbtc_server_create": (st, [vp, C.POINTER(bbtc_server_opts), pvp]),
        "bbtc_server_reset_clock": (st, [vp]),
        "bbtc_server_now": (C.c_double, [vp]),
        "bbtc_server_submit": (st, [vp, vp, vp, u64, i32, C.c_double, vp, C.POINTER(u64)]),
        "bbtc_server_poll": (st, [vp, C.POINTER(bbtc_server_result), u64, i32, C.POINTER(u64)]),
        "bbtc_server_wait_idle": (st, [vp]),
        "bbtc_server_get_stats": (st, [vp, C.POINTER(bbtc_server_stats)]),
        "bbtc_server_destroy": (None, [vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


def lib():
    """Load libbbtc.so (built by __graft_entry__.build / make); raise if absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        _lib = _declare(C.CDLL(LIB_PATH))
    return _lib


def _check(status: int):
    if status != 0:
        L = lib()
        raise BbtcError(status, f"{L.bbtc_strerror(status).decode()}: {L.bbtc_last_error().decode()}")


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def tasks_array(tasks) -> np.ndarray:
    """(n,3) integers or a TASK_DTYPE array -> contiguous TASK_DTYPE array."""
    if isinstance(tasks, np.ndarray) and tasks.dtype == TASK_DTYPE:
        return np.ascontiguousarray(tasks)
    a = np.asarray(tasks, dtype=np.uint32).reshape(-1, 3)
    out = np.empty(len(a), dtype=TASK_DTYPE)
    out["i"], out["j"], out["k"] = a[:, 0], a[:, 1], a[:, 2]
    return out


class Graph:
    """Degree-ordered upper-triangular CSR held by the native library."""

    def __init__(self, handle):
        self._h = handle

    @staticmethod
    def _make(fn, *args):
        h = C.c_void_p()
        _check(fn(*args, C.byref(h)))
        return Graph(h)

    @classmethod
    def load(cls, path: str) -> "Graph":
        return cls._make(lib().bbtc_hgraph_load, path.encode())

    @classmethod
    def load_text(cls, path: str) -> "Graph":
        """Edge list or Matrix Market text, converted as the reference does
        (duplicate entries dropped)."""
        return cls._make(lib().bbtc_hgraph_load_text, path.encode())

    @classmethod
    def rmat(cls, scale, edge_factor, a=0.57, b=0.19, c=0.19, seed=1) -> "Graph":
        return cls._make(lib().bbtc_hgraph_gen_rmat, scale, edge_factor, a, b, c, seed)

    @classmethod
    def gnm(cls, n, m, seed) -> "Graph":
        return cls._make(lib().bbtc_hgraph_gen_gnm, n, m, seed)

    @classmethod
    def rmat_device(cls, dev, scale, edge_factor, a=0.57, b=0.19, c=0.19, seed=1) -> "Graph":
        """Same graph as rmat(), generated on the GPU of `dev`."""
        return cls._make(lib().bbtc_hgraph_gen_rmat_device, dev._h, scale, edge_factor, a, b, c, seed)

    @classmethod
    def gnm_device(cls, dev, n, m, seed) -> "Graph":
        """Same graph as gnm(), generated on the GPU of `dev`."""
        return cls._make(lib().bbtc_hgraph_gen_gnm_device, dev._h, n, m, seed)

    @classmethod
    def from_csr(cls, offsets, cols, n=None, ncols=None) -> "Graph":
        off = np.ascontiguousarray(offsets, dtype=np.uint64)
        col = np.ascontiguousarray(cols, dtype=np.uint32)
        n = len(off) - 1 if n is None else n
        ncols = n if ncols is None else ncols
        return cls._make(lib().bbtc_hgraph_from_csr, n, ncols, int(off[-1]) if len(off) else 0,
                         _ptr(off), _ptr(col))

    @classmethod
    def from_edges(cls, n, src, dst) -> "Graph":
        s = np.ascontiguousarray(src, dtype=np.uint32)
        d = np.ascontiguousarray(dst, dtype=np.uint32)
        return cls._make(lib().bbtc_hgraph_from_edges, n, len(s), _ptr(s), _ptr(d))

    def save(self, path: str):
        _check(lib().bbtc_hgraph_save(self._h, path.encode()))

    @property
    def info(self):
        n, nc, m = C.c_uint32(), C.c_uint32(), C.c_uint64()
        lib().bbtc_hgraph_info(self._h, C.byref(n), C.byref(nc), C.byref(m))
        return n.value, nc.value, m.value

    @property
    def n(self):
        return self.info[0]

    @property
    def m(self):
        return self.info[2]

    def offsets(self) -> np.ndarray:
        n = self.n
        p = lib().bbtc_hgraph_offsets(self._h)
        return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint64)), shape=(n + 1,)).copy()

    def cols(self) -> np.ndarray:
        m = self.m
        if m == 0:
            return np.zeros(0, dtype=np.uint32)
        p = lib().bbtc_hgraph_cols(self._h)
        return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint32)), shape=(m,)).copy()

    def partition(self, kind: str, p: int) -> np.ndarray:
        cuts = np.zeros(p + 1, dtype=np.uint32)
        _check(lib().bbtc_partition(self._h, PART[kind], p, _ptr(cuts)))
        return cuts

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.bbtc_hgraph_free(self._h)
            self._h = None


@dataclass
class TileInfo:
    nnz: int
    height: int
    width: int
    nz_rows: int
    avg_degree: float
    max_row: int


class Layout:
    """Block CSR of a graph for given cuts (BlockCSR)."""

    def __init__(self, graph: Graph, cuts):
        self.graph = graph
        c = np.ascontiguousarray(cuts, dtype=np.uint32)
        h = C.c_void_p()
        _check(lib().bbtc_layout_build(graph._h, _ptr(c), len(c) - 1, C.byref(h)))
        self._h = h
        self.p = len(c) - 1
        self.cuts = c.copy()

    @classmethod
    def load_bcsr(cls, path: str) -> "Layout":
        """A stored block layout (.bcsr) and its graph, with the stored cuts."""
        gh, lh = C.c_void_p(), C.c_void_p()
        _check(lib().bbtc_layout_load_bcsr(path.encode(), C.byref(gh), C.byref(lh)))
        L = cls.__new__(cls)
        L.graph = Graph(gh)
        L._h = lh
        L.p = int(lib().bbtc_layout_ncut(lh))
        ptr = C.cast(lib().bbtc_layout_cuts(lh), C.POINTER(C.c_uint32))
        L.cuts = np.array([ptr[x] for x in range(L.p + 1)], dtype=np.uint32)
        return L

    def save_bcsr(self, path: str):
        _check(lib().bbtc_layout_save_bcsr(self._h, path.encode()))

    def tile(self, i, k) -> TileInfo:
        t = bbtc_tile_info()
        _check(lib().bbtc_layout_tile(self._h, i, k, C.byref(t)))
        return TileInfo(t.nnz, t.height, t.width, t.nz_rows, t.avg_degree, t.max_row)

    @property
    def device_bytes(self) -> int:
        return lib().bbtc_layout_device_bytes(self._h)

    @property
    def pinned(self) -> bool:
        """Whether the arrays are in page-locked host memory (Device.pin)."""
        return bool(lib().bbtc_layout_pinned(self._h))

    def arrays(self):
        """Copies of (rowptr, cols, nzrows) (bbtc_layout_arrays); None when
        the layout holds no host arrays."""
        ps = [C.c_void_p() for _ in range(3)]
        ns = [C.c_uint64() for _ in range(3)]
        _check(lib().bbtc_layout_arrays(self._h, C.byref(ps[0]), C.byref(ns[0]), C.byref(ps[1]),
                                        C.byref(ns[1]), C.byref(ps[2]), C.byref(ns[2])))
        if not ps[0].value:
            return None
        return tuple(np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint32)), shape=(n.value,)).copy()
                     for p, n in zip(ps, ns))

    def tasks(self, order: str = "lex") -> np.ndarray:
        n = C.c_uint64()
        cap = lib().bbtc_task_count(self.p)
        out = np.empty(cap, dtype=TASK_DTYPE)
        _check(lib().bbtc_tasks_build(self._h, ORDER[order], _ptr(out), cap, C.byref(n)))
        return out[: n.value]

    def task_bytes(self, tasks, offset_bytes: int):
        t = tasks_array(tasks)
        per = np.zeros(len(t), dtype=np.uint64)
        tot, wed = C.c_uint64(), C.c_uint64()
        _check(lib().bbtc_tasks_bytes(self._h, _ptr(t), len(t), offset_bytes, _ptr(per),
                                      C.byref(tot), C.byref(wed)))
        return per, tot.value, wed.value

    def task_stats(self, tasks, sample_units: int = 0):
        """Per task (n x 6): rows, wedges, sum d_ac(u), probes of bc lists,
        probes of ac lists, bitmap-row lines met by ac probes
        (bbtc_tasks_stats; sample_units > 0: estimates from at most that
        many 64-row units per task)."""
        t = tasks_array(tasks)
        out = np.zeros((len(t), 6), dtype=np.uint64)
        _check(lib().bbtc_tasks_stats(self._h, _ptr(t), len(t), sample_units, _ptr(out)))
        return out

    def count_cpu(self, tasks, nthreads: int = 0):
        """Per-task and total counts on the host (the --ngpu 0 path). Workers
        claim from the back of `tasks`, so pass tasks("asc") to start with the
        heaviest; nthreads 0 means all OpenMP threads."""
        t = tasks_array(tasks)
        per = np.zeros(len(t), dtype=np.uint64)
        tot = C.c_uint64()
        _check(lib().bbtc_count_cpu(self._h, _ptr(t), len(t), nthreads, _ptr(per), C.byref(tot)))
        return per, tot.value

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.bbtc_layout_free(self._h)
            self._h = None


def convert_text(in_path: str, out_path: str, dedup: bool = True):
    """Writes the .bbtc the reference's driver writes for a text graph
    (byte for byte with dedup=False)."""
    _check(lib().bbtc_convert_text(in_path.encode(), out_path.encode(), int(dedup)))


def device_count() -> int:
    n = C.c_int()
    _check(lib().bbtc_device_count(C.byref(n)))
    return n.value


class Device:
    """A bbtc context on one GPU (stream, counters, memory accounting)."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0):
        o = bbtc_opts(hbm_budget_bytes)
        h = C.c_void_p()
        _check(lib().bbtc_ctx_create(device, C.byref(o), C.byref(h)))
        self._h = h

    @property
    def bytes_in_use(self) -> int:
        return lib().bbtc_ctx_bytes_in_use(self._h)

    def synchronize(self):
        _check(lib().bbtc_ctx_synchronize(self._h))

    def read_bandwidth(self, nbytes: int = 8 << 30, iters: int = 5) -> np.ndarray:
        """GB/s of a streaming read over nbytes of device memory, per pass."""
        out = np.zeros(iters, dtype=np.float64)
        _check(lib().bbtc_ctx_read_bandwidth(self._h, nbytes, iters, _ptr(out)))
        return out

    def upload(self, layout: Layout) -> "DeviceGraph":
        return DeviceGraph(self, layout)

    def pin(self, layout: Layout) -> Layout:
        """Moves the layout's block-CSR arrays into page-locked host memory
        (bbtc_layout_pin, as the reference's BlockCSR holds its tiles), so
        uploads, batches and served requests copy them straight over PCIe.
        Returns the layout."""
        _check(lib().bbtc_layout_pin(self._h, layout._h))
        return layout

    def build(self, graph: Graph, cuts, host_arrays: bool = True) -> "DeviceGraph":
        """The block layout of graph built on this device (bbtc_graph_build):
        a resident DeviceGraph whose .layout is the host side, with the
        arrays copied back when host_arrays is set."""
        c = np.ascontiguousarray(cuts, dtype=np.uint32)
        lh, gh = C.c_void_p(), C.c_void_p()
        _check(lib().bbtc_graph_build(self._h, graph._h, _ptr(c), len(c) - 1,
                                      BUILD_HOST_ARRAYS if host_arrays else 0, C.byref(lh), C.byref(gh)))
        L = Layout.__new__(Layout)
        L.graph = graph
        L._h = lh
        L.p = len(c) - 1
        L.cuts = c.copy()
        g = DeviceGraph.__new__(DeviceGraph)
        g.dev = self
        g.layout = L
        g._h = gh
        return g

    def close(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.bbtc_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


class DeviceGraph:
    def __init__(self, dev: Device, layout: Layout):
        self.dev = dev
        self.layout = layout
        h = C.c_void_p()
        _check(lib().bbtc_graph_upload(dev._h, layout._h, C.byref(h)))
        self._h = h

    @property
    def bytes(self) -> int:
        return lib().bbtc_graph_bytes(self._h)

    def autotune(self, tasks):
        """Per-task form choice from measured clocks: (choice[n], cycles[n,5]),
        forms in the order merge, bitmap, dense, rdense, lane. choice & 7 is
        the form (1-5), choice >> 4 the times the rows per work item were
        halved (0-2)."""
        t = tasks_array(tasks)
        choice = np.zeros(len(t), dtype=np.uint8)
        cyc = np.zeros((len(t), 5), dtype=np.uint64)
        _check(lib().bbtc_autotune(self.dev._h, self._h, _ptr(t), len(t), _ptr(choice), _ptr(cyc)))
        return choice, cyc

    def plan(self, tasks, kernel: str = "auto", choice=None) -> "Plan":
        return Plan(self, tasks, kernel, choice)

    def task_stats(self, tasks, sample_units: int = 0):
        """Layout.task_stats computed on the device (bbtc_graph_tasks_stats)."""
        t = tasks_array(tasks)
        out = np.zeros((len(t), 6), dtype=np.uint64)
        _check(lib().bbtc_graph_tasks_stats(self._h, _ptr(t), len(t), sample_units, _ptr(out)))
        return out

    def shard(self, tasks, shards: int):
        """(share of each task, its estimated cost) for `shards` GPUs
        (bbtc_tasks_shard: longest estimated first onto the least loaded
        share)."""
        t = tasks_array(tasks)
        out = np.zeros(len(t), dtype=np.uint32)
        cost = np.zeros(len(t), dtype=np.float64)
        _check(lib().bbtc_tasks_shard(self._h, _ptr(t), len(t), shards, _ptr(out), _ptr(cost)))
        return out, cost

    def count(self, tasks, kernel: str = "auto"):
        t = tasks_array(tasks)
        per = np.zeros(len(t), dtype=np.uint64)
        tot = C.c_uint64()
        _check(lib().bbtc_count(self.dev._h, self._h, _ptr(t), len(t), KERNEL[kernel], _ptr(per),
                                C.byref(tot), None))
        return per, tot.value

    def count_noalloc(self, tasks, kernel: str = "auto", per_task: bool = True):
        """bbtc_count_noalloc: the same count with no device allocation (a
        fixed per-device scratch array; at most 65,536 tasks; calls on one
        device must not overlap). per_task False asks for the total only."""
        t = tasks_array(tasks)
        per = np.zeros(len(t), dtype=np.uint64)
        tot = C.c_uint64()
        _check(lib().bbtc_count_noalloc(self.dev._h, self._h, _ptr(t), len(t), KERNEL[kernel],
                                        _ptr(per) if per_task else None, C.byref(tot), None))
        return (per if per_task else None), tot.value

    def count_hybrid(self, tasks, separator, nthreads=0, nstreams=1, kernel="auto", choice=None):
        """Device from the front, host workers from the back (--border)."""
        t = tasks_array(tasks)
        per = np.zeros(len(t), dtype=np.uint64)
        tot, gt = C.c_uint64(), C.c_uint64()
        ch = None if choice is None else np.ascontiguousarray(choice, dtype=np.uint8)
        _check(lib().bbtc_count_hybrid(self.dev._h, self._h, self.layout._h, _ptr(t), len(t), separator,
                                       nthreads, nstreams, KERNEL[kernel], None if ch is None else _ptr(ch),
                                       _ptr(per), C.byref(tot), C.byref(gt)))
        return per, tot.value, gt.value

    def release(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.bbtc_graph_release(self._h)
            self._h = None

    def __del__(self):
        self.release()


class Plan:
    def __init__(self, g: DeviceGraph, tasks, kernel: str = "auto", choice=None):
        self.g = g
        self.tasks = tasks_array(tasks)
        h = C.c_void_p()
        if choice is None:
            _check(lib().bbtc_plan_create(g.dev._h, g._h, _ptr(self.tasks), len(self.tasks),
                                          KERNEL[kernel], C.byref(h)))
        else:
            self.choice = np.ascontiguousarray(choice, dtype=np.uint8)
            _check(lib().bbtc_plan_create_tuned(g.dev._h, g._h, _ptr(self.tasks), len(self.tasks),
                                                KERNEL[kernel], _ptr(self.choice), C.byref(h)))
        self._h = h

    @property
    def bitmap_tasks(self) -> int:
        return lib().bbtc_plan_bitmap_tasks(self._h)

    def launch(self):
        _check(lib().bbtc_plan_launch(self._h, None))

    def results(self):
        per = np.zeros(len(self.tasks), dtype=np.uint64)
        tot = C.c_uint64()
        _check(lib().bbtc_plan_results(self._h, _ptr(per), C.byref(tot)))
        return per, tot.value

    def time(self, iters: int) -> np.ndarray:
        ms = np.zeros(iters, dtype=np.float32)
        _check(lib().bbtc_plan_time(self._h, iters, _ptr(ms)))
        return ms

    def time_tasks(self) -> np.ndarray:
        ms = np.zeros(len(self.tasks), dtype=np.float32)
        _check(lib().bbtc_plan_time_tasks(self._h, _ptr(ms)))
        return ms

    def profile(self) -> np.ndarray:
        c = np.zeros(len(self.tasks), dtype=np.uint64)
        _check(lib().bbtc_plan_profile(self._h, _ptr(c)))
        return c

    def kinds(self) -> np.ndarray:
        """Form per task: 1 merge, 2 bitmap, 3 dense, 4 rdense, 5 lane, 6 rows
        (a task with work counted by the row form's launch)."""
        k = np.zeros(len(self.tasks), dtype=np.uint8)
        _check(lib().bbtc_plan_task_kinds(self._h, _ptr(k)))
        return k

    def destroy(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.bbtc_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        self.destroy()


# ---- batches -------------------------------------------------------------------
def layouts_build(graphs, ncuts, kind: str = "pal"):
    """Block layouts of many graphs, built in parallel (one graph per thread)."""
    n = len(graphs)
    gh = (C.c_void_p * max(n, 1))(*[g._h for g in graphs])
    nc = np.ascontiguousarray(ncuts, dtype=np.uint32)
    out = (C.c_void_p * max(n, 1))()
    _check(lib().bbtc_layouts_build(gh, n, _ptr(nc), PART[kind], out))
    layouts = []
    for x in range(n):
        L = Layout.__new__(Layout)
        L.graph = graphs[x]
        L._h = C.c_void_p(out[x])
        L.p = int(nc[x])
        ptr = C.cast(lib().bbtc_layout_cuts(L._h), C.POINTER(C.c_uint32))
        L.cuts = np.array([ptr[c] for c in range(L.p + 1)], dtype=np.uint32)
        layouts.append(L)
    return layouts


def tune_ncut(dev: "Device", sample, cands=(1, 2, 4, 8), kind: str = "pal", tune_forms: bool = True,
              by_time: bool = False):
    """Cut-count rule from measured device work on sample graphs (forms
    autotuned per graph, or the AUTO rule's forms; work as summed clocks, or
    with by_time the launch time of each graph alone): returns (rule[64]
    indexed by floor(log2(m+1)), work[n, ncand])."""
    n = len(sample)
    gh = (C.c_void_p * max(n, 1))(*[g._h for g in sample])
    cc = np.ascontiguousarray(cands, dtype=np.uint32)
    rule = np.zeros(64, dtype=np.uint32)
    work = np.zeros((max(n, 1), len(cc)), dtype=np.uint64)
    flags = (1 if tune_forms else 0) | (2 if by_time else 0)
    _check(lib().bbtc_tune_ncut(dev._h, gh, n, _ptr(cc), len(cc), PART[kind], flags, _ptr(rule), _ptr(work)))
    return rule, work[:n]


def rule_ncut(rule, m: int) -> int:
    r = np.ascontiguousarray(rule, dtype=np.uint32)
    return int(lib().bbtc_rule_ncut(_ptr(r), m))


class Batch:
    """Many graphs resident in one arena, counted by one persistent launch."""

    ORDER = {"heavy_first": 0, "small_first": 1}

    def __init__(self, dev: "Device", layouts, kernel: str = "auto", order: str = "heavy_first"):
        self.dev = dev
        self.layouts = list(layouts)
        n = len(self.layouts)
        lh = (C.c_void_p * max(n, 1))(*[L._h for L in self.layouts])
        h = C.c_void_p()
        _check(lib().bbtc_batch_create(dev._h, lh, n, KERNEL[kernel], self.ORDER[order], C.byref(h)))
        self._h = h
        self.ngraphs = n

    @property
    def ntasks(self) -> int:
        return lib().bbtc_batch_ntasks(self._h)

    @property
    def bytes(self) -> int:
        return lib().bbtc_batch_bytes(self._h)

    def tasks(self):
        n = self.ntasks
        g = np.zeros(max(n, 1), dtype=np.uint32)
        t = np.zeros(max(n, 1), dtype=TASK_DTYPE)
        _check(lib().bbtc_batch_tasks(self._h, _ptr(g), _ptr(t)))
        return g[:n], t[:n]

    def autotune(self):
        c = np.zeros(max(self.ntasks, 1), dtype=np.uint8)
        _check(lib().bbtc_batch_autotune(self._h, _ptr(c)))
        return c[: self.ntasks]

    def kinds(self) -> np.ndarray:
        """Form per task (batch order): 1 merge, 2 bitmap, 3 dense, 4 rdense, 5 lane."""
        k = np.zeros(max(self.ntasks, 1), dtype=np.uint8)
        _check(lib().bbtc_batch_task_kinds(self._h, _ptr(k)))
        return k[: self.ntasks]

    def set_forms(self, kernel: str = "auto", choice=None):
        ch = None if choice is None else np.ascontiguousarray(choice, dtype=np.uint8)
        _check(lib().bbtc_batch_set_forms(self._h, KERNEL[kernel], None if ch is None else _ptr(ch)))

    def capture(self):
        _check(lib().bbtc_batch_capture(self._h))

    def launch(self):
        _check(lib().bbtc_batch_launch(self._h))

    def results(self):
        """(per-graph counts, per-task counts in batch order, per-graph latency ms)."""
        pg = np.zeros(max(self.ngraphs, 1), dtype=np.uint64)
        pt = np.zeros(max(self.ntasks, 1), dtype=np.uint64)
        lat = np.zeros(max(self.ngraphs, 1), dtype=np.float64)
        _check(lib().bbtc_batch_results(self._h, _ptr(pg), _ptr(pt), _ptr(lat)))
        return pg[: self.ngraphs], pt[: self.ntasks], lat[: self.ngraphs]

    def destroy(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.bbtc_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.destroy()


# ---- serving -------------------------------------------------------------------
@dataclass
class ServerResult:
    id: int
    total: int
    status: int
    launches: int
    bytes: int
    arrival_s: float
    admit_s: float
    ready_s: float
    start_s: float
    done_s: float
    per_task: np.ndarray = None

    @property
    def latency_s(self) -> float:
        return self.done_s - self.arrival_s


class Server:
    """Continuous-batching request loop on one device (bbtc_server_*).

    Options (0 / False: the library's defaults): heap_bytes, the device heap
    (200 GB); quantum_items, a fixed item budget per launch instead of the
    default budget in estimated work sized to launch_us (40 ms); oldest_pct,
    the oldest request's share of each launch (90); list_chunks, deal a
    launch's items to the claim parts in 4,096-item chunks instead of one
    contiguous range per part; defer_steps, the wave steps after which an item
    is split; lane_form, let AUTO requests whose tasks all admit the lane form
    run in launches of their own; sequential, one request at a time (the plain
    path's order)."""

    def __init__(self, dev: "Device", heap_bytes: int = 0, quantum_items: int = 0, max_inflight: int = 0,
                 max_segments: int = 0, sequential: bool = False, lane_form: bool = False,
                 launch_us: int = 0, oldest_pct: int = 0, list_chunks: bool = False, defer_steps: int = 0):
        o = bbtc_server_opts(heap_bytes, quantum_items, max_inflight, max_segments, int(sequential),
                             int(lane_form), launch_us, oldest_pct, int(list_chunks), defer_steps)
        h = C.c_void_p()
        _check(lib().bbtc_server_create(dev._h, C.byref(o), C.byref(h)))
        self._h = h
        self.dev = dev
        self._keep = {}  # id -> (layout, tasks, per-task buffer) until polled

    def reset_clock(self):
        _check(lib().bbtc_server_reset_clock(self._h))

    def now(self) -> float:
        return lib().bbtc_server_now(self._h)

    def submit(self, layout: "Layout", tasks, arrival_s: float = 0.0, kernel: str = "auto") -> int:
        t = tasks_array(tasks)
        per = np.zeros(max(len(t), 1), dtype=np.uint64)
        rid = C.c_uint64()
        _check(lib().bbtc_server_submit(self._h, layout._h, _ptr(t), len(t), KERNEL[kernel], float(arrival_s),
                                        _ptr(per), C.byref(rid)))
        self._keep[rid.value] = (layout, t, per)
        return rid.value

    def poll(self, max_results: int = 1024, wait_ms: int = 0):
        buf = (bbtc_server_result * max(max_results, 1))()
        got = C.c_uint64()
        _check(lib().bbtc_server_poll(self._h, buf, max_results, wait_ms, C.byref(got)))
        out = []
        for x in range(got.value):
            r = buf[x]
            layout, t, per = self._keep.pop(r.id)
            out.append(ServerResult(r.id, r.total, r.status, r.launches, r.bytes, r.arrival_s, r.admit_s,
                                    r.ready_s, r.start_s, r.done_s, per[: len(t)]))
        return out

    def wait_idle(self):
        _check(lib().bbtc_server_wait_idle(self._h))

    def drain(self):
        """Waits for every outstanding request; returns their results."""
        self.wait_idle()
        return self.poll(max(len(self._keep), 1))

    def stats(self) -> dict:
        s = bbtc_server_stats()
        _check(lib().bbtc_server_get_stats(self._h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in s._fields_}

    def destroy(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.bbtc_server_destroy(self._h)
            self._h = None
            self._keep.clear()

    def __del__(self):
        self.destroy()
