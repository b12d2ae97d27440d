"""The C-ABI boundary: libbbtc.so loads without a GPU and exports every
function include/bbtc.h declares; the declarations stay plain C."""
import ctypes as C
import os
import re
import subprocess
import sys

import pytest

import bbtc_amd as B

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
HEADER = os.path.join(ROOT, "include", "bbtc.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    names = re.findall(r"\b(bbtc_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_functions():
    names = declared_functions()
    assert len(names) > 60
    for must in ["bbtc_hgraph_load", "bbtc_partition", "bbtc_layout_build", "bbtc_count", "bbtc_count_hybrid",
                 "bbtc_batch_create", "bbtc_server_create", "bbtc_convert_text", "bbtc_layout_save_bcsr"]:
        assert must in names


def test_every_declared_function_is_exported():
    lib = C.CDLL(B.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_exported_symbols_are_c_names():
    out = subprocess.run(["nm", "-D", "--defined-only", B.LIB_PATH], capture_output=True, text=True, check=True)
    exported = {ln.split()[-1] for ln in out.stdout.splitlines() if " T " in ln}
    for n in declared_functions():
        assert n in exported, n  # unmangled: extern "C"


def test_header_is_plain_c(tmp_path):
    src = tmp_path / "use.c"
    src.write_text('#include "bbtc.h"\nint main(void) { bbtc_status s = BBTC_OK; (void)s; return 0; }\n')
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), "-c", str(src),
                        "-o", str(tmp_path / "use.o")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_status_strings_and_errors_without_gpu():
    lib = B.lib()
    for s in range(9):
        assert lib.bbtc_strerror(s)
    with pytest.raises(B.BbtcError) as e:
        B.Graph.load("/nonexistent/graph.bbtc")
    assert e.value.status == 2  # BBTC_ERR_IO
    g = B.Graph.rmat(8, 4, seed=1)
    with pytest.raises(B.BbtcError) as e:
        g.partition("pal", 0)
    assert e.value.status == 1  # BBTC_ERR_INVALID


STRUCTS = {
    "bbtc_task": B.bbtc_task, "bbtc_tile_info": B.bbtc_tile_info, "bbtc_opts": B.bbtc_opts,
    "bbtc_server_opts": B.bbtc_server_opts, "bbtc_server_result": B.bbtc_server_result,
    "bbtc_server_stats": B.bbtc_server_stats,
}


def test_python_structs_match_the_header(tmp_path):
    # sizes and field offsets of every struct the ctypes layer mirrors, as the
    # C compiler lays them out from include/bbtc.h
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "bbtc.h"', "int main(void) {"]
    for name, py in STRUCTS.items():
        lines.append(f'  printf("{name} size %zu\\n", sizeof({name}));')
        for field, _ in py._fields_:
            lines.append(f'  printf("{name} {field} %zu\\n", offsetof({name}, {field}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    r = subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {}
    for ln in out:
        if ln:
            name, what, val = ln.split()
            got[(name, what)] = int(val)
    for name, py in STRUCTS.items():
        assert got[(name, "size")] == C.sizeof(py), name
        for field, _ in py._fields_:
            assert got[(name, field)] == getattr(py, field).offset, (name, field)


SWEEP = r'''
import ctypes as C, re, sys
import bbtc_amd as B
hdr = re.sub(r"/\*.*?\*/", "", open(sys.argv[1]).read(), flags=re.S)
lib = C.CDLL(B.LIB_PATH)
size = int(sys.argv[2])
for ret, name, params in re.findall(r"^\s*([A-Za-z_][\w\s\*]*?)\b(bbtc_\w+)\s*\(([^)]*)\)\s*;", hdr, flags=re.M):
    args = []
    for p in (p.strip() for p in params.split(",")):
        if not p or p == "void":
            continue
        args.append(C.c_void_p(None) if "*" in p else C.c_double(0.0) if "double" in p
                    else C.c_float(0.0) if "float" in p else C.c_uint64(size))
    f = getattr(lib, name)
    status = ret.strip().startswith("bbtc_status")
    f.restype = C.c_int if status else (C.c_void_p if "*" in ret else C.c_uint64)
    print(name, end=" ", flush=True)
    print(int(status), f(*args), flush=True)
'''


@pytest.mark.parametrize("size", [0, 1])
def test_null_arguments_are_errors(size):
    """Every entry point, called with NULL pointers and sizes 0 or 1 and no
    GPU, returns an error status (or, for the getters and destroyers, does
    nothing) instead of crashing; the sweep runs in a child process so a crash
    names the function."""
    r = subprocess.run([sys.executable, "-c", SWEEP, HEADER, str(size)], capture_output=True, text=True,
                       timeout=120, cwd=ROOT)
    assert r.returncode == 0, r.stdout[-300:] + r.stderr[-2000:]
    seen = 0
    for ln in r.stdout.splitlines():
        name, status, value = ln.split()
        seen += 1
        if status == "1" and not (name == "bbtc_layouts_build" and size == 0):  # n = 0: nothing to build
            assert int(value) != 0, name
    assert seen == len(declared_functions())
