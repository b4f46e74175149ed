"""The C-ABI library loads and exports every function the public headers
declare (no device calls: this runs without a GPU)."""
import ctypes
import os
import re
import subprocess
import sys

import pytest

import hsrq3in_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("hsrq.h", "inverse_iteration.h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"\b([a-z_][a-z0-9_]*)\s*\(", txt):
            name = m.group(1)
            if name.startswith("hsrq_") or name == "tiled_inverse_iteration":
                names.add(name)
    return names


def test_library_loads_and_exports_all_declared():
    lib = ctypes.CDLL(hsrq3in_amd.LIB_PATH)
    decl = declared_functions()
    assert "tiled_inverse_iteration" in decl and len(decl) >= 15
    missing = [n for n in sorted(decl) if not hasattr(lib, n)]
    assert not missing, missing
    assert set(hsrq3in_amd.EXPORTS) <= decl


def test_exports_are_c_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", hsrq3in_amd.LIB_PATH],
                         stdout=subprocess.PIPE, text=True, check=True).stdout
    syms = {line.split()[-1] for line in out.splitlines() if line.strip()}
    for n in declared_functions():
        assert n in syms, n      # unmangled, extern "C"


def test_strerror_is_host_only():
    lib = hsrq3in_amd.lib()
    assert lib.hsrq_strerror(-1) == b"invalid argument"
    assert lib.hsrq_strerror(-5).startswith(b"device self-test")


def test_invalid_arguments_rejected_without_device():
    lib = hsrq3in_amd.lib()
    assert lib.hsrq_context_set_tile(None, 64) == -1
    assert lib.hsrq_solve_device(None, 10, None, 10, 1, None, None, 10, 64, None) == -1
    assert lib.hsrq_plan_execute(None) == -1
    assert lib.hsrq_gaussian_hessenberg_device(None, 10, 1, None, 10) == -1
    assert lib.hsrq_grade_device(None, 10, None, 10, None, 0, None, None) == -1
    assert lib.hsrq_constructed_hessenberg_device(None, 10, None, None, None, None, None,
                                                  10) == -1
    assert lib.hsrq_strerror(-8) == b"result outside the floating-point range"


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    assert r.returncode == 2 and "one GPU" in r.stderr
