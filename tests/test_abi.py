"""The C-ABI library: it loads, exports every function declared in
include/*.h, and reports errors through return codes.  CPU only (no compute
call reaches a GPU here)."""
import ctypes as C
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(beep_\w+)\s*\(", txt)))


@pytest.mark.parametrize("header", ["beep_hip.h", "beep_host.h", "beep_serve.h"])
def test_every_declared_symbol_is_exported(header):
    from beep3_amd._lib import LIB_PATH

    lib = C.CDLL(LIB_PATH)
    names = declared(header)
    assert len(names) >= 8
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_headers_compile_as_c():
    src = ('#include "beep_hip.h"\n#include "beep_host.h"\n#include "beep_serve.h"\n'
           'int main(void){return 0;}\n')
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                        "-x", "c", "-", "-o", "/dev/null"], input=src, text=True,
                       capture_output=True)
    assert r.returncode == 0, r.stderr


def test_null_context_is_rejected():
    from beep3_amd._lib import lib

    L = lib()
    assert L.beep_hip_matvec(None, None, None, 0) != 0
    assert L.beep_hip_precompute(None, 0.1, 0) != 0
    assert L.beep_hip_destroy(None) == 0


def test_create_without_gpu_reports_an_error():
    """On a machine without a GPU, beep_hip_create fails with a message
    instead of crashing; on a GPU machine it succeeds."""
    from beep3_amd._lib import lib, vp

    L = lib()
    h = vp()
    rc = L.beep_hip_create(0, C.byref(h))
    if rc == 0:
        L.beep_hip_destroy(h)
    else:
        assert L.beep_hip_last_error(None).decode()


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
