"""The C-ABI libraries load and export every symbol their headers declare."""

import ctypes
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "grico.jl_amd", "lib")


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:gmf|gh)_[a-z0-9_]+)\s*\(", text)))


def exported(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


def test_host_library_exports_header():
    path = os.path.join(LIB, "libgrico_host.so")
    ctypes.CDLL(path)
    missing = [s for s in declared("grico_host.h") if s not in exported(path)]
    assert not missing, missing


def test_device_library_exports_header():
    path = os.path.join(LIB, "libgrico_mf.so")
    if not os.path.exists(path):
        subprocess.run([sys.executable, os.path.join(ROOT, "grico.jl_amd", "build.py"), "device"], check=True)
    syms = declared("grico_mf.h")
    assert "gmf_apply" in syms and "gmf_cg_solve" in syms
    missing = [s for s in syms if s not in exported(path)]
    assert not missing, missing


def test_device_library_loads_without_gpu():
    path = os.path.join(LIB, "libgrico_mf.so")
    if not os.path.exists(path):
        pytest.skip("device library not built")
    lib = ctypes.CDLL(path)
    lib.gmf_last_error.restype = ctypes.c_char_p
    assert lib.gmf_last_error() is not None


def test_null_arguments_rejected_without_touching_the_device():
    path = os.path.join(LIB, "libgrico_mf.so")
    if not os.path.exists(path):
        pytest.skip("device library not built")
    lib = ctypes.CDLL(path)
    lib.gmf_last_error.restype = ctypes.c_char_p
    assert lib.gmf_apply(None, None, None) == -1
    assert b"null" in lib.gmf_last_error()
