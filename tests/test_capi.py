"""The C-ABI library loads and exports exactly what include/biofvm_amd.h declares.

CPU-only: no entry point that touches the device is called here, except
argument validation that returns before any HIP call.
"""
import ctypes
import os
import re
import subprocess

import pytest

from biofvm_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "biofvm_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bfvm_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_bound_signatures():
    assert declared_functions() == sorted(_capi.SIGNATURES)


def test_library_exports_every_declared_symbol():
    lib = _capi.load()
    for name in declared_functions():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", _capi.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (bfvm_[a-z_0-9]+)", out))
    assert exported == set(declared_functions())


def test_one_hip_runtime_in_the_process():
    # torch is loaded before the library, so the library binds to the HIP
    # runtime torch brought in (same soname) instead of loading a second one
    import sys
    _capi.load()
    assert "torch" in sys.modules
    maps = open("/proc/self/maps").read()
    runtimes = {line.split()[-1] for line in maps.splitlines() if "libamdhip64" in line}
    assert len(runtimes) == 1, runtimes


def test_library_carries_gfx950_code_objects():
    # the offload bundle embedded by hipcc names its target triple
    blob = open(_capi.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob


@pytest.mark.parametrize("args", [(0, 4, 4, 1), (4, 4, 4, 0), (4, -1, 4, 2)])
def test_create_rejects_bad_sizes_before_touching_the_device(args):
    lib = _capi.load()
    h = ctypes.c_void_p()
    rc = lib.bfvm_create(*args, 10.0, _capi.FP64, ctypes.byref(h))
    assert rc == -1 and not h.value
    assert b">= 1" in lib.bfvm_last_error(None)


def test_create_rejects_bad_precision_and_edges():
    lib = _capi.load()
    h = ctypes.c_void_p()
    assert lib.bfvm_create(4, 4, 4, 1, 10.0, 7, ctypes.byref(h)) == -1
    assert lib.bfvm_create(4, 4, 4, 1, -1.0, 0, ctypes.byref(h)) == -1
    assert lib.bfvm_create(4, 4, 4, 1, 10.0, 0, None) == -1


def test_null_handle_is_rejected():
    lib = _capi.load()
    assert lib.bfvm_step(None, 0.01, 7) == -1
    assert lib.bfvm_sync(None) == -1
    assert lib.bfvm_destroy(None) == 0
