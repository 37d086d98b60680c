"""The C-ABI library: it is built, loads, exports every entry point include/zmpm.h declares,
and reports errors as status codes.  No kernel runs here (CPU-only check)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "zmpm.h")
LIB = os.path.join(ROOT, "zeno_amd", "lib", "libzmpm.so")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^ZMPM_API\s+[\w\s\*]*?\b(zmpm_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "zmpm_create" in syms and "zmpm_substeps" in syms and "zmpm_p2g" in syms
    assert len(syms) >= 20


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build libzmpm.so first (python -m zeno_amd.build)"
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT\s+(zmpm_\w+)", out))
    missing = [s for s in declared_symbols() if s not in exported]
    assert not missing, missing


def test_binding_covers_header():
    from zeno_amd import zmpm
    assert sorted(zmpm.EXPORTED) == declared_symbols()


def test_load_and_error_paths_without_gpu():
    import torch
    from zeno_amd import zmpm
    L = zmpm.load_library()
    for name in declared_symbols():
        assert hasattr(L, name)
    # argument errors never reach the device
    h = ctypes.c_void_p()
    assert L.zmpm_create(ctypes.byref(h), None) == 1
    assert b"null" in L.zmpm_last_error()
    cfg = zmpm.CConfig(0.0, 0, 0, 0)
    assert L.zmpm_create(ctypes.byref(h), ctypes.byref(cfg)) == 1
    assert L.zmpm_sync(None) == 1
    if not torch.cuda.is_available():
        cfg = zmpm.CConfig(1 / 64, 64, 1000, 0)
        rc = L.zmpm_create(ctypes.byref(h), ctypes.byref(cfg))
        assert rc == 6  # ZMPM_ERR_NODEVICE
        with pytest.raises(zmpm.ZmpmError):
            zmpm.Context(1 / 64, 64, 1000)
