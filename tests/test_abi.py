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


def test_check_particles_rejects_views_outside_the_tile():
    """zmpm_check_particles (run by import and export on every view) turns an AoSoA32 channel past
    the tile, a negative channel, a missing tile width and unpaired optional attributes into
    ZMPM_ERR_ARG on the host; no pointer is dereferenced (these are never device addresses)."""
    from zeno_amd import zmpm
    L = zmpm.load_library()
    P = 0x1000  # a stand-in address: the check never reads through it

    def view(**kw):
        v = zmpm.CParticles()
        v.n, v.layout, v.tile_channels = kw.pop("n", 10), kw.pop("layout", zmpm.LAYOUT_AOSOA32), kw.pop("tc", 20)
        v.x, v.x_chn = P, 1
        for name, chn in kw.items():
            setattr(v, name, P)
            setattr(v, name + "_chn", chn)
        return v

    ok = [view(J=19), view(tc=28, v=4, C=8, F=19), view(tc=28, m=0, vol=7), view(layout=zmpm.LAYOUT_SOA, tc=0, F=99),
          view(layout=zmpm.LAYOUT_AOS, tc=0, C=-5)]
    for v in ok:
        assert L.zmpm_check_particles(ctypes.byref(v)) == 0, L.zmpm_last_error()
    bad = [view(F=19),                 # 9 channels from 19 in a 20-channel J-fluid tile
           view(tc=28, a=26),          # 3 channels from 26
           view(J=-1), view(tc=0, J=3), view(tc=-4), view(n=-1), view(layout=7),
           view(tc=28, mu=28), view(tc=28, vol=28, m=0)]
    for v in bad:
        assert L.zmpm_check_particles(ctypes.byref(v)) == 1
    v = view(tc=28, m=0)  # m without vol
    assert L.zmpm_check_particles(ctypes.byref(v)) == 1
    v = view(F=19)
    v.x = None  # positions missing
    v.F = None
    assert L.zmpm_check_particles(ctypes.byref(v)) == 1
    assert L.zmpm_check_particles(None) == 1
    with pytest.raises(zmpm.ZmpmError):
        zmpm.check_particles(view(F=19))
