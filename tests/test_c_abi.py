"""The C ABI as a plain C caller sees it: tests/c_abi/abi_probe.c is compiled by gcc
against include/zmpm.h alone and linked to libzmpm.so.  Its struct sizes and field
offsets must equal the Python binding's ctypes layouts, and its calls return the
documented status codes (CPU only: no kernel runs)."""
import ctypes
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "c_abi", "abi_probe.c")
LIBDIR = os.path.join(ROOT, "zeno_amd", "lib")


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    if not os.path.exists(os.path.join(LIBDIR, "libzmpm.so")):
        pytest.fail("build libzmpm.so first (python -m zeno_amd.build)")
    exe = str(tmp_path_factory.mktemp("c_abi") / "abi_probe")
    hip_lib = "/opt/rocm/lib"
    cmd = ["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"), SRC, "-o", exe,
           "-L", LIBDIR, "-lzmpm", f"-Wl,-rpath,{LIBDIR}", f"-Wl,-rpath,{hip_lib}"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    env = dict(os.environ)
    env["LD_LIBRARY_PATH"] = ":".join([LIBDIR, hip_lib, env.get("LD_LIBRARY_PATH", "")])
    out = subprocess.run([exe], check=True, capture_output=True, text=True, env=env, timeout=120).stdout
    layouts, calls, errors = {}, {}, []
    for line in out.splitlines():
        parts = line.split()
        if parts[0] == "call":
            calls[parts[1]] = int(parts[2])
        elif parts[0] == "error":
            errors.append(line[6:])
        else:
            layouts.setdefault(parts[0], {})[parts[1]] = int(parts[2])
    return layouts, calls, errors


def ctypes_structs():
    from zeno_amd import zmpm
    return {"zmpm_material": zmpm.CMaterial, "zmpm_config": zmpm.CConfig, "zmpm_scene": zmpm.CScene,
            "zmpm_step_params": zmpm.CStepParams, "zmpm_collider": zmpm.CCollider,
            "zmpm_particles": zmpm.CParticles, "zmpm_boundary_particles": zmpm.CBoundaryParticles,
            "zmpm_adjoint": zmpm.CAdjoint}


def test_struct_layouts_match_ctypes(probe):
    layouts, _, _ = probe
    structs = ctypes_structs()
    assert set(layouts) == set(structs)
    for name, cls in structs.items():
        c = layouts[name]
        assert c.pop("sizeof") == ctypes.sizeof(cls), name
        fields = [f[0] for f in cls._fields_]
        assert sorted(c) == sorted(fields), (name, sorted(set(c) ^ set(fields)))
        for f in fields:
            assert c[f] == getattr(cls, f).offset, (name, f)


def test_plain_c_calls(probe):
    _, calls, errors = probe
    import torch
    if torch.cuda.is_available():
        assert calls["create"] == 0 and calls["destroy"] == 0
    else:
        assert calls["create"] == 6  # ZMPM_ERR_NODEVICE
    assert calls["create_null"] == 1  # ZMPM_ERR_ARG
    assert calls["check_ok"] == 0
    assert calls["check_bad"] == 1
    assert '"F" outside the tile' in errors[-1]
