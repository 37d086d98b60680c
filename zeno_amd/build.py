"""Build libzmpm.so (hand-written HIP for gfx950) in-tree with hipcc.

    python -m zeno_amd.build      # or zeno_amd.build.build()

Output: zeno_amd/lib/libzmpm.so.  The library is plain C ABI (include/zmpm.h);
nothing here depends on torch.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "zeno_amd", "csrc")
LIBDIR = os.path.join(ROOT, "zeno_amd", "lib")
LIB = os.path.join(LIBDIR, "libzmpm.so")
SOURCES = ["zmpm.hip", "sort.hip", "mpm.hip", "adjoint.hip"]
ARCH = "gfx950"

# -ffp-contract=off: products and sums round as written (fmaf() where a fused
# multiply-add is wanted), so the binning key and the two G2P passes are
# bit-reproducible.  Packed FP32 instructions are disabled for gfx950.
HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-fvisibility=hidden", "-Wall", "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    objs = []
    for src in SOURCES:
        obj = os.path.join(LIBDIR, src.replace(".hip", ".o"))
        cmd = [hipcc(), *HIPCC_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"hipcc failed on {src}")
        errs = [l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and "warnings generated" not in l
                and "ignoring feature" not in l and l.strip()]
        if verbose and errs:
            sys.stderr.write("\n".join(errs) + "\n")
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
