"""Builds libfastatari_hip.so (HIP kernels + C ABI) for gfx950 with hipcc.

Each translation unit is compiled to an object with its own options (in
parallel), then the objects are linked into the shared library."""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(ROOT, "csrc")
LIB = os.path.join(PKG, "libfastatari_hip.so")
OBJ = os.path.join(ROOT, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "-std=c++17", "-O3", "--offload-arch=gfx950", "-fPIC",
    # packed FP32 instructions are disabled for every HIP translation unit
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # the SLP vectoriser packs wave-uniform and per-lane values into one vector,
    # which makes the uniformity analysis treat the scalar control state of the
    # frame loop's uniform mode as per-lane (VGPRs and exec-masked branches)
    "-fno-slp-vectorize",
    "-Wno-unused-result", "-Wno-unused-value", "-pthread",
]
# Both frame kernels: control flow whose branches are all wave-uniform is
# left as written (the opcode dispatch and the uniform mode's handlers stay
# scalar compare-and-branch code instead of flag-steered flow blocks).  The
# option mishandles uniform regions nested inside divergent ones, so the
# opcode-grouping loop puts its scalar dispatch around the per-lane group test
# rather than inside it.
UNIFORM_SKIP = ["-mllvm", "-structurizecfg-skip-uniform-regions=true"]

# (source, extra options); fa_kernels*.hip hold the frame kernels, the others
# are host code over their launchers, each with its own small kernels
# (fa_render.hip: the rendering console's kernel and C ABI)
SOURCES = [
    ("fa_kernels.hip", UNIFORM_SKIP),
    ("fa_kernels_sorted.hip", UNIFORM_SKIP),
    ("fa_lib.hip", []),
    ("fa_batch.hip", []),
    ("fa_serve.hip", []),
    ("fa_adapter.hip", []),
    ("fa_noalloc.hip", []),
    ("fa_probe.hip", []),
    ("fa_render.hip", []),
]


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    procs, objs = [], []
    for src, extra in SOURCES:
        obj = os.path.join(OBJ, src.replace(".hip", ".o"))
        cmd = [HIPCC, *FLAGS, *extra, "-c", "-o", obj, os.path.join(CSRC, src)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        objs.append(obj)
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out)
            failed.append(src)
    if failed:
        raise RuntimeError("hipcc failed on " + ", ".join(failed))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed linking libfastatari_hip.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
