"""Build recipe for libmaze_poisson.so (hipcc, gfx950) and the CPU oracle.

``python -m maze_poisson_amd.build`` compiles in-tree; the shared library is placed
next to this file, where the ctypes binding (and the reference's
``maze_poisson/c_api/__init__.py:44``-style loader) looks for it.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libmaze_poisson.so")
SOURCES = ["grid_kernels.hip", "mg_kernels.hip", "particle_kernels.hip", "pb_kernels.hip", "sim.hip",
           "abi.hip", "engine.hip", "noalloc.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # Keep a*b+c as two rounded operations, as in the reference's C, so that
    # point-wise results match the CPU oracle bit for bit.
    "-ffp-contract=off",
    # Packed FP32 instructions are disabled on this target.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build_library(verbose: bool = True) -> str:
    objs = []
    os.makedirs(os.path.join(PKG, "_build"), exist_ok=True)
    procs = []
    for src in SOURCES:
        obj = os.path.join(PKG, "_build", src.replace(".hip", ".o"))
        cmd = [hipcc()] + [f for f in HIPCC_FLAGS if f != "-shared"] + [
            "-c", os.path.join(CSRC, src), "-o", obj]
        procs.append((src, cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                                 stderr=subprocess.STDOUT)))
        objs.append(obj)
    for src, cmd, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out.decode(errors="replace"))
            raise RuntimeError(f"hipcc failed on {src}: {' '.join(cmd)}")
        if verbose and out.strip():
            msgs = [ln for ln in out.decode(errors="replace").splitlines()
                    if "packed-fp32-ops" not in ln and ln.strip()]
            if msgs:
                sys.stderr.write("\n".join(msgs) + "\n")
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


def build_oracle(verbose: bool = True) -> None:
    """The CPU oracle, and the reference C when its sources are present."""
    oracle = os.path.join(ROOT, "oracle")
    subprocess.check_call(["make", "-s", "-C", oracle, "-B", "all"])
    if os.path.isdir("/root/reference/library/src"):
        subprocess.check_call(["make", "-s", "-C", oracle, "ref"])
    if verbose:
        print("built oracle")


if __name__ == "__main__":
    build_library()
    build_oracle()
