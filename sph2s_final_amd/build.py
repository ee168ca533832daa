"""Build step of libsph2s.so (HIP, gfx950) -- run by ``__graft_entry__.build()``.

The library is compiled in-tree with hipcc so that the shared object sits next
to the Python package that loads it.  Flags:

* ``--offload-arch=gfx950``: MI355X only.
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32 VALU
  instructions are generated.
* ``-fhip-fp32-correctly-rounded-divide-sqrt``: IEEE FP32 division, so cell
  coordinates ``(int)((x - o) / spacing)`` match the FP32 reference arithmetic
  bit for bit.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB_NAME = "libsph2s.so"
SOURCES = ["api.hip", "radix_sort.hip", "step.hip", "two_scale.hip", "batch.hip", "adjoint.hip",
           "walls.hip", "diagnostics.hip"]
HEADERS = ["internal.hpp", "neighbors.hpp", "pairwalk.hpp", "batch.hpp", "walls.hpp",
           "diagnostics.hpp"]


def hipcc_path() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build libsph2s.so")


def compile_flags() -> list[str]:
    return [
        "--offload-arch=gfx950",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-fhip-fp32-correctly-rounded-divide-sqrt",
        "-Wall",
        "-Wno-unused-command-line-argument",
    ]


def build_library(verbose: bool = False) -> str:
    """Compile every HIP source into ``sph2s_final_amd/libsph2s.so``; return its path."""
    out = os.path.join(HERE, LIB_NAME)
    tmp = out + ".tmp"
    cmd = [hipcc_path(), *compile_flags(), "-shared", "-o", tmp,
           *[os.path.join(CSRC, s) for s in SOURCES]]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed with exit code {res.returncode}")
    if verbose and res.stderr.strip():
        sys.stderr.write(res.stderr)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build_library(verbose=True))
