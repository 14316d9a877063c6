"""Compile libpcm.so for gfx950 with hipcc (in-tree, no JIT cache).

Flags:
  -ffp-contract=off          keep the reference's separate multiply and add roundings
                             (neighbor sets must be bit-exact, velocities within 1e-5/1e-12);
  -Xclang -target-feature -Xclang -packed-fp32-ops
                             no packed FP32 instructions are generated;
  -fgpu-flush-denormals-to-zero is not used: denormals are kept as on the host.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
OUT_DIR = os.path.join(PKG, "_build")
LIB = os.path.join(OUT_DIR, "libpcm.so")
SOURCES = ["pcm_lib.hip"]
HEADERS = ["pcm_device.h", "pcm_kernels.h", "pcm_rng.h", os.path.join("..", "..", "include", "pcm.h")]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def flags(arch: str = "gfx950"):
    return [
        f"--offload-arch={arch}",
        "-O3",
        "-std=c++20",
        "-fPIC",
        "-shared",
        "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wno-unused-command-line-argument",
    ]


def build(verbose: bool = False) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *flags(), "-o", tmp, *[os.path.join(CSRC, s) for s in SOURCES]]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libpcm.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
