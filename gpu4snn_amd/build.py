"""Build the gfx950 library and CLI in-tree with hipcc.

Outputs (git-ignored, shipped with the tree):
  gpu4snn_amd/libgpu4snn.so   -- kernels + host runtime + C ABI
  gpu4snn_amd/bin/gpu4snn     -- the reference-compatible CLI
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libgpu4snn.so")
CLI = os.path.join(PKG, "bin", "gpu4snn")

LIB_SOURCES = ["g4s_kernels.hip", "g4s_build.hip", "g4s_runtime.hip", "g4s_batch.hip", "g4s_serve.hip"]
CLI_SOURCES = ["g4s_cli.cpp"]

# -ffp-contract=off: neuron arithmetic must equal the host's binary32 ops.
# -packed-fp32-ops off: no v_pk_{add,mul,fma}_f32 in generated code.
COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def hipcc() -> str:
    p = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(p):
        raise RuntimeError("hipcc not found; the library is built with ROCm's hipcc for gfx950")
    return p


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build failed: {' '.join(cmd[:3])} ... (exit {r.returncode})")
    return r


def build(verbose: bool = False, arch: str = "gfx950") -> None:
    srcs = [os.path.join(CSRC, s) for s in LIB_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = LIB + ".tmp"
    _run([hipcc(), f"--offload-arch={arch}", *COMMON, "-shared", "-o", tmp, *srcs, "-lpthread"], verbose)
    os.replace(tmp, LIB)
    cli_srcs = [os.path.join(CSRC, s) for s in CLI_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    if cli_srcs:
        os.makedirs(os.path.dirname(CLI), exist_ok=True)
        _run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-o", CLI + ".tmp", *cli_srcs,
              "-L", PKG, "-lgpu4snn", f"-Wl,-rpath,$ORIGIN/..", "-lpthread"], verbose)
        os.replace(CLI + ".tmp", CLI)


if __name__ == "__main__":
    build(verbose=True)
