"""Build the gfx950 library and CLI in-tree with hipcc.

Outputs (git-ignored, shipped with the tree):
  gpu4snn_amd/libgpu4snn.so   -- kernels + host runtime + C ABI
  gpu4snn_amd/bin/gpu4snn     -- the reference-compatible CLI

Each source compiles to its own object in gpu4snn_amd/_build.  The kernel
file is compiled with the compiler's resource-usage report on; the report
(VGPRs, SGPRs, static LDS and the occupancy in waves per SIMD of every kernel
instantiation) becomes a generated table, g4s_resources.cpp, linked into the
library.  The runtime sizes every grid whose workgroups wait on each other
from that table, so a persistent launch never asks for more workgroups than
the device holds at once.
"""
from __future__ import annotations

import os
import re
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libgpu4snn.so")
CLI = os.path.join(PKG, "bin", "gpu4snn")

LIB_SOURCES = ["g4s_kernels.hip", "g4s_build.hip", "g4s_runtime.hip", "g4s_batch.hip", "g4s_serve.hip"]
CLI_SOURCES = ["g4s_cli.cpp"]
# kernels whose resource usage the runtime needs (persistent launches)
REPORTED = "g4s_kernels.hip"

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


_FIELDS = {"TotalSGPRs": "sgprs", "VGPRs": "vgprs", "AGPRs": "agprs", "ScratchSize [bytes/lane]": "scratch",
           "Occupancy [waves/SIMD]": "occupancy", "LDS Size [bytes/block]": "lds"}


def parse_resource_remarks(text: str) -> list[dict]:
    """Kernel records from hipcc's -Rpass-analysis=kernel-resource-usage
    report: {"mangled", "sgprs", "vgprs", "agprs", "scratch", "occupancy",
    "lds"} per function, in report order."""
    out, cur = [], None
    for line in text.splitlines():
        m = re.search(r"remark:\s+Function Name: (\S+)", line)
        if m:
            cur = {"mangled": m.group(1)}
            out.append(cur)
            continue
        m = re.search(r"remark:\s+(.+?): (\d+) \[-Rpass-analysis", line)
        if m and cur is not None and m.group(1) in _FIELDS:
            cur[_FIELDS[m.group(1)]] = int(m.group(2))
    return [r for r in out if "occupancy" in r]


def demangle(names: list[str]) -> list[str]:
    tool = shutil.which("c++filt") or shutil.which("llvm-cxxfilt")
    if not tool:
        raise RuntimeError("c++filt not found: needed to name the kernels of the resource report")
    r = subprocess.run([tool], input="\n".join(names) + "\n", capture_output=True, text=True, check=True)
    return r.stdout.splitlines()


def kernel_key(demangled: str) -> tuple[str, str] | None:
    """('k_skl_slice', '1,false,false') from 'void g4s::(anonymous
    namespace)::k_skl_slice<1, false, false>(g4s::DevNet, int, int, int)';
    None for anything that is not a kernel of this library."""
    m = re.match(r"(?:void )?g4s::\(anonymous namespace\)::(k_\w+)(?:<([^>]*)>)?\(", demangled)
    if not m:
        return None
    return m.group(1), (m.group(2) or "").replace(" ", "")


def resource_table_source(records: list[dict]) -> str:
    names = demangle([r["mangled"] for r in records])
    rows = []
    for r, dn in zip(records, names):
        key = kernel_key(dn)
        if key is None:
            continue
        rows.append(f'    {{"{key[0]}", "{key[1]}", {r["occupancy"]}, {r.get("vgprs", 0)}, {r.get("agprs", 0)}, '
                    f'{r.get("sgprs", 0)}, {r.get("lds", 0)}, {r.get("scratch", 0)}}},')
    if not rows:
        raise RuntimeError("the compiler's resource report named no kernel")
    return ("// Generated by gpu4snn_amd/build.py from hipcc's kernel resource-usage report\n"
            "// of g4s_kernels.hip (gfx950).  Do not edit.\n"
            '#include "g4s_internal.h"\n\nnamespace g4s {\n'
            "const KernelRes kKernelRes[] = {\n" + "\n".join(rows) + "\n};\n"
            f"const int kKernelResCount = {len(rows)};\n}}  // namespace g4s\n")


def build(verbose: bool = False, arch: str = "gfx950") -> None:
    os.makedirs(BUILD, exist_ok=True)
    cc = hipcc()
    srcs = [s for s in LIB_SOURCES if os.path.exists(os.path.join(CSRC, s))]

    def compile_one(s):
        flags = ["-Rpass-analysis=kernel-resource-usage"] if s == REPORTED else []
        obj = os.path.join(BUILD, s + ".o")
        r = _run([cc, f"--offload-arch={arch}", *COMMON, *flags, "-c", os.path.join(CSRC, s), "-o", obj], verbose)
        return obj, (r.stderr if s == REPORTED else "")

    with ThreadPoolExecutor(max_workers=min(len(srcs), max(1, os.cpu_count() or 1))) as ex:
        results = list(ex.map(compile_one, srcs))
    objs = [o for o, _ in results]
    report = "".join(t for _, t in results)
    gen = os.path.join(BUILD, "g4s_resources.cpp")
    with open(gen, "w") as f:
        f.write(resource_table_source(parse_resource_remarks(report)))
    gen_obj = gen + ".o"
    _run([cc, f"--offload-arch={arch}", *COMMON, "-I", CSRC, "-x", "hip", "-c", gen, "-o", gen_obj], verbose)
    tmp = LIB + ".tmp"
    rocm_lib = os.path.join(os.path.dirname(os.path.dirname(os.path.realpath(cc))), "lib")
    _run([cc, f"--offload-arch={arch}", "-shared", "-fPIC", "-o", tmp, *objs, gen_obj, "-lpthread",
          f"-L{rocm_lib}", "-lrocprofiler-sdk-roctx", f"-Wl,-rpath,{rocm_lib}"], verbose)
    os.replace(tmp, LIB)
    cli_srcs = [os.path.join(CSRC, s) for s in CLI_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    if cli_srcs:
        os.makedirs(os.path.dirname(CLI), exist_ok=True)
        _run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-o", CLI + ".tmp", *cli_srcs,
              "-L", PKG, "-lgpu4snn", f"-Wl,-rpath,$ORIGIN/..", "-lpthread", "-ldl"], verbose)
        os.replace(CLI + ".tmp", CLI)


if __name__ == "__main__":
    build(verbose=True)
