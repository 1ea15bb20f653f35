"""Drop-in driver: `python gpu4snn.py GPU_NAME ITERATIONS [extra CLI flags]`.

Same contract as the reference driver (Crypto-TII/GPU4SNN gpu4snn.py:19-47):
N and Nsyn are module constants; it removes old results and builds
(reference: make clean / rmdata / all, py:27-29), creates ./GPU_NAME
(py:37), runs `gpu4snn N Nsyn ITERATIONS` -- all 3 regimes x 5 schemes by
default (py:42) -- and moves ./Results to ./GPU_NAME/ITERATIONS/Results
(py:45-47).  Differences: importing this module runs nothing; the build runs
only when the library or CLI is missing (or with --rebuild); extra flags
(--alg, --regime, --mode det|atomic, --seed) pass through to the CLI; the
exit code is the CLI's.
"""
import os
import shutil
import subprocess
import sys

# Set the number of neurons and synapses
N = 2500
Nsyn = 1000

ROOT = os.path.dirname(os.path.abspath(__file__))
CLI = os.path.join(ROOT, "gpu4snn_amd", "bin", "gpu4snn")
LIB = os.path.join(ROOT, "gpu4snn_amd", "libgpu4snn.so")


def main(argv):
    if len(argv) < 3:
        print("usage: python gpu4snn.py GPU_NAME ITERATIONS [--rebuild] [CLI flags]", file=sys.stderr)
        return 2
    gpu_name, iterations = argv[1], argv[2]
    extra = [a for a in argv[3:] if a != "--rebuild"]
    shutil.rmtree(os.path.join(os.getcwd(), "Results"), ignore_errors=True)
    if "--rebuild" in argv or not (os.path.exists(CLI) and os.path.exists(LIB)):
        sys.path.insert(0, ROOT)
        from gpu4snn_amd import build
        build.build(verbose=True)
    os.makedirs(gpu_name, exist_ok=True)
    rc = subprocess.call([CLI, str(N), str(Nsyn), str(iterations), *extra])
    dest = os.path.join(".", gpu_name, str(iterations))
    os.makedirs(dest, exist_ok=True)
    if os.path.isdir("Results"):
        shutil.rmtree(os.path.join(dest, "Results"), ignore_errors=True)
        shutil.move("Results", dest)
    return rc


if __name__ == "__main__":
    sys.exit(main(sys.argv))
