"""Builds libqgk (the HIP kernels + C-ABI) in-tree for gfx950.

    python pyqg_amd/build.py        (runs without importing the package)

The shared library lands next to this file as _qgk.so and is loaded by
pyqg_amd._lib with ctypes.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_qgk.so")
# (source, shared library) pairs, built in parallel
TARGETS = [("qgk.hip", "_qgk.so"), ("cnn.hip", "_qgcnn.so")]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # keep the compiler from forming packed-FP32 instructions: some of their
    # forms have given wrong results on this hardware, without an error
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build libqgk")


def build(verbose=False, extra=()):
    procs = []
    for src, lib in TARGETS:
        out = os.path.join(HERE, lib)
        tmp = out + ".tmp"
        cmd = [hipcc(), *HIPCC_FLAGS, "-I", os.path.join(ROOT, "include"), *extra,
               os.path.join(CSRC, src), "-o", tmp]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                       text=True), tmp, out))
    failed = False
    for p, tmp, out in procs:
        log, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(log)
            failed = True
        else:
            os.replace(tmp, out)
    if failed:
        raise RuntimeError("hipcc failed")
    return [out for _, _, out in procs]


if __name__ == "__main__":
    print(build(verbose=True))
