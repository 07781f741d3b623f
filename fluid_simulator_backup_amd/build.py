"""Build libfs_hip.so (the product) and the CPU checker library.

    python -m fluid_simulator_backup_amd.build

hipcc compiles every csrc/*.hip file for gfx950 with FMA contraction off
(operation order is part of the parity contract) and packed-FP32 ops
disabled, g++ the host-only csrc/*.cpp files (the scene layer), then hipcc
links one shared library in the package directory.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libfs_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

CXX = os.environ.get("CXX_HOST", "g++")
# host-only C++ (the scene layer and its script reader): no device code
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function"]

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed: " + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def build_hip(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip") or f.endswith(".cpp"))
    objs = []
    jobs = []
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        for s in srcs:
            base, ext = os.path.splitext(s)
            o = os.path.join(BUILD, base + (".o" if ext == ".hip" else ".host.o"))
            objs.append(o)
            if ext == ".hip":
                cmd = [HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, s), "-o", o]
            else:
                cmd = [CXX] + CXX_FLAGS + ["-c", os.path.join(CSRC, s), "-o", o]
            jobs.append(ex.submit(_run, cmd))
        for j in jobs:
            r = j.result()
            if verbose and r.stderr:
                print(r.stderr, file=sys.stderr)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


def build_oracle():
    _run(["make", "-s", "-B", "-C", os.path.join(ROOT, "oracle")])
    return os.path.join(ROOT, "oracle", "libfs_oracle.so")


if __name__ == "__main__":
    print(build_hip(verbose=True))
    print(build_oracle())
