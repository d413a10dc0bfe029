"""Build libifgf_amd.so in-tree: host C++ with g++, HIP kernels with hipcc for gfx950."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libifgf_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-pthread", "-ffp-contract=off", "-fno-fast-math",
              "-Wall", "-Wno-sign-compare"]
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build(verbose=True):
    """Compile every translation unit (in parallel) and link libifgf_amd.so."""
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    for src in ("plan.cpp", "plan_io.cpp", "capi.cpp"):
        jobs.append((["g++", *HOST_FLAGS, "-c", os.path.join(CSRC, src)],
                     os.path.join(BUILD, src + ".o")))
    for src in ("device.hip", "batched.hip", "tree.hip", "cones.hip", "pairs.hip"):
        jobs.append(([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src)],
                     os.path.join(BUILD, src + ".o")))
    procs = []
    for cmd, obj in jobs:
        full = cmd + ["-o", obj]
        print(" ".join(full), flush=True)
        procs.append((full, subprocess.Popen(full)))
    failed = [full for full, p in procs if p.wait() != 0]
    if failed:
        raise subprocess.CalledProcessError(1, failed[0])
    objs = [obj for _, obj in jobs]
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp, "-pthread"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
