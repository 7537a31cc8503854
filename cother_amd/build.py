"""Build libcother_hip.so in-tree for gfx950.

HIP sources (*.hip) are compiled by hipcc for gfx950, host sources (*.cpp)
by g++, and everything is linked by hipcc into cother_amd/libcother_hip.so.
FP contraction is disabled everywhere: the single-precision operation order
of the kernels is part of their contract (explicit fmaf() where a fused
multiply-add is intended).
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
LIB = os.path.join(HERE, "libcother_hip.so")
ARCH = os.environ.get("CTH_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", f"-I{INCLUDE}",
]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-Wall",
             "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__", f"-I{INCLUDE}"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    for f in sorted(os.listdir(CSRC)):
        src = os.path.join(CSRC, f)
        obj = os.path.join(BUILD, f + ".o")
        if f.endswith(".hip"):
            jobs.append([HIPCC, *HIP_FLAGS, "-c", src, "-o", obj])
        elif f.endswith(".cpp"):
            jobs.append(["g++", *CXX_FLAGS, "-c", src, "-o", obj])
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        for r in ex.map(_run, jobs):
            if verbose and (r.stdout or r.stderr):
                print(r.stdout + r.stderr, file=sys.stderr)
    objs = [j[-1] for j in jobs]
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lgomp"])
    os.replace(tmp, LIB)
    return LIB


def build_oracle():
    d = os.path.join(os.path.dirname(HERE), "oracle")
    _run(["make", "-s", "-B", "-C", d])
    return os.path.join(d, "liboracle.so")


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_oracle())
