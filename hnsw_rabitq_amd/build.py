"""Builds libhnswgpu.so in-tree for gfx950.

Host C++ (graph construction, distance order) is compiled with g++ and -ffp-contract=off so
only the explicit fmaf calls fuse; HIP sources are compiled with hipcc for gfx950 with
contraction off and packed FP32 ops disabled; everything links into one shared library with a
plain C ABI (include/hnswgpu.h).
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libhnswgpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HOST_SOURCES = ["host/hnsw_index.cpp"]
HIP_SOURCES = ["search.hip", "build.hip", "mvsearch.hip", "bruteforce.hip", "mlp.hip", "capi.cpp", "capi_bf.cpp", "capi_stop.cpp"]

GXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-march=x86-64-v3", "-pthread",
             "-fvisibility=hidden", "-Wall", "-Wno-unused-function"]
HIP_FLAGS = [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-fvisibility=hidden", "-Wno-unused-result", "-Wno-unused-command-line-argument"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def _obj(src):
    return os.path.join(OBJ, src.replace("/", "_") + ".o")


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    jobs = []
    for s in HOST_SOURCES:
        jobs.append(["g++"] + GXX_FLAGS + ["-c", os.path.join(CSRC, s), "-o", _obj(s)])
    for s in HIP_SOURCES:
        path = os.path.join(CSRC, s)
        if not os.path.exists(path):
            continue
        extra = ["-x", "hip"] if s.endswith(".cpp") else []
        jobs.append([HIPCC] + HIP_FLAGS + extra + ["-c", path, "-o", _obj(s)])
    workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", "8"))))
    with concurrent.futures.ThreadPoolExecutor(workers) as ex:
        outs = list(ex.map(_run, jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    objs = [_obj(s) for s in HOST_SOURCES] + [_obj(s) for s in HIP_SOURCES
                                               if os.path.exists(os.path.join(CSRC, s))]
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-pthread"])
    os.replace(tmp, LIB)
    # the C++ drop-in's test program (include/hnswgpu.hpp), linked against the library in-tree
    root = os.path.dirname(HERE)
    src = os.path.join(root, "tests", "cpp", "closer_first_test.cpp")
    if os.path.exists(src):
        _run(["g++", "-O2", "-std=c++17", "-Wall", "-I" + os.path.join(root, "include"), src, "-o",
              os.path.join(root, "tests", "cpp", "closer_first_test"), "-L" + HERE, "-lhnswgpu",
              "-Wl,-rpath,$ORIGIN/../../hnsw_rabitq_amd"])
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
