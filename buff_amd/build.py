"""Builds libbuffgpu.so (HIP, gfx950) in-tree: buff_amd/libbuffgpu.so.

Run as `python -m buff_amd.build` or call build(). Always recompiles.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libbuffgpu.so")
CLI = os.path.join(HERE, "comp_profiler")
SOURCES = ["buff_kernels.hip", "buff_major.hip", "buff_batch.hip", "buff_cond.hip", "buff_wmax.hip", "buff_api.hip",
           "buff_major_api.hip", "buff_batch_api.hip", "buff_slice_api.hip", "buff_csv.hip", "buff_noalloc.hip", "buff_segment.hip"]
ARCH = os.environ.get("BUFF_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=" + ARCH,
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    # packed FP32 ops are kept out of generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-command-line-argument",
]


def _compile(src, verbose):
    obj = os.path.join(CSRC, src.replace(".hip", ".o"))
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return obj


def build(verbose=False):
    from concurrent.futures import ThreadPoolExecutor

    jobs = max(1, min(len(SOURCES), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))))
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda src: _compile(src, verbose), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    # the reference-shaped CLI on top of the C ABI (host C++ only)
    cli = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-Wall", "-I", os.path.join(HERE, "..", "include"),
           os.path.join(HERE, "host", "comp_profiler.cpp"), "-o", CLI, "-L", HERE, "-lbuffgpu",
           "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cli), flush=True)
    subprocess.run(cli, check=True)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
