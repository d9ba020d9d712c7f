"""Builds libpharos_amd.so in-tree: gfx950 HIP kernels (hipcc) plus the host C++ (g++).

Run as ``python -m pharos_amd.build``.  Every HIP object is built with -ffp-contract=off (the reference fuses
only where it writes @mulAdd) and without packed FP32 instructions.
"""
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpharos_amd.so")
ROOT = os.path.dirname(HERE)
C_EXAMPLE = os.path.join(ROOT, "tests", "c", "run_pull")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = os.environ.get("PHAROS_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["synth_engine.hip"]
CXX_SOURCES = ["signals.cpp", "scenario.cpp", "output.cpp", "run_api.cpp", "trajectory.cpp"]

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # Branches on wave-uniform values (a channel's kind) stay plain scalar branches instead of structurised flow
    # blocks, so each channel's loop body updates the channel sums in place: no copies of the 64 sums per channel.
    "-mllvm", "-structurizecfg-skip-uniform-regions=true",
    # Single-lane atomics stay plain: the optimizer's lane scan reads an atomic's result right after it, so the wave
    # would wait out the work queue's atomic where the kernel means it to return during the channel loop.
    "-mllvm", "-amdgpu-atomic-optimizer-strategy=None",
    "-Wall", "-Wno-unused-parameter", "-Wno-unused-function",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall", "-Wno-unused-parameter",
             f"-I{ROCM}/include", "-D__HIP_PLATFORM_AMD__"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose=True):
    """Compiles every source and links libpharos_amd.so next to this file (replaced atomically)."""
    if not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    tmp = tempfile.mkdtemp(prefix="pharos_build_")
    try:
        objs = []
        for src in HIP_SOURCES:
            obj = os.path.join(tmp, src + ".o")
            _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
            objs.append(obj)
        for src in CXX_SOURCES:
            path = os.path.join(CSRC, src)
            obj = os.path.join(tmp, src + ".o")
            _run(["g++", *CXX_FLAGS, "-c", path, "-o", obj])
            objs.append(obj)
        out_tmp = LIB + ".tmp"
        _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", out_tmp, *objs, "-lpthread", "-ldl"])
        os.replace(out_tmp, LIB)
        # A plain C caller of include/pharos_run.h (tests/c/run_pull.c), linked against the library as a C program
        # written for the reference's run API would be.
        c_src = os.path.join(ROOT, "tests", "c", "run_pull.c")
        if os.path.exists(c_src):
            exe_tmp = C_EXAMPLE + ".tmp"
            _run(["gcc", "-std=c11", "-O2", "-Wall", "-Wextra", "-I", os.path.join(ROOT, "include"), c_src, "-o",
                  exe_tmp, "-L", HERE, "-lpharos_amd", f"-Wl,-rpath,{HERE}", f"-Wl,-rpath,{ROCM}/lib"])
            os.replace(exe_tmp, C_EXAMPLE)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
