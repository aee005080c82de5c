"""Build the in-tree shared library libdefmm_amd.so for gfx950.

The host plan builder (csrc/plan.cpp) is compiled by g++ without
floating-point contraction (its discrete choices must not depend on the
compiler); the device executor (csrc/fmm.hip), the direct sum (csrc/direct.hip) and
the device set-up (csrc/setup.hip, also without contraction: it repeats the
plan builder's decisions) by hipcc for gfx950 with packed FP32 instructions
disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libdefmm_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("DEFMM_OFFLOAD_ARCH", "gfx950")


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    plan_o = os.path.join(BUILD, "plan.o")
    fmm_o = os.path.join(BUILD, "fmm.o")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-fopenmp",
          "-c", os.path.join(CSRC, "plan.cpp"), "-o", plan_o])
    direct_o = os.path.join(BUILD, "direct.o")
    setup_o = os.path.join(BUILD, "setup.o")
    for src, obj, contract in (("fmm.hip", fmm_o, "fast"), ("direct.hip", direct_o, "fast"),
                               ("setup.hip", setup_o, "off")):
        _run([HIPCC, "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
              "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
              f"-ffp-contract={contract}", "-Wno-unused-result",
              "-c", os.path.join(CSRC, src), "-o", obj])
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", plan_o, fmm_o, direct_o, setup_o,
          "-o", tmp, "-lgomp", "-lpthread"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
