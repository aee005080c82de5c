"""Build recipe of the oracle (test infrastructure only).

liboracle.so          strict IEEE double build (no contraction), portable
                      x86-64-v3 code: the checker used by tests and the CPU
                      baseline of bench.py.
liboracle_refround.so the same source with ORACLE_REFERENCE_ROUNDING, built
                      with the reference's own makefile flags (makefile:4,
                      g++ -O3 -ffast-math -mfma -march=native ...), used only
                      here to show that the restatement reproduces the
                      reference's printed known answer digit for digit.  It is
                      tied to this machine (-march=native) and never shipped.

The reference itself is not built: its FFT header (fftw3.h) is not part of
this image.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "oracle.cpp")
LIB = os.path.join(HERE, "liboracle.so")
LIB_REFROUND = os.path.join(HERE, "liboracle_refround.so")

STRICT = ["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
          "-march=x86-64-v3"]
REFERENCE_FLAGS = ["g++", "-Wall", "-mprefer-vector-width=512", "-O3", "-ffast-math", "-mfma",
                   "-fno-trapping-math", "-march=native", "-mtune=native", "-fopenmp-simd",
                   "-mavx", "-std=c++17", "-fPIC", "-shared", "-fopenmp",
                   "-DORACLE_REFERENCE_ROUNDING"]


def _build(cmd, out):
    tmp = out + ".tmp"
    print(" ".join(cmd + [SRC, "-o", tmp]), flush=True)
    subprocess.check_call(cmd + [SRC, "-o", tmp])
    os.replace(tmp, out)
    return out


def build(refround=False):
    _build(STRICT, LIB)
    if refround:
        _build(REFERENCE_FLAGS, LIB_REFROUND)
    return LIB


if __name__ == "__main__":
    build(refround="--refround" in sys.argv)
