"""The 15-point DFT of M2F and F2L at order 8 (defmm_amd/csrc/dft.hpp,
prime-factor algorithm), compiled for the CPU and checked against the direct
sum in long double: forward and inverse, 15 inputs and the 8-input form of
M2F (whose zero inputs must not be read: they hold NaN in the check), on
impulses and random inputs, in double and float."""
import os
import re
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_dft15(tmp_path):
    exe = tmp_path / "dft_check"
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-Wall",
                           os.path.join(ROOT, "tests", "cpp", "dft_check.cpp"), "-o", str(exe)])
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120, check=True).stdout
    errs = dict((k, float(v)) for k, v in re.findall(r"(\w+) ([0-9.eE+-]+)", out))
    assert len(errs) == 8, out
    for k, e in errs.items():
        # a few units in the last place of the largest output
        assert e < (1e-15 if k.startswith("f64") else 5e-7), (k, e)
