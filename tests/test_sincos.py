"""The double-precision sin/cos used by the FP64 near field and the FP64
P2M/L2P phases, and its float counterpart for the mixed path (defmm_amd/csrc/sincos_red.hpp), compiled for the CPU and
checked against long double sinl/cosl over 10M arguments up to 1e6 radians,
near multiples of pi/2 and tiny arguments."""
import os
import re
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_sincos_red_accuracy(tmp_path):
    exe = tmp_path / "sincos_check"
    subprocess.check_call(["g++", "-O2", "-std=c++17",
                           os.path.join(ROOT, "tests", "cpp", "sincos_check.cpp"), "-o", str(exe)])
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120, check=True).stdout
    es, ec = (float(v) for v in re.findall(r"max_abs_err_(?:sin|cos) ([0-9.eE+-]+)", out))
    # about one unit in the last place of a value near 1 (2.2e-16)
    assert es < 2.5e-16 and ec < 2.5e-16, out
    # float variant: about one float unit in the last place (1.2e-7)
    ef = float(re.findall(r"max_abs_err_float ([0-9.eE+-]+)", out)[0])
    assert ef < 1.5e-7, out
