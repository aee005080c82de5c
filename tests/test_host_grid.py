"""Launch-grid arithmetic of the executor, checked on the host (no GPU): a
plain C++ program includes defmm_amd/csrc/grid.hpp and checks coverage and
the 2^32 work-item bound of every grid the executor builds."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_grid_arithmetic(tmp_path):
    exe = tmp_path / "grid_check"
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-Wall",
                           os.path.join(ROOT, "tests", "cpp", "grid_check.cpp"), "-o", str(exe)])
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stdout
