"""The device set-up's algorithm (defmm_amd/csrc/setup.hip) on the CPU: a
C++ model of its steps (level-synchronous octree, level-synchronous walk in
the wavefronts' order, smallest pair per offset, offsetCodes, marks, stable
per-slot placement, group merge, per-leaf P2P merge) plugged into the plan
builder's SetupDevice hook must give the host builder's plan array for array
(trees, marks, slots, tables, M2L CSR and groups, P2P ranges, M2M/L2L lists,
statistics) on eight workloads: HF and
LF levels, Laplace, fixed depth, distinct clouds, order 8, a leaf target
root.  This pins, without a GPU, the host functions the device path uses
(m2lWindow, offsetCodes) and the order claims it rests on."""
import os
import re
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_setup_model_equals_host_builder(tmp_path):
    exe = tmp_path / "setup_model"
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-Wall", "-fopenmp", "-ffp-contract=off",
                           os.path.join(ROOT, "tests", "cpp", "setup_model.cpp"),
                           os.path.join(ROOT, "defmm_amd", "csrc", "plan.cpp"), "-o", str(exe)])
    env = dict(os.environ, OMP_NUM_THREADS=os.environ.get("OMP_NUM_THREADS", "8"))
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300, env=env)
    print(run.stdout)
    ok = re.findall(r"^case (\S+) OK", run.stdout, re.M)
    assert run.returncode == 0 and len(ok) == 8, run.stdout + run.stderr
