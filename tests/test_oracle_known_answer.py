"""Pins the oracle to the reference's own known answers.

The reference's only test (test/testinterface.cpp) evaluates the refined-cube
mesh (mesh/refinedcube.inp, 41,334 points, kappa = 1, Ncrit = 32, charges
rand()/RAND_MAX from glibc's default seed) and prints the relative L-inf /
L1 / L2 errors of the FMM against the direct sum.  Its values at L = 6 are
1.38628e-05 / 1.22574e-05 / 1.13715e-05; at L = 3 and 4 the L2 errors are
6.30e-3 and 1.12e-3; the per-pass call counts of that run are 129,928 P2P,
336,776 HF M2L, 17,334 LF M2L, 747 kernel tables, 2,193 HF / 805 LF leaves,
2,508 / 733 / 96 M2M (hf / lf->hf / lf) and 36,773 target slots.

The strict-IEEE oracle breaks exact ties of the best-cone search toward the
first cone; the reference breaks them by its -ffast-math rounding, which
changes 4 of 36,773 slots and the errors in the fifth or sixth digit.  Built
with the reference's flags (ORACLE_REFERENCE_ROUNDING, this machine only),
the same source reproduces all printed digits and the slot count.
"""
import os
import platform
import subprocess

import numpy as np
import pytest

import oracle_api as oa

pytestmark = pytest.mark.slow

KNOWN_L6 = (1.38628e-05, 1.22574e-05, 1.13715e-05)
COUNTS_L6 = dict(p2p_calls=129928, hf_m2l=336776, lf_m2l=17334, hf_leaves=2193, lf_leaves=805,
                 hf_m2m=2508, lf2hf_m2m=733, lf_m2m=96, hf_l2l=2508, hf2lf_l2l=733, lf_l2l=96,
                 nlvl=9, cpt=6, max_stage=3, ncells_t=3698, ncells_s=3698)


def test_refined_cube_L6_strict(cube, cube_charges, cube_exact):
    p, st, _, _ = oa.fmm(cube, cube, cube_charges, 6, 32, 1.0)
    errs = oa.rel_errors(cube_exact, p)
    for k, v in COUNTS_L6.items():
        assert st[k] == v, k
    assert st["hf_tables"] + st["lf_tables"] == 747
    assert abs(st["trg_slots"] - 36773) <= 8
    np.testing.assert_allclose(errs, KNOWN_L6, rtol=2e-4)


@pytest.mark.parametrize("L,l2", [(3, 6.30e-3), (4, 1.12e-3)])
def test_refined_cube_error_sweep(cube, cube_charges, cube_exact, L, l2):
    p, _, _, _ = oa.fmm(cube, cube, cube_charges, L, 32, 1.0)
    errs = oa.rel_errors(cube_exact, p)
    assert abs(errs[2] - l2) / l2 < 5e-3


@pytest.mark.skipif(platform.machine() != "x86_64", reason="reference flags target x86-64")
def test_refined_cube_L6_reference_rounding(cube, cube_charges, cube_exact):
    path = os.path.join(oa.ORACLE_DIR, "liboracle_refround.so")
    try:
        lib = oa.load(path)
    except (subprocess.CalledProcessError, OSError) as e:  # no g++ / no such CPU features
        pytest.skip(f"reference-flag build unavailable: {e}")
    p, st, _, _ = oa.fmm(cube, cube, cube_charges, 6, 32, 1.0, lib=lib)
    errs = oa.rel_errors(cube_exact, p)
    assert st["trg_slots"] == 36773
    np.testing.assert_allclose(errs, KNOWN_L6, rtol=1e-5)
