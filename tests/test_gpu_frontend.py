"""The front-end's single point (runtype 1) on the GPU against the reference:
the input file c1.inp holds the geometry, grid and Davidson settings of the
reference Davidson fixture davidson_c1 (32^3, (H2O)6, one state, pTol 9)."""
import os
import re
import subprocess

import pytest

from conftest import golden

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "pisces-1_amd", "bin", "pisces_amd_run")
C1 = os.path.join(ROOT, "tests", "golden", "inputs", "c1.inp")
AU2MEV = 27211.38386


@pytest.mark.parametrize("precision, tol", [("fp64", 1e-8), ("bf16x2", 1e-6)])
def test_single_point_matches_reference(precision, tol):
    ref = golden("davidson_c1.npz")
    r = subprocess.run([BIN, C1, "--precision", precision], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert re.search(r"State 0\s+-?\d+\.\d{8} Hartree =\s+-?\d+\.\d{2} meV", r.stdout)
    ebe = float(re.search(r"EBE\s+=\s+(\S+) meV", r.stdout).group(1)) / AU2MEV
    assert abs(ebe - float(ref["evals"][0])) <= tol
