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


def test_single_point_fft_kinetic(tmp_path, oracle_mod):
    """DVRType = 3 in the input file: the front-end's single point on the FFT
    kinetic energy against the oracle's Davidson on the same grid and potential."""
    ref = golden("davidson_c1.npz")
    xyz = ref["xyz"]
    lines = ["BeginJob", "  runtype = 1,", "EndJob", "BeginGridDef", "  DVRType = 3,",
             "  NoOfGridPoints = 24, 24, 24,", "  Length = 36.0, 36.0, 36.0,", "EndGridDef",
             "BeginDiag", "  nStates = 1,", "  maxSubspace = 20,", "  maxIter = 200,",
             "  pTol = 9,", "EndDiag", "BeginWaters"]
    for w in xyz.reshape(-1, 9):
        for a, sym in enumerate(("O", "H", "H")):
            lines.append("%s %.17g %.17g %.17g" % ((sym,) + tuple(w[3 * a:3 * a + 3])))
    lines.append("EndWaters")
    path = tmp_path / "fft.inp"
    path.write_text("\n".join(lines) + "\n")
    r = subprocess.run([BIN, str(path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    ebe = float(re.search(r"EBE\s+=\s+(\S+) meV", r.stdout).group(1)) / AU2MEV
    cl = oracle_mod.Cluster(xyz)
    g = oracle_mod.Grid((24, 24, 24), (36.0, 36.0, 36.0), dvr_type=3)
    ro = g.solve(cl.potential_grid(g), 1, 20, 200, 9, 1, g.piab().reshape(1, -1))
    assert ro["nconv"] == 1
    assert abs(ebe - float(ro["evals"][0])) <= 1e-8


def test_single_point_poltype3(tmp_path, oracle_mod):
    """Polarization = 3 in the input file: self-consistent polarizable sites
    through the front-end against the oracle's PolType 3 potential and Davidson."""
    ref = golden("davidson_c1.npz")
    xyz = ref["xyz"]
    lines = ["BeginJob", "  runtype = 1,", "EndJob", "BeginElectronPotential",
             "  Polarization = 3,", "EndElectronPotential", "BeginGridDef",
             "  NoOfGridPoints = 24, 24, 24,", "  Length = 36.0, 36.0, 36.0,", "EndGridDef",
             "BeginDiag", "  nStates = 1,", "  maxSubspace = 20,", "  maxIter = 200,",
             "  pTol = 9,", "EndDiag", "BeginWaters"]
    for w in xyz.reshape(-1, 9):
        for a, sym in enumerate(("O", "H", "H")):
            lines.append("%s %.17g %.17g %.17g" % ((sym,) + tuple(w[3 * a:3 * a + 3])))
    lines.append("EndWaters")
    path = tmp_path / "pol3.inp"
    path.write_text("\n".join(lines) + "\n")
    r = subprocess.run([BIN, str(path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    ebe = float(re.search(r"EBE\s+=\s+(\S+) meV", r.stdout).group(1)) / AU2MEV
    cl = oracle_mod.Cluster(xyz)
    g = oracle_mod.Grid((24, 24, 24), (36.0, 36.0, 36.0))
    V = cl.potential_grid_pol3(g, (3, 2, 0, 3, 0, 0))
    ro = g.solve(V, 1, 20, 200, 9, 1, g.piab().reshape(1, -1))
    assert ro["nconv"] == 1
    # the GPU water model's SCF and the oracle's stop at a mean change < 1e-12
    assert abs(ebe - float(ro["evals"][0])) <= 1e-8
