"""The front-end's single point (runtype 1) on the GPU against the reference:
the input file c1.inp holds the geometry, grid and Davidson settings of the
reference Davidson fixture davidson_c1 (32^3, (H2O)6, one state, pTol 9); the
same geometry's neutral-cluster energy E0 comes from the reference's own
WaterCluster::CalcEnergy (water_c1.npz)."""
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
    ebe = float(re.search(r"EBE (\S+) Eh", r.stdout).group(1))
    assert abs(ebe - float(ref["evals"][0])) <= tol
    # the energy block of src/ClusterAnion.cpp:306-311, six significant digits
    # as the reference's cout prints them
    e0_ref = float(golden("water_c1.npz")["e0"])
    e0 = float(re.search(r"E0 (\S+) Eh", r.stdout).group(1))
    assert abs(e0 - e0_ref) <= 1e-12 * abs(e0_ref)
    blk = {k: float(re.search(r"%s\s+=\s+(\S+) meV" % k, r.stdout).group(1))
           for k in ("Eneutral", "EBE", "Etotal")}
    assert abs(blk["Eneutral"] - e0_ref * AU2MEV) <= 1e-5 * abs(e0_ref * AU2MEV)
    assert abs(blk["EBE"] - ebe * AU2MEV) <= 1e-5 * abs(ebe * AU2MEV)
    assert abs(blk["Etotal"] - (ebe + e0_ref) * AU2MEV) <= 1e-5 * abs((ebe + e0_ref) * AU2MEV)
    assert float(re.search(r"Energy of the neutral cluster E0 = (\S+)", r.stdout).group(1)) \
        == pytest.approx(e0_ref, rel=1e-5)
    # DVR::ExpectationValues: the state's norm and its size
    assert re.search(r"normalization of state0 is 1\b", r.stdout), r.stdout
    m = re.search(r"\n\s+0\s+(\S+)\s+(\S+)\s+(\S+)\n", r.stdout)
    r_mean, r_rms, spread = (float(v) for v in m.groups())
    assert 0 <= r_mean <= r_rms and spread > 0


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
    ebe = float(re.search(r"EBE (\S+) Eh", r.stdout).group(1))
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
    ebe = float(re.search(r"EBE (\S+) Eh", r.stdout).group(1))
    cl = oracle_mod.Cluster(xyz)
    g = oracle_mod.Grid((24, 24, 24), (36.0, 36.0, 36.0))
    V = cl.potential_grid_pol3(g, (3, 2, 0, 3, 0, 0))
    ro = g.solve(V, 1, 20, 200, 9, 1, g.piab().reshape(1, -1))
    assert ro["nconv"] == 1
    # the GPU water model's SCF and the oracle's stop at a mean change < 1e-12
    assert abs(ebe - float(ro["evals"][0])) <= 1e-8


def test_single_point_dual_grid(tmp_path):
    """Polarization = 6 in the input file: the dual-grid single point (a coarse
    solve, the potential carried over beyond Rtol, the fine solve started from
    the coarse ground state) agrees with the direct PolType 3 single point on
    the same grid when Rtol is large enough that every point is exact (then
    only the start vector differs), in fewer matvecs."""
    ref = golden("davidson_c1.npz")
    xyz = ref["xyz"]

    def run(pol, rtol):
        lines = ["BeginJob", "  runtype = 1,", "EndJob", "BeginElectronPotential",
                 "  Polarization = %d," % pol, "  Rtol = %g," % rtol, "EndElectronPotential",
                 "BeginGridDef", "  NoOfGridPoints = 33, 33, 33,",
                 "  Length = 36.0, 36.0, 36.0,", "EndGridDef", "BeginDiag", "  nStates = 1,",
                 "  maxSubspace = 20,", "  maxIter = 200,", "  pTol = 9,", "EndDiag",
                 "BeginWaters"]
        for w in xyz.reshape(-1, 9):
            for a, sym in enumerate(("O", "H", "H")):
                lines.append("%s %.17g %.17g %.17g" % ((sym,) + tuple(w[3 * a:3 * a + 3])))
        lines.append("EndWaters")
        path = tmp_path / ("p%d.inp" % pol)
        path.write_text("\n".join(lines) + "\n")
        r = subprocess.run([BIN, str(path)], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        ebe = float(re.search(r"EBE (\S+) Eh", r.stdout).group(1))
        mv = int(re.search(r"matvecs (\d+), iterations", r.stdout).group(1))
        return ebe, mv, r.stdout

    e3, mv3, _ = run(3, 20.0)
    e6, mv6, out6 = run(6, 60.0)   # every fine point within 60 bohr of a site: all exact
    assert "Energies For pre-calculations" in out6
    assert abs(e6 - e3) <= 1e-8
    assert mv6 < mv3
