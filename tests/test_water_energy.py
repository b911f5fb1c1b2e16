"""The library's neutral-cluster energy (pisces_water_energy, host code: no GPU)
against the reference's own WaterCluster::CalcEnergy (src/DPP.cpp:251-262,
1648-1700) on the reference's sites and induced dipoles (water_*.npz)."""
import ctypes

import numpy as np
import pytest

from conftest import golden


@pytest.mark.parametrize("tag", ["c1", "c2", "c3", "nonrigid"])
def test_water_energy_matches_reference(pisces, oracle_mod, tag):
    f = golden("water_%s.npz" % tag)
    nw = f["xyz"].shape[0]
    sites = np.ascontiguousarray(f["sites"].reshape(-1))
    mu = np.ascontiguousarray(f["mu"])
    e = np.zeros(4)
    L = pisces.lib()
    D = ctypes.POINTER(ctypes.c_double)
    L.pisces_water_energy.argtypes = [ctypes.c_int, D, D, D]
    assert L.pisces_water_energy(nw, sites.ctypes.data_as(D), mu.ctypes.data_as(D),
                                 e.ctypes.data_as(D)) == 0
    assert abs(e[0] - f["e0"]) <= 1e-12 * max(1.0, abs(f["e0"]))
    assert abs(e[2] - f["e_pol"]) <= 1e-12 * abs(f["e_pol"])
    # the oracle's split of the same energy
    eo = oracle_mod.water_energy(f["sites"], f["mu"])
    assert np.abs(e - eo).max() <= 1e-12 * max(1.0, np.abs(eo).max())
