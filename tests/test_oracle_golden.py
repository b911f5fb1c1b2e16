"""The CPU restatement (oracle) against the reference's own outputs.

The fixtures in tests/golden/ were written by tests/golden/make_golden.py, which
ran the reference's sine_dvr, Potential::Setup/Evaluate and Davidson (compiled
from the reference sources by oracle/Makefile.ref) on seeded inputs.
"""
import numpy as np
import pytest

from conftest import golden


def test_tkin_matches_reference(oracle_mod):
    f = golden("tkin.npz")
    for i in range(int(f["ncases"])):
        n, L = f["n%d" % i]
        x, T = oracle_mod.sine_dvr(int(n), float(L), packed=True)
        np.testing.assert_array_equal(x, f["x%d" % i])
        np.testing.assert_array_equal(T, f["T%d" % i])


@pytest.mark.parametrize("tag", ["c1", "c2"])
@pytest.mark.parametrize("flagset", ["default", "adiabatic", "gaussdamp", "gto6", "gto12",
                                     "gto12_sigo", "norep", "nopol", "internal1"])
def test_potential_matches_reference(oracle_mod, tag, flagset):
    f = golden("potential_%s.npz" % tag)
    cl = oracle_mod.Cluster(f["xyz"])
    np.testing.assert_array_equal(cl.sites, f["sites"])
    v, parts = cl.potential_points(f["pts"], f["flags_" + flagset], f["potpara_" + flagset])
    E = f["E_" + flagset]
    scale = np.abs(E[:, 4]).max()
    assert np.abs(parts[:, :4] - E[:, :4]).max() <= 1e-14 * scale
    assert np.abs(v - E[:, 4]).max() <= 1e-14 * scale


@pytest.mark.parametrize("tag", ["g10_r1", "g12_r4", "g10_r3_small", "g10_r2_jd", "c1"])
def test_davidson_matches_reference(oracle_mod, tag):
    f = golden("davidson_%s.npz" % tag)
    cl = oracle_mod.Cluster(f["xyz"])
    g = oracle_mod.Grid(tuple(f["n"]), tuple(f["L"]))
    V = cl.potential_grid(g)
    p = [int(v) for v in f["params"]]
    r = g.solve(V, *p, f["start"])
    assert r["nconv"] == int(f["nconv"])
    assert abs(r["n_mtx"] - int(f["n_mtx"])) <= 2
    if r["nconv"] == p[0]:
        np.testing.assert_allclose(r["evals"], f["evals"], rtol=0, atol=1e-10)
    for i in range(p[0]):
        ov = abs(np.dot(r["vecs"][i], f["vecs"][i]))
        assert ov > 1 - 1e-8, (i, ov)


def test_water_damping_matches_reference(oracle_mod):
    """The water model's charge-dipole and dipole-dipole damping
    (ElectrostaticDCDamp / ElectrostaticDDDamp, src/DPP.cpp:1798-1812),
    evaluated by the reference's own functions."""
    f = golden("damp.npz")
    dc, dd = oracle_mod.damp(f["r"], f["ai"], f["aj"], f["a"])
    np.testing.assert_allclose(dc, f["dc"], rtol=1e-14, atol=1e-300)
    np.testing.assert_allclose(dd, f["dd"], rtol=1e-13, atol=1e-300)
