"""The oracle's water model against the reference's own WaterCluster.

The fixtures water_*.npz hold the output of the reference's water model
(WaterCluster + Water, src/DPP.cpp, src/Water.cpp) run by
oracle/_ref/ref_driver in the order ClusterAnion runs it
(src/ClusterAnion.cpp:59-66, 165-197): the projection of O,H,H input onto
rigid DPP monomers (Water::Reposition_A, src/Water.cpp:679-825), the
configuration round trip (GetConfiguration / SetConfiguration), the
induced-dipole SCF (CalcInducedDipoles, src/DPP.cpp:760-1065), the
neutral-cluster energy E0 (CalcIntermolecularPotential, src/DPP.cpp:1648-1700)
and the lists the electron's Potential receives (GetLists,
ReportInducedDipoles, ReportPolSitesAndField, src/DPP.cpp:402-642).

The ion term (Bloomfield's erf-damped charge, src/NaCl.cpp:285-318, which no
reference path combines with waters) is pinned by two invariants: with no
waters it is -q erf(R / r_Na) / R (checked against scipy's erf), and an ion
far from the cluster adds -q / R to the water potential.
"""
import numpy as np
import pytest

from conftest import golden

TAGS = ["c1", "c2", "c3", "nonrigid"]


@pytest.mark.parametrize("tag", TAGS)
def test_projection_matches_reference(oracle_mod, tag):
    f = golden("water_%s.npz" % tag)
    sites, conf = oracle_mod.water_sites(f["xyz"])
    scale = np.abs(f["sites"]).max()
    assert np.abs(sites - f["sites"]).max() <= 1e-13 * scale
    # CoM position and Euler angles of every monomer (Water::Reposition)
    assert np.abs(conf[:, :3] - f["conf"][:, :3]).max() <= 1e-13 * scale
    assert np.abs(conf[:, 3:] - f["conf"][:, 3:]).max() <= 1e-12


@pytest.mark.parametrize("tag", TAGS)
def test_dipoles_field_and_energy_match_reference(oracle_mod, tag):
    f = golden("water_%s.npz" % tag)
    ind = oracle_mod.water_induced(f["sites"])
    assert ind["converged"]
    mu_ref = f["mu"]
    assert np.abs(ind["mu"] - mu_ref).max() <= 1e-10 * np.abs(mu_ref).max()
    epc_ref = f["epc"]
    assert np.abs(ind["epc"] - epc_ref).max() <= 1e-12 * np.abs(epc_ref).max()
    assert abs(ind["e_pol"] - f["e_pol"]) <= 1e-12 * abs(f["e_pol"])
    e0, elec, pol, disp = oracle_mod.water_energy(f["sites"], ind["mu"])
    assert abs(e0 - f["e0"]) <= 1e-12 * max(1.0, abs(f["e0"]))
    assert abs(pol - f["e_pol"]) <= 1e-12 * abs(f["e_pol"])
    assert abs(elec + pol + disp - e0) <= 1e-14 * max(1.0, abs(e0))


@pytest.mark.parametrize("tag", TAGS)
def test_lists_match_reference(oracle_mod, tag):
    f = golden("water_%s.npz" % tag)
    nw = f["xyz"].shape[0]
    ind = oracle_mod.water_induced(f["sites"])
    q, iq, d, idd = oracle_mod.water_lists(f["sites"], ind["mu"])
    np.testing.assert_array_equal(q, f["q"])
    np.testing.assert_array_equal(iq, f["iq"])
    # the electron sees every induced dipole (the ClusterAnion override of the
    # one-per-water list, src/ClusterAnion.cpp:196-202)
    np.testing.assert_array_equal(idd, f["id_mu"])
    assert np.abs(d - f["mu"]).max() <= 1e-10 * np.abs(f["mu"]).max()
    # PolType 3 lists: polarizable sites O,H1,H2 (alpha, site index) and the
    # permanent-charge field there
    alpha, ip, epc, vww = oracle_mod.Cluster(f["xyz"]).pol3_lists()
    np.testing.assert_array_equal(ip, f["ipp"])
    assert np.abs(alpha - f["alpha"]).max() <= 1e-14 * alpha.max()
    assert np.abs(epc - f["epc"]).max() <= 1e-12 * np.abs(f["epc"]).max()
    assert abs(vww - f["e_pol"]) <= 1e-12 * abs(f["e_pol"])
    # the reference's collapsed one-per-water dipole at the M site is the sum
    # of that water's three induced dipoles
    dw = f["mu"].reshape(nw, 3, 3).sum(axis=1)
    assert np.abs(dw - f["d_water"]).max() <= 1e-15 * np.abs(dw).max() * 3
    np.testing.assert_array_equal(f["id_water"], 4 * np.arange(nw) + 3)


def test_projection_catches_a_wrong_euler_convention(oracle_mod):
    """A plausible slip, transposing the rotation (the inverse Euler rotation),
    moves the projected sites of the non-rigid geometry by far more than the
    comparison above allows."""
    f = golden("water_nonrigid.npz")
    sites, conf = oracle_mod.water_sites(f["xyz"])
    ref = sites.reshape(-1, 4, 3)
    com = conf[:, None, :3]
    # rotate back about each CoM by twice the Euler rotation's inverse: the
    # transposed convention puts sites at R^T (s - c) + c instead of R (s - c) + c
    c1, c2, c3 = np.cos(conf[:, 3]), np.cos(conf[:, 4]), np.cos(conf[:, 5])
    s1, s2, s3 = np.sin(conf[:, 3]), np.sin(conf[:, 4]), np.sin(conf[:, 5])
    R = np.stack([c1 * c2 * c3 - s1 * s3, -c1 * c2 * s3 - s1 * c3, c1 * s2,
                  s1 * c2 * c3 + c1 * s3, -s1 * c2 * s3 + c1 * c3, s1 * s2,
                  -s2 * c3, s2 * s3, c2], axis=1).reshape(-1, 3, 3)
    body = np.einsum("wji,waj->wai", R, ref - com)  # R^T (s - c): standard frame
    wrong = np.einsum("wji,waj->wai", R, body) + com  # R^T applied instead of R
    assert np.abs(wrong.reshape(-1, 3) - f["sites"]).max() > 1e-3


def _bare_cluster(oracle_mod, sites, ion_xyz, ion_q, ion_r):
    """A Cluster with the given site list and no water lists (no charges,
    dipoles or repulsive cores reach the potential)."""
    cl = oracle_mod.Cluster.__new__(oracle_mod.Cluster)
    cl.sites = np.ascontiguousarray(np.asarray(sites, dtype=np.float64).reshape(-1, 3))
    cl.q = np.zeros(0)
    cl.iq = np.zeros(0, dtype=np.int32)
    cl.d = np.zeros(0)
    cl.id = np.zeros(0, dtype=np.int32)
    cl.ion_xyz = np.ascontiguousarray(np.asarray(ion_xyz, dtype=np.float64).reshape(-1, 3))
    cl.ion_q = np.ascontiguousarray(np.asarray(ion_q, dtype=np.float64).reshape(-1))
    cl.ion_r = float(ion_r)
    return cl


def test_ion_alone_is_bloomfield_closed_form(oracle_mod):
    from scipy.special import erf
    rng = np.random.default_rng(5)
    ion = np.array([0.7, -1.1, 0.4])
    pts = ion + rng.normal(size=(300, 3)) * rng.uniform(0.05, 12.0, size=(300, 1))
    # the potential with Potential 3 flags but no water sites: only the ion
    cl = _bare_cluster(oracle_mod, np.zeros((0, 3)), ion, [1.0], 2.9)
    v, parts = cl.potential_points(pts, flags=(3, 2, 0, 0, 0, 0))
    R = np.linalg.norm(pts - ion, axis=1)
    ref = -1.0 * erf(R / 2.9) / R
    assert np.abs(v - ref).max() <= 1e-14 * np.abs(ref).max()
    # a doubled charge doubles it; a smaller radius is a different function
    cl2 = _bare_cluster(oracle_mod, np.zeros((0, 3)), ion, [2.0], 2.9)
    assert np.abs(cl2.potential_points(pts, flags=(3, 2, 0, 0, 0, 0))[0] - 2 * ref).max() \
        <= 1e-14 * np.abs(ref).max()
    cl3 = _bare_cluster(oracle_mod, np.zeros((0, 3)), ion, [1.0], 1.0)
    assert np.abs(cl3.potential_points(pts, flags=(3, 2, 0, 0, 0, 0))[0] - ref).max() > 1e-3


def test_far_ion_adds_its_coulomb_term(oracle_mod):
    f = golden("water_c1.npz")
    rng = np.random.default_rng(9)
    far = np.array([3.0e3, -2.0e3, 1.5e3])
    ions = (far, [1.0])
    with_ion = oracle_mod.Cluster(f["xyz"], ions=ions, ion_r=2.9)
    without = oracle_mod.Cluster(f["xyz"])
    pts = rng.uniform(-15, 15, size=(200, 3))
    for flags in [(3, 2, 0, 1, 0, 0), (3, 1, 0, 1, 0, 0), (2, 2, 0, 1, 0, 0)]:
        dv = with_ion.potential_points(pts, flags=flags)[0] - without.potential_points(pts, flags=flags)[0]
        R = np.linalg.norm(pts - far, axis=1)
        assert np.abs(dv - (-1.0 / R)).max() <= 1e-12 * (1.0 / R).max()
