"""Self-consistent polarizable sites (PolType 3) in the oracle.

Pinned four ways:
* against the reference's own Potential::Evaluate with PolType 3 (fixtures
  potential_pol3_*.npz, written by tests/golden/make_golden.py through the
  reference's Potential.cpp + MolPolAux.cpp, InvA by LAPACK dsytrf/dsytri);
* InvA E_pc must be the water model's self-consistent dipoles, which the
  oracle finds by a different method (0.5-mixed fixed-point iteration,
  src/DPP.cpp:953-1006) from the same Thole tensors;
* with VpolWaterWater = the water-water polarization energy the field-free
  limit of Vpol is 0 (the electron far away polarizes nothing);
* one polarizable site with no permanent field is the non-interacting
  PolType 1 term -alpha R^2 / (2 Reff^6) (the same effective-distance damping).
"""
import numpy as np
import pytest

from conftest import golden

POL3 = ["scp", "scp_gaussdamp", "scp_gto12"]


@pytest.mark.parametrize("tag", ["c1", "c2"])
@pytest.mark.parametrize("flagset", POL3)
def test_pol3_matches_reference(oracle_mod, tag, flagset):
    f = golden("potential_pol3_%s.npz" % tag)
    cl = oracle_mod.Cluster(f["xyz"])
    np.testing.assert_array_equal(cl.sites, f["sites"])
    np.testing.assert_array_equal(cl.epc, f["epc"])
    v, parts = cl.potential_points_pol3(f["pts"], f["flags_" + flagset], f["potpara_" + flagset])
    E = f["E_" + flagset]
    scale = max(1.0, np.abs(E).max())
    # Vpc, Vind (0: dipoles ignored), Vrep, Vpol, Vtotal
    assert np.abs(parts - E).max() <= 1e-13 * scale
    assert np.all(E[:, 1] == 0.0)


@pytest.mark.parametrize("tag", ["c1", "c2"])
def test_inva_epc_is_the_scf_dipoles(oracle_mod, tag):
    f = golden("potential_pol3_%s.npz" % tag)
    cl = oracle_mod.Cluster(f["xyz"])
    alpha, ip, epc, _ = cl.pol3_lists()
    A = oracle_mod.pol3_inva(cl.sites[ip], alpha)
    mu = A @ epc
    # the SCF stops at a mean change below 1e-12: ~1e-10 left in mu
    assert np.abs(mu - cl.mu).max() <= 1e-8 * np.abs(cl.mu).max()
    # and InvA is symmetric positive definite (a physical polarizability)
    assert np.abs(A - A.T).max() <= 1e-12 * np.abs(A).max()
    assert np.linalg.eigvalsh(0.5 * (A + A.T)).min() > 0.0


@pytest.mark.parametrize("tag", ["c1", "c2"])
def test_field_free_limit(oracle_mod, tag):
    f = golden("potential_pol3_%s.npz" % tag)
    cl = oracle_mod.Cluster(f["xyz"])
    alpha, ip, epc, e_pol = cl.pol3_lists()
    A = oracle_mod.pol3_inva(cl.sites[ip], alpha)
    c0 = -0.5 * epc @ (A @ epc) - e_pol
    assert abs(c0) <= 1e-8 * abs(e_pol)
    # points 1e4 bohr away: what remains is the electron's field against the
    # water-induced dipoles, -sum_i e_i . mu_i (O(1e-9)); the next term,
    # -e^T InvA e / 2, is O(r^-4) ~ 1e-14
    far = f["pts"][-8:]
    _, parts = cl.potential_points_pol3(far, f["flags_scp"], f["potpara_scp"])
    mu = (A @ epc).reshape(-1, 3)
    R = far[:, None, :] - cl.sites[ip][None, :, :]
    e = R / np.linalg.norm(R, axis=2)[:, :, None] ** 3
    lead = -np.einsum("pic,ic->p", e, mu)
    assert np.abs(lead).max() > 1e-10
    np.testing.assert_allclose(parts[:, 3], lead + c0, rtol=0, atol=1e-13)


def test_one_site_is_poltype1(oracle_mod):
    f = golden("potential_pol3_c1.npz")
    cl = oracle_mod.Cluster(f["xyz"][:1])  # one water: its M site polarizable
    pp = oracle_mod.default_potpara()
    rng = np.random.default_rng(5)
    m = cl.sites[3]
    d = rng.normal(size=(200, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    pts = m + d * rng.uniform(0.05, 12.0, size=(200, 1))  # inside and outside pp[1]
    pol = (np.array([pp[0]]), np.array([3], dtype=np.int32), np.zeros(3), 0.0)
    _, p3 = cl.potential_points_pol3(pts, (3, 2, 0, 3, 0, 0), pp, pol=pol)
    _, p1 = cl.potential_points(pts, (3, 2, 0, 1, 0, 0), pp)
    np.testing.assert_allclose(p3[:, 3], p1[:, 3], rtol=1e-13, atol=0)
    assert np.any(np.linalg.norm(pts - m, axis=1) < pp[1])


def test_flags_without_sites_are_rejected(oracle_mod):
    f = golden("potential_pol3_c1.npz")
    cl = oracle_mod.Cluster(f["xyz"])
    with pytest.raises(ValueError):
        cl.potential_points(f["pts"][:4], (3, 2, 0, 3, 0, 0))
