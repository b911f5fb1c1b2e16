"""Grid potential and water model on the GPU against the oracle (whose
Potential restatement is pinned bit-for-bit to the reference's
Potential::Evaluate by tests/test_oracle_golden.py)."""
import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu

FLAGSETS = ["default", "adiabatic", "gaussdamp", "gto6", "gto12", "gto12_sigo", "norep",
            "nopol", "internal1"]


@pytest.mark.parametrize("flagset", FLAGSETS)
def test_build_potential_lists(pisces, oracle_mod, flagset):
    f = golden("potential_c1.npz")
    cl = oracle_mod.Cluster(f["xyz"])
    n, L = (32, 32, 32), (36.0, 36.0, 36.0)
    g = oracle_mod.Grid(n, L)
    flags, pp = f["flags_" + flagset], f["potpara_" + flagset]
    Vo = cl.potential_grid(g, flags, pp)
    with pisces.Context(n, L) as ctx:
        ctx.build_potential(flags, pp, cl.sites, cl.q, cl.iq, cl.d, cl.id)
        V = ctx.get_potential()
    assert np.abs(V - Vo).max() <= 1e-13 * np.abs(Vo).max()


@pytest.mark.parametrize("flagset", FLAGSETS)
def test_site_lists_in_any_order(pisces, oracle_mod, flagset):
    """The charge and dipole lists in the cluster's order and reversed (the
    same terms, summed in the other order) give the oracle's potential to
    rounding, and the same potential as each other."""
    f = golden("potential_c1.npz")
    cl = oracle_mod.Cluster(f["xyz"])
    n, L = (24, 24, 24), (30.0, 30.0, 30.0)
    g = oracle_mod.Grid(n, L)
    flags, pp = f["flags_" + flagset], f["potpara_" + flagset]
    Vo = cl.potential_grid(g, flags, pp)
    iq, idd = np.asarray(cl.iq), np.asarray(cl.id)
    assert np.all(np.diff(iq) > 0) and np.all(np.diff(idd) > 0)
    d = np.asarray(cl.d).reshape(-1, 3)
    with pisces.Context(n, L) as ctx:
        ctx.build_potential(flags, pp, cl.sites, cl.q, iq, d, idd)
        Va = ctx.get_potential()
        ctx.build_potential(flags, pp, cl.sites, np.asarray(cl.q)[::-1], iq[::-1], d[::-1], idd[::-1])
        Vr = ctx.get_potential()
    scale = np.abs(Vo).max()
    assert np.abs(Va - Vo).max() <= 1e-13 * scale
    assert np.abs(Vr - Vo).max() <= 1e-13 * scale
    assert np.abs(Va - Vr).max() <= 1e-13 * scale


@pytest.mark.parametrize("which", ["iq", "id"])
def test_site_on_a_grid_point(pisces, oracle_mod, which):
    """A charge site (which = iq) or a dipole site (id) exactly on a grid point
    (zero distance): the damped forms stay finite and match the oracle."""
    f = golden("potential_c1.npz")
    cl = oracle_mod.Cluster(f["xyz"])
    n, L = (33, 33, 33), (36.0, 36.0, 36.0)
    g = oracle_mod.Grid(n, L)
    s0 = int(np.asarray(getattr(cl, which))[0])
    mid = np.array([g.x[a][n[a] // 2] for a in range(3)])
    sites = np.asarray(cl.sites, dtype=np.float64).reshape(-1, 3)
    sites = (sites - sites[s0]) + mid
    assert np.array_equal(sites[s0], mid)
    cl.sites = sites.reshape(np.shape(cl.sites))
    flags, pp = f["flags_default"], f["potpara_default"]
    Vo = cl.potential_grid(g, flags, pp)
    with pisces.Context(n, L) as ctx:
        ctx.build_potential(flags, pp, cl.sites, cl.q, cl.iq, cl.d, cl.id)
        V = ctx.get_potential()
    assert np.all(np.isfinite(V)) and np.all(np.isfinite(Vo))
    assert np.abs(V - Vo).max() <= 1e-13 * np.abs(Vo).max()


@pytest.mark.parametrize("nw,rad,seed,n,L", [(6, 3.5, 1000, (32, 32, 32), 36.0),
                                             (24, 6.0, 2000, (64, 64, 64), 40.0),
                                             (45, 8.0, 3000, (48, 48, 48), 48.0)])
def test_cluster_potential_gpu_water_model(pisces, oracle_mod, nw, rad, seed, n, L):
    import workload
    print("cluster", nw, n, flush=True)
    xyz, _ = workload.water_cluster(nw, rad, seed)
    cl = oracle_mod.Cluster(xyz)
    g = oracle_mod.Grid(n, (L, L, L))
    Vo = cl.potential_grid(g)
    with pisces.Context(n, (L, L, L)) as ctx:
        sites, mu, rc = ctx.build_cluster_potential(xyz)
        V = ctx.get_potential()
    assert rc == 0
    np.testing.assert_array_equal(sites, cl.sites)
    assert np.abs(mu - cl.mu).max() <= 1e-10 * np.abs(cl.mu).max()
    assert np.abs(V - Vo).max() <= 1e-11 * np.abs(Vo).max()


def test_cluster_potential_with_ion(pisces, oracle_mod):
    import workload
    xyz, ion = workload.water_cluster(10, 5.0, 3100, ion_min=2.3)
    ion_b = np.asarray(ion) * workload.ANGS2BOHR
    n, L = (32, 32, 32), (40.0, 40.0, 40.0)
    cl = oracle_mod.Cluster(xyz, ions=(ion_b, [1.0]), ion_r=2.9)
    g = oracle_mod.Grid(n, L)
    Vo = cl.potential_grid(g)
    with pisces.Context(n, L) as ctx:
        ctx.build_cluster_potential(xyz, ions=(ion_b, [1.0]), ion_r=2.9)
        V = ctx.get_potential()
    assert np.abs(V - Vo).max() <= 1e-11 * np.abs(Vo).max()


def test_diag_and_piab(pisces, oracle_mod):
    n, L = (20, 18, 16), (30.0, 28.0, 26.0)
    g = oracle_mod.Grid(n, L)
    rng = np.random.default_rng(4)
    V = rng.normal(size=g.ngp)
    with pisces.Context(n, L) as ctx:
        ctx.set_potential(V)
        np.testing.assert_array_equal(ctx.diag(), g.diag(V))
        np.testing.assert_array_equal(ctx.piab(), g.piab())
        xs, Ts = ctx.grid()
    for a in range(3):
        np.testing.assert_array_equal(xs[a], g.x[a])
        np.testing.assert_array_equal(Ts[a], g.T[a])
