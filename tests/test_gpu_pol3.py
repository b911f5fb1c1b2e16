"""Self-consistent polarizable sites (PolType 3) on the GPU against the oracle,
whose PolType 3 restatement is pinned to the reference's Potential::Evaluate by
tests/test_oracle_pol3.py."""
import ctypes

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu

POL3 = ["scp", "scp_gaussdamp", "scp_gto12"]


@pytest.mark.parametrize("flagset", POL3)
def test_build_potential_scp_c1(pisces, oracle_mod, flagset):
    """(H2O)6 on a 32^3 grid, the whole grid."""
    f = golden("potential_pol3_c1.npz")
    cl = oracle_mod.Cluster(f["xyz"])
    n, L = (32, 32, 32), (36.0, 36.0, 36.0)
    g = oracle_mod.Grid(n, L)
    flags, pp = f["flags_" + flagset], f["potpara_" + flagset]
    Vo = cl.potential_grid_pol3(g, flags, pp)
    alpha, ip, epc, e_pol = cl.pol3_lists()
    with pisces.Context(n, L) as ctx:
        ctx.build_potential_scp(flags, pp, cl.sites, cl.q, cl.iq, alpha, ip, epc, e_pol)
        V = ctx.get_potential()
    assert np.all(np.isfinite(V))
    assert np.abs(V - Vo).max() <= 1e-12 * np.abs(Vo).max()


def test_build_potential_scp_c2(pisces, oracle_mod):
    """(H2O)24 on a 64^3 grid (216 polarizable dimensions), against the oracle
    at 4096 grid points."""
    f = golden("potential_pol3_c2.npz")
    cl = oracle_mod.Cluster(f["xyz"])
    n, L = (64, 64, 64), (40.0, 40.0, 40.0)
    g = oracle_mod.Grid(n, L)
    flags, pp = f["flags_scp"], f["potpara_scp"]
    alpha, ip, epc, e_pol = cl.pol3_lists()
    with pisces.Context(n, L) as ctx:
        ctx.build_potential_scp(flags, pp, cl.sites, cl.q, cl.iq, alpha, ip, epc, e_pol)
        V = ctx.get_potential()
    rng = np.random.default_rng(3)
    idx = rng.choice(g.ngp, size=4096, replace=False)
    i, j, k = idx % 64, (idx // 64) % 64, idx // 4096
    pts = np.stack([g.x[0][i], g.x[1][j], g.x[2][k]], axis=1)
    Vo, _ = cl.potential_points_pol3(pts, flags, pp)
    assert np.abs(V[idx] - Vo).max() <= 1e-12 * np.abs(Vo).max()


def test_cluster_potential_poltype3(pisces, oracle_mod):
    """Water model on the GPU (dipole SCF, Epc, VpolWaterWater) + PolType 3
    potential, against the oracle's water model and potential."""
    f = golden("potential_pol3_c1.npz")
    cl = oracle_mod.Cluster(f["xyz"])
    n, L = (32, 32, 32), (36.0, 36.0, 36.0)
    g = oracle_mod.Grid(n, L)
    flags = (3, 2, 0, 3, 0, 0)
    Vo = cl.potential_grid_pol3(g, flags)
    with pisces.Context(n, L) as ctx:
        sites, mu, rc = ctx.build_cluster_potential(f["xyz"], flags=flags)
        V = ctx.get_potential()
    assert rc == 0
    np.testing.assert_allclose(sites, cl.sites, rtol=0, atol=1e-12)
    # two converged SCFs (stop at a mean change < 1e-12) feed Epc and VpolWW
    assert np.abs(V - Vo).max() <= 1e-9 * np.abs(Vo).max()


def test_polarizable_site_on_a_grid_point(pisces, oracle_mod):
    """A polarizable site exactly on a grid point: DampType 2 gives the field
    0 there (Reff is finite), as in the reference."""
    f = golden("potential_pol3_c1.npz")
    cl = oracle_mod.Cluster(f["xyz"])
    n, L = (33, 33, 33), (36.0, 36.0, 36.0)
    g = oracle_mod.Grid(n, L)
    alpha, ip, epc, e_pol = cl.pol3_lists()
    mid = np.array([g.x[a][n[a] // 2] for a in range(3)])
    sites = np.asarray(cl.sites, dtype=np.float64).reshape(-1, 3)
    sites = (sites - sites[ip[0]]) + mid
    assert np.abs(sites[ip[0]] - mid).max() == 0.0
    flags, pp = f["flags_scp"], f["potpara_scp"]
    pts = np.stack(np.meshgrid(g.x[0], g.x[1], g.x[2], indexing="ij"), axis=-1)
    pts = pts.transpose(2, 1, 0, 3).reshape(-1, 3)
    c2 = oracle_mod.Cluster(f["xyz"])
    c2.sites = sites
    Vo, _ = c2.potential_points_pol3(pts, flags, pp, pol=(alpha, ip, epc, e_pol))
    with pisces.Context(n, L) as ctx:
        ctx.build_potential_scp(flags, pp, sites, cl.q, cl.iq, alpha, ip, epc, e_pol)
        V = ctx.get_potential()
    assert np.all(np.isfinite(V))
    assert np.abs(V - Vo).max() <= 1e-12 * np.abs(Vo).max()


def test_scp_argument_errors(pisces, oracle_mod):
    f = golden("potential_pol3_c1.npz")
    cl = oracle_mod.Cluster(f["xyz"])
    alpha, ip, epc, e_pol = cl.pol3_lists()
    pp = oracle_mod.default_potpara()
    with pisces.Context((16, 16, 16), (36.0, 36.0, 36.0)) as ctx:
        with pytest.raises(pisces.PiscesError):  # flags[3] must be 3
            ctx.build_potential_scp((3, 2, 0, 1, 0, 0), pp, cl.sites, cl.q, cl.iq, alpha, ip,
                                    epc, e_pol)
        with pytest.raises(pisces.PiscesError):  # PolType 3 without the site lists
            ctx.build_potential((3, 2, 0, 3, 0, 0), pp, cl.sites, cl.q, cl.iq, cl.d, cl.id)
        zero = np.array(alpha)
        zero[1] = 0.0
        with pytest.raises(pisces.PiscesError, match="positive"):
            ctx.build_potential_scp((3, 2, 0, 3, 0, 0), pp, cl.sites, cl.q, cl.iq, zero, ip,
                                    epc, e_pol)
        bad = np.array(ip)
        bad[0] = 10 ** 6
        with pytest.raises(ValueError):  # the binding's own check
            ctx.build_potential_scp((3, 2, 0, 3, 0, 0), pp, cl.sites, cl.q, cl.iq, alpha, bad,
                                    epc, e_pol)
        # the library's check of the same list, called past the binding
        f, pp_, s, q, iq, al, ip_, ep, ix, iqq, nion = pisces.scp_arrays(
            (3, 2, 0, 3, 0, 0), pp, cl.sites, cl.q, cl.iq, alpha, ip, epc)
        ip_[0] = 10 ** 6
        rc = pisces.lib().pisces_build_potential_scp(
            ctx.h, pisces._pi(f), pisces._p(pp_), s.size // 3, pisces._p(s), q.size,
            pisces._p(q), pisces._pi(iq), ip_.size, pisces._p(al), pisces._pi(ip_),
            pisces._p(ep), ctypes.c_double(e_pol), 0, pisces._p(ix), pisces._p(iqq),
            ctypes.c_double(2.9))
        assert rc == -1


def test_solve_on_poltype3_potential(pisces, oracle_mod):
    """FP64 eigensolve on the GPU-built PolType 3 potential against the
    oracle's Davidson on the oracle's potential (same grid, same start)."""
    f = golden("potential_pol3_c1.npz")
    cl = oracle_mod.Cluster(f["xyz"])
    n, L = (24, 24, 24), (36.0, 36.0, 36.0)
    g = oracle_mod.Grid(n, L)
    flags = (3, 2, 0, 3, 0, 0)
    Vo = cl.potential_grid_pol3(g, flags)
    start = g.piab().reshape(1, -1)
    ro = g.solve(Vo, 1, 20, 100, 9, 1, start)
    alpha, ip, epc, e_pol = cl.pol3_lists()
    with pisces.Context(n, L) as ctx:
        ctx.build_potential_scp(flags, oracle_mod.default_potpara(), cl.sites, cl.q, cl.iq,
                                alpha, ip, epc, e_pol)
        r = ctx.solve(1, start, maxsub=20, maxmacro=100, ptol=9, corrflag=1)
    assert ro["nconv"] == 1 and r["nconv"] == 1
    assert abs(r["evals"][0] - ro["evals"][0]) <= 1e-8


@pytest.mark.parametrize("rb", ["1", "2", "4"])
def test_tile_sizes(pisces, oracle_mod, rb, monkeypatch):
    """16-, 32- and 64-point tiles (PISCES_POL3_RB) give the same potential."""
    monkeypatch.setenv("PISCES_POL3_RB", rb)
    f = golden("potential_pol3_c1.npz")
    cl = oracle_mod.Cluster(f["xyz"])
    n, L = (32, 32, 32), (36.0, 36.0, 36.0)
    g = oracle_mod.Grid(n, L)
    flags, pp = f["flags_scp"], f["potpara_scp"]
    Vo = cl.potential_grid_pol3(g, flags, pp)
    alpha, ip, epc, e_pol = cl.pol3_lists()
    with pisces.Context(n, L) as ctx:
        ctx.build_potential_scp(flags, pp, cl.sites, cl.q, cl.iq, alpha, ip, epc, e_pol)
        V = ctx.get_potential()
    assert np.abs(V - Vo).max() <= 1e-12 * np.abs(Vo).max()


@pytest.mark.parametrize("nt", ["256", "512", "1024"])
def test_thread_counts(pisces, oracle_mod, nt, monkeypatch):
    """Workgroups of 256, 512 and 1024 threads (PISCES_POL3_NT) on the 64^3
    (H2O)24 grid (14 column blocks: 1024 threads leave waves without one)."""
    monkeypatch.setenv("PISCES_POL3_NT", nt)
    f = golden("potential_pol3_c2.npz")
    cl = oracle_mod.Cluster(f["xyz"])
    n, L = (64, 64, 64), (40.0, 40.0, 40.0)
    g = oracle_mod.Grid(n, L)
    flags, pp = f["flags_scp"], f["potpara_scp"]
    alpha, ip, epc, e_pol = cl.pol3_lists()
    with pisces.Context(n, L) as ctx:
        ctx.build_potential_scp(flags, pp, cl.sites, cl.q, cl.iq, alpha, ip, epc, e_pol)
        V = ctx.get_potential()
    rng = np.random.default_rng(5)
    idx = rng.choice(g.ngp, size=2048, replace=False)
    i, j, k = idx % 64, (idx // 64) % 64, idx // 4096
    pts = np.stack([g.x[0][i], g.x[1][j], g.x[2][k]], axis=1)
    Vo, _ = cl.potential_points_pol3(pts, flags, pp)
    assert np.abs(V[idx] - Vo).max() <= 1e-12 * np.abs(Vo).max()


@pytest.mark.parametrize("nw, radius, seed", [(45, 8.0, 3000), (64, 9.0, 21)])
def test_large_polarizable_sets(pisces, oracle_mod, nw, radius, seed):
    """(H2O)45 (405 polarizable dimensions, the largest cluster a served
    configuration names) and (H2O)64 (576, the library's limit) on a 24^3 grid:
    the default 1024-thread workgroups and dynamic LDS above 64 KB, against the
    oracle at 1024 sampled points."""
    import workload
    xyz, _ = workload.water_cluster(nw, radius, seed)
    cl = oracle_mod.Cluster(xyz)
    n, L = (24, 24, 24), (48.0, 48.0, 48.0)
    g = oracle_mod.Grid(n, L)
    f = golden("potential_pol3_c1.npz")
    flags, pp = f["flags_scp"], f["potpara_scp"]
    alpha, ip, epc, e_pol = cl.pol3_lists()
    assert 3 * len(ip) == 9 * nw
    with pisces.Context(n, L) as ctx:
        ctx.build_potential_scp(flags, pp, cl.sites, cl.q, cl.iq, alpha, ip, epc, e_pol)
        V = ctx.get_potential()
    rng = np.random.default_rng(seed)
    idx = rng.choice(g.ngp, size=1024, replace=False)
    i, j, k = idx % 24, (idx // 24) % 24, idx // 576
    pts = np.stack([g.x[0][i], g.x[1][j], g.x[2][k]], axis=1)
    Vo, _ = cl.potential_points_pol3(pts, flags, pp)
    assert np.abs(V[idx] - Vo).max() <= 1e-12 * np.abs(Vo).max()


def test_too_many_polarizable_sites(pisces, oracle_mod):
    """3 np above 576 is refused before any launch."""
    import workload
    xyz, _ = workload.water_cluster(65, 9.5, 22)
    cl = oracle_mod.Cluster(xyz)
    f = golden("potential_pol3_c1.npz")
    alpha, ip, epc, e_pol = cl.pol3_lists()
    with pisces.Context((16, 16, 16), (48.0, 48.0, 48.0)) as ctx:
        with pytest.raises(pisces.PiscesError, match="polarizable"):
            ctx.build_potential_scp(f["flags_scp"], f["potpara_scp"], cl.sites, cl.q, cl.iq,
                                    alpha, ip, epc, e_pol)
