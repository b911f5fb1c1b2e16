"""The dual-grid method on the GPU (pisces_dual_grid_*): the coarse -> fine
transfer against the oracle's restatement of ExtendWfDouble / ExtendWfThree and
InterpolVdouble / InterpolVtriple (src/DVR.cpp:1240-1722), the potential of
Polarization 6 (src/DVR.cpp:776-830) against the oracle's exact potential and
transfer with the reference's distance test, and the fine solve started from
the carried-over coarse ground state (SVFlag 3) against the same solve from the
particle-in-a-box start."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [33, 31, 32, 48, 64, 96])
@pytest.mark.parametrize("kind", [0, 1])
def test_transfer_matches_oracle(pisces, oracle_mod, n, kind):
    print("dual transfer", n, kind, flush=True)
    L = 0.6 * (n + 1)
    nc, _ = pisces.dual_coarse((n,) * 3, (L,) * 3)
    assert nc[0] == oracle_mod.dual_coarse(n, L)[0]
    coarse = np.random.default_rng(n + kind).uniform(-1, 1, nc[0] ** 3)
    with pisces.Context((n,) * 3, (L,) * 3) as ctx:
        fine = ctx.dual_extend(nc, coarse, kind)
    ref = oracle_mod.dual_extend(n, nc[0], kind, coarse)
    assert np.abs(fine - ref).max() <= 4e-16 * np.abs(ref).max() * 4


@pytest.mark.parametrize("n", [33, 32])
def test_poltype6_potential(pisces, oracle_mod, n):
    """Exact PolType 3 potential near the sites, the coarse one carried over
    past rtol (a small rtol so that both kinds of points occur)."""
    import workload
    xyz, _ = workload.water_cluster(6, 3.5, 1000)
    L = 36.0
    flags = (3, 2, 0, 3, 0, 0)
    nc, Lc = pisces.dual_coarse((n,) * 3, (L,) * 3)
    rtol = 6.0
    with pisces.Context(nc, Lc) as cc:
        cc.build_cluster_potential(xyz, flags=flags)
        Vc = cc.get_potential()
    with pisces.Context((n,) * 3, (L,) * 3) as ctx:
        sites, _, _ = ctx.build_cluster_potential(xyz, flags=flags)
        Vexact = ctx.get_potential()
        ctx.dual_potential(nc, Vc, rtol)
        V6 = ctx.get_potential()
    g = oracle_mod.Grid((n,) * 3, (L,) * 3)
    X, Y, Z = np.meshgrid(g.x[0], g.x[1], g.x[2], indexing="ij")
    pts = np.stack([A.transpose(2, 1, 0).reshape(-1) for A in (X, Y, Z)], axis=1)
    mind = np.minimum(50.0, np.min(np.linalg.norm(pts[:, None, :] - sites[None, :, :], axis=2),
                                   axis=1))
    carried = oracle_mod.dual_extend(n, nc[0], 1, Vc)
    idx = np.arange(n)
    if n % 2:
        on = (idx % 2 == 0)
        inside = np.ones(n, bool)
    else:
        check = n - (nc[0] * 3 // 2 - 1) * 2
        q = idx - check // 2
        inside = (q >= 0) & (q < n - check)
        on = inside & (q % 3 == 0)
    onl = (on[None, None, :] & on[None, :, None] & on[:, None, None]).reshape(-1)
    ins = (inside[None, None, :] & inside[None, :, None] & inside[:, None, None]).reshape(-1)
    use = (~ins) | onl | (mind > rtol)
    expect = np.where(use, carried, Vexact)
    assert 0 < use.sum() < use.size
    assert np.abs(V6 - expect).max() <= 1e-12 * np.abs(expect).max()


def test_fine_solve_from_the_coarse_ground_state(pisces):
    """SVFlag 3: the fine solve started from the coarse ground state carried
    over converges to the eigenvalue of the particle-in-a-box start, in fewer
    matvecs."""
    import workload
    xyz, _ = workload.water_cluster(6, 3.5, 1000)
    n, L = 33, 36.0
    flags = (3, 2, 0, 3, 0, 0)
    nc, Lc = pisces.dual_coarse((n,) * 3, (L,) * 3)
    with pisces.Context(nc, Lc) as cc:
        cc.build_cluster_potential(xyz, flags=flags)
        rc_ = cc.solve(1, pisces.start_vectors(cc, 1), maxsub=20, maxmacro=200, ptol=8, corrflag=1)
        Vc = cc.get_potential()
    assert rc_["nconv"] == 1
    with pisces.Context((n,) * 3, (L,) * 3) as ctx:
        ctx.build_cluster_potential(xyz, flags=flags)
        ctx.dual_potential(nc, Vc, 20.0)
        cold = ctx.solve(1, pisces.start_vectors(ctx, 1), maxsub=20, maxmacro=200, ptol=8,
                         corrflag=1)
        start = ctx.dual_extend(nc, rc_["vecs"][0], 0).reshape(1, -1)
        warm = ctx.solve(1, start, maxsub=20, maxmacro=200, ptol=8, corrflag=1)
    print("dual: coarse %.10f fine cold %.10f (%d matvecs) warm %.10f (%d matvecs)"
          % (rc_["evals"][0], cold["evals"][0], cold["n_matvec"], warm["evals"][0],
             warm["n_matvec"]), flush=True)
    assert cold["nconv"] == 1 and warm["nconv"] == 1
    assert abs(warm["evals"][0] - cold["evals"][0]) <= 1e-8
    assert warm["n_matvec"] < cold["n_matvec"]
