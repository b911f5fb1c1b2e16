"""FP64 Davidson on the GPU against the reference's own Davidson (golden
fixtures written by running src/Davidson.cpp) and against the oracle."""
import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("tag", ["g10_r1", "g12_r4", "g10_r2_jd", "g10_r3_small", "c1"])
def test_fp64_solve_matches_reference(pisces, oracle_mod, tag):
    print("solve", tag, flush=True)
    f = golden("davidson_%s.npz" % tag)
    n, L = tuple(int(v) for v in f["n"]), tuple(float(v) for v in f["L"])
    cl = oracle_mod.Cluster(f["xyz"])
    g = oracle_mod.Grid(n, L)
    V = cl.potential_grid(g)
    nroots, maxsub, maxmacro, ptol, corrflag = (int(v) for v in f["params"])
    with pisces.Context(n, L) as ctx:
        ctx.set_potential(V)
        r = ctx.solve(nroots, f["start"], maxsub, maxmacro, ptol, corrflag)
    assert r["nconv"] == int(f["nconv"])
    if r["nconv"] == nroots:
        # eigenvalues within 1e-8 Eh, eigenvectors with overlap >= 1 - 1e-8
        np.testing.assert_allclose(r["evals"], f["evals"], rtol=0, atol=1e-8)
        for i in range(nroots):
            assert abs(r["vecs"][i] @ f["vecs"][i]) >= 1 - 1e-8
        assert np.all(r["resid"] < 10.0 ** -ptol)
        # matvec counts stay close to the reference's (rounding-level changes
        # in the orthogonalisation can shift a restart by an iteration)
        assert abs(r["n_matvec"] - int(f["n_mtx"])) <= 0.1 * int(f["n_mtx"]) + 8


def test_fp64_solve_c2_reproduces_reference(pisces):
    """The benchmark workload ((H2O)24, 64^3, 4 states, maxsub 32, pTol 7) with
    the reference's one-vector-per-root Davidson: the reference itself converges
    only 2 of the 4 roots in 200 macro iterations (5531 H*x products; the 2nd
    and 3rd states are 1.8e-4 Eh apart).  The GPU form reproduces that outcome
    and the two converged eigenpairs."""
    import workload
    f = golden("davidson_c2_meta.npz")
    xyz, _ = workload.water_cluster(24, 6.0, 2000)
    n, L = (64, 64, 64), (40.0, 40.0, 40.0)
    with pisces.Context(n, L) as ctx:
        ctx.build_cluster_potential(xyz)
        start = pisces.start_vectors(ctx, 4, seed=1)
        r = ctx.solve(4, start, maxsub=32, maxmacro=200, ptol=7, corrflag=1)
        U = r["vecs"]
        HU = ctx.apply_h(U)
    assert r["nconv"] == int(f["nconv"])
    assert abs(r["n_matvec"] - int(f["n_mtx"])) <= 0.1 * int(f["n_mtx"])
    rq = np.einsum("ij,ij->i", U, HU) / np.einsum("ij,ij->i", U, U)
    nc = int(f["nconv"])
    np.testing.assert_allclose(rq[:nc], f["rq"][:nc], rtol=0, atol=1e-8)
    res = np.linalg.norm(HU - rq[:, None] * U, axis=1)
    assert np.all(res[:nc] < 1e-7)


@pytest.mark.parametrize("tag", ["g10_r1", "g12_r4", "g10_r2_jd"])
def test_reverse_communication_matches_reference(pisces, oracle_mod, tag):
    """pisces_davidson driven like DVR::davdriver, with the caller's own matvec
    (the oracle's CPU H*x): the reference's eigenvalues and matvec count."""
    print("rc", tag, flush=True)
    f = golden("davidson_%s.npz" % tag)
    n, L = tuple(int(v) for v in f["n"]), tuple(float(v) for v in f["L"])
    g = oracle_mod.Grid(n, L)
    V = oracle_mod.Cluster(f["xyz"]).potential_grid(g)
    nroots, maxsub, maxmacro, ptol, corrflag = (int(v) for v in f["params"])
    with pisces.Context(n, L) as ctx:
        nconv, ev, vecs, n_mtx = ctx.davidson_rc(
            f["start"], g.diag(V), lambda X: g.apply_h(V, X), maxsub=maxsub, maxmacro=maxmacro,
            ptol=ptol, corrflag=corrflag)
    assert nconv == int(f["nconv"]) == nroots
    assert n_mtx == int(f["n_mtx"])
    np.testing.assert_allclose(ev, f["evals"], rtol=0, atol=1e-8)
    for i in range(nroots):
        assert abs(vecs[i] @ f["vecs"][i]) >= 1 - 1e-8
