"""DVRType 3 (the FFT kinetic energy) on the GPU: the same matrix-core axis
contractions with the circulant Fourier kinetic matrices, against the oracle
(pinned to numpy's FFT form by tests/test_oracle_fft.py)."""
import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu


def test_grid_matrices(pisces, oracle_mod):
    n, L = (16, 20, 24), (20.0, 24.0, 28.0)
    g = oracle_mod.Grid(n, L, dvr_type=3)
    with pisces.Context(n, L, dvr_type=3) as ctx:
        xyz, T = ctx.grid()
    for a in range(3):
        np.testing.assert_allclose(xyz[a], g.x[a], rtol=0, atol=1e-13)
        np.testing.assert_allclose(T[a], g.T[a], rtol=0, atol=1e-13 * np.abs(g.T[a]).max())


@pytest.mark.parametrize("precision, tol", [("fp64", 1e-12), ("fp32", 2e-6), ("bf16x2", 2e-4)])
def test_apply_h(pisces, oracle_mod, precision, tol):
    n, L = (32, 24, 16), (36.0, 30.0, 24.0)
    g = oracle_mod.Grid(n, L, dvr_type=3)
    rng = np.random.default_rng(2)
    V = rng.uniform(-0.5, 0.5, g.ngp)
    X = rng.uniform(-1, 1, (4, g.ngp))
    Yo = g.apply_h(V, X)
    with pisces.Context(n, L, dvr_type=3) as ctx:
        ctx.set_potential(V)
        Y = ctx.apply_h(X) if precision == "fp64" else ctx.apply_h_ex(X, pisces.PRECISIONS[precision])
    assert np.abs(Y - Yo).max() <= tol * np.abs(Yo).max()


def _cluster(oracle_mod, n, L):
    f = golden("davidson_c1.npz")
    cl = oracle_mod.Cluster(f["xyz"])
    g = oracle_mod.Grid(n, L, dvr_type=3)
    return cl, g, cl.potential_grid(g)


def test_solve_matches_oracle(pisces, oracle_mod):
    """(H2O)6 anion with the FFT kinetic energy, 24^3 on 36 bohr: the FP64
    Davidson of the reference algorithm on the GPU against the oracle's."""
    n, L = (24, 24, 24), (36.0, 36.0, 36.0)
    cl, g, V = _cluster(oracle_mod, n, L)
    start = g.piab().reshape(1, -1)
    ro = g.solve(V, 1, 20, 200, 9, 1, start)
    with pisces.Context(n, L, dvr_type=3) as ctx:
        ctx.set_potential(V)
        r = ctx.solve(1, start, maxsub=20, maxmacro=200, ptol=9, corrflag=1)
    assert ro["nconv"] == 1 and r["nconv"] == 1
    assert abs(r["evals"][0] - ro["evals"][0]) <= 1e-8
    assert abs(r["vecs"][0] @ ro["vecs"][0]) >= 1 - 1e-8


@pytest.mark.parametrize("precision", ["fp64", "bf16x2"])
def test_block_engine(pisces, oracle_mod, precision):
    """The block engine (kinetic preconditioner through the circulant's
    eigenbasis) on the FFT kinetic energy: 3 states against the dense
    eigensolve of the oracle's H on a 12^3 grid."""
    n, L = (12, 12, 12), (24.0, 24.0, 24.0)
    cl, g, V = _cluster(oracle_mod, n, L)
    H = g.apply_h(V, np.eye(g.ngp))
    w = np.linalg.eigvalsh(0.5 * (H + H.T))[:3]
    rng = np.random.default_rng(4)
    start = rng.uniform(-1, 1, (3, g.ngp))
    with pisces.Context(n, L, dvr_type=3) as ctx:
        ctx.set_potential(V)
        r = ctx.solve_block(3, start, block=3, maxsub=15, max_iter=3000, ptol=9,
                            precision=pisces.PRECISIONS[precision])
    assert r["nconv"] == 3
    tol = 1e-8 if precision == "fp64" else 1e-6
    np.testing.assert_allclose(r["evals"], w, rtol=0, atol=tol)


def test_odd_sizes_are_refused(pisces):
    with pytest.raises(pisces.PiscesError):
        pisces.Context((15, 16, 16), (20.0, 20.0, 20.0), dvr_type=3)
    with pytest.raises(pisces.PiscesError):
        pisces.Context((16, 16, 16), (20.0, 20.0, 20.0), dvr_type=2)
