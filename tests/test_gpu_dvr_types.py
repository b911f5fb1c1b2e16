"""The other kinetic DVRs of DVR::SetupDVR on the GPU: the harmonic-oscillator
DVR (DVRType 1, src/ho_dvr.cpp) and the Colbert-Miller DVR (20, src/cm_dvr.cpp)
against the reference's own matrices (fixture tkin_types.npz, written through
the reference sources by tests/golden/make_golden.py) and closed-form 3-D
spectra."""
import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu


def test_matrices_match_reference(pisces):
    f = golden("tkin_types.npz")
    for i in range(int(f["ncases"])):
        t, n, p = f["case%d" % i]
        t, n = int(t), int(n)
        with pisces.Context((n, 8, 8), (p, p, p), dvr_type=t) as ctx:
            xs, Ts = ctx.grid()
        Tr = f["T%d" % i]
        np.testing.assert_allclose(xs[0], f["x%d" % i], rtol=0, atol=1e-12 * np.abs(xs[0]).max())
        np.testing.assert_allclose(Ts[0], Tr, rtol=0, atol=1e-12 * np.abs(Tr).max())


def test_harmonic_oscillator_spectrum(pisces):
    """H = T + sum_a w_a^2 x_a^2 / 2 on oscillator DVRs of the same frequencies
    (one per axis, so that the levels are distinct): E = sum_a w_a (n_a + 1/2).
    The reference Davidson algorithm (FP64, diagonal preconditioner)."""
    w, n = np.array([0.30, 0.37, 0.45]), 16
    with pisces.Context((n, n, n), tuple(w), dvr_type=1) as ctx:
        xs, _ = ctx.grid()
        X, Y, Z = np.meshgrid(xs[0], xs[1], xs[2], indexing="ij")
        V = 0.5 * (w[0] ** 2 * X ** 2 + w[1] ** 2 * Y ** 2 + w[2] ** 2 * Z ** 2)
        ctx.set_potential(V.transpose(2, 1, 0).reshape(-1))
        rng = np.random.default_rng(7)
        r = ctx.solve(4, rng.uniform(-1, 1, (4, n ** 3)), maxsub=24, maxmacro=200, ptol=9)
    assert r["nconv"] == 4
    e0 = 0.5 * w.sum()
    np.testing.assert_allclose(r["evals"], e0 + np.array([0.0, w[0], w[1], w[2]]), rtol=0,
                               atol=1e-9)


def _dense_h(Ts, V):
    nx, ny, nz = Ts[0].shape[0], Ts[1].shape[0], Ts[2].shape[0]
    Ix, Iy, Iz = np.eye(nx), np.eye(ny), np.eye(nz)
    # x fastest: index i + nx (j + ny k) -> kron(z, y, x)
    H = (np.kron(Iz, np.kron(Iy, Ts[0])) + np.kron(Iz, np.kron(Ts[1], Ix))
         + np.kron(Ts[2], np.kron(Iy, Ix)))
    return H + np.diag(V)


@pytest.mark.parametrize("precision", ["fp64", "bf16x2"])
def test_colbert_miller_well(pisces, precision):
    """A Gaussian well in a box of sides 20, 23, 27 bohr on the Colbert-Miller
    DVR: the 3 lowest states against the dense eigensolve of the same H (the
    context's matrices, which equal the reference's).  FP64: the reference
    Davidson algorithm; BF16X2: the block engine."""
    n, L = 12, (20.0, 23.0, 27.0)
    with pisces.Context((n, n, n), L, dvr_type=20) as ctx:
        xs, Ts = ctx.grid()
        X, Y, Z = np.meshgrid(xs[0], xs[1], xs[2], indexing="ij")
        r2 = (X - xs[0].mean()) ** 2 + (Y - xs[1].mean()) ** 2 + (Z - xs[2].mean()) ** 2
        V = (-0.5 * np.exp(-r2 / 16.0)).transpose(2, 1, 0).reshape(-1)
        ctx.set_potential(V)
        rng = np.random.default_rng(8)
        start = rng.uniform(-1, 1, (3, n ** 3))
        if precision == "fp64":
            r = ctx.solve(3, start, maxsub=20, maxmacro=300, ptol=9)
        else:
            # the mixed forms converge residuals to the workload's 1e-7; their
            # floor on this well lies near 1e-8, below the FP64 form's 1e-9
            r = ctx.solve_block(3, start, block=3, maxsub=15, ptol=7, precision=pisces.BF16X2)
    w = np.linalg.eigvalsh(_dense_h(Ts, V))[:3]
    assert r["nconv"] == 3
    np.testing.assert_allclose(r["evals"], w, rtol=0, atol=1e-8 if precision == "fp64" else 1e-6)
