"""DVRType 3 (the FFT kinetic energy, src/KE_diag.cpp + src/VectorFFTW.cpp) in
the oracle: the circulant grid matrix against closed forms and against numpy's
FFT (the reference's own FFT path needs FFTW, absent here)."""
import numpy as np
import pytest


def ek_numpy(n, L):
    dx = L / (n + 1)
    k = 2.0 * np.pi * np.fft.fftfreq(n, d=dx)
    return 0.5 * k * k


@pytest.mark.parametrize("n, L", [(8, 10.0), (16, 20.0), (32, 36.0), (64, 40.0)])
def test_spectrum_is_the_fourier_energies(oracle_mod, n, L):
    x, T, ek = oracle_mod.fft_kinetic(n, L)
    np.testing.assert_allclose(ek, ek_numpy(n, L), rtol=1e-14, atol=0)
    np.testing.assert_allclose(np.linalg.eigvalsh(T), np.sort(ek), rtol=0,
                               atol=1e-12 * ek.max())
    np.testing.assert_array_equal(T, T.T)
    dx = L / (n + 1)
    np.testing.assert_allclose(x, -0.5 * L + dx * np.arange(1, n + 1), rtol=0, atol=1e-14)


@pytest.mark.parametrize("m", [0, 1, 3, 7])
def test_plane_wave_is_an_eigenvector(oracle_mod, m):
    n, L = 16, 20.0
    x, T, ek = oracle_mod.fft_kinetic(n, L)
    dx = L / (n + 1)
    k = 2.0 * np.pi * m / (n * dx)
    c = np.cos(k * x)
    np.testing.assert_allclose(T @ c, 0.5 * k * k * c, rtol=0, atol=1e-12 * ek.max())


def test_apply_h_is_the_fft_form(oracle_mod):
    n, L = (6, 8, 10), (9.0, 11.0, 13.0)
    g = oracle_mod.Grid(n, L, dvr_type=3)
    rng = np.random.default_rng(1)
    V = rng.uniform(-1, 1, g.ngp)
    X = rng.uniform(-1, 1, (2, g.ngp))
    Y = g.apply_h(V, X)
    e = [ek_numpy(n[a], L[a]) for a in range(3)]
    KE = e[2][:, None, None] + e[1][None, :, None] + e[0][None, None, :]  # (z, y, x)
    for v in range(2):
        Xg = X[v].reshape(n[2], n[1], n[0])
        Yf = np.fft.ifftn(KE * np.fft.fftn(Xg)).real.reshape(-1) + V * X[v]
        np.testing.assert_allclose(Y[v], Yf, rtol=0, atol=1e-12 * np.abs(Yf).max())
