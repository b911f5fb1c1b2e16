"""The oracle's dual-grid transfer (coarse -> fine, src/DVR.cpp:1240-1722 with
the coarse grid of src/ClusterAnion.cpp:209-232), pinned by what the geometry
and the interpolation fix rather than by a reference run (DVR.cpp needs FFTW
headers that this image lacks):
* the coarse grid's points coincide with every second (odd n) or third (even n)
  fine point, shifted by the reference's checkM / 2 for triple spacing;
* double spacing averages the 2, 4 or 8 coarse corners: it reproduces every
  multilinear function exactly;
* triple spacing weighs the corners by inverse distance, normalised (the
  reference's constants carry six digits); ExtendWfThree differs from that in
  three weights, at src/DVR.cpp:1570 and :1615-1620, and nowhere else;
* points outside the coarse grid's hull are 0."""
import itertools

import numpy as np
import pytest


@pytest.mark.parametrize("n", [15, 31, 33, 95, 32, 48, 64, 96])
def test_coarse_points_coincide_with_fine_points(oracle_mod, n):
    L = 0.6154 * (n + 1)
    nc, Lc = oracle_mod.dual_coarse(n, L)
    xf, _ = oracle_mod.sine_dvr(n, L)
    xc, _ = oracle_mod.sine_dvr(nc, Lc)
    h = L / (n + 1)
    if n % 2:
        assert nc == (n + 1) // 2
        np.testing.assert_allclose(xc, xf[0::2], rtol=0, atol=1e-12 * L)
    else:
        check = n - (nc * 3 // 2 - 1) * 2
        assert check in (0, 2, 4)
        sh = check // 2
        np.testing.assert_allclose(xc, xf[sh:sh + 3 * nc - 2:3], rtol=0, atol=1e-12 * L)
        np.testing.assert_allclose(np.diff(xc), 3 * h, rtol=1e-12)


@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("n", [9, 17])
def test_double_spacing_reproduces_multilinear_functions(oracle_mod, n, kind):
    nc = (n + 1) // 2
    rng = np.random.default_rng(n)
    a = rng.normal(size=(3, 2))
    c = rng.normal(size=8)
    def f(i, j, k):  # multilinear in the index coordinates (the grids are uniform)
        xs = [(a[0, 0] + a[0, 1] * i), (a[1, 0] + a[1, 1] * j), (a[2, 0] + a[2, 1] * k)]
        return (c[0] + c[1] * xs[0] + c[2] * xs[1] + c[3] * xs[2] + c[4] * xs[0] * xs[1] +
                c[5] * xs[1] * xs[2] + c[6] * xs[0] * xs[2] + c[7] * xs[0] * xs[1] * xs[2])
    K, J, I = np.meshgrid(np.arange(nc), np.arange(nc), np.arange(nc), indexing="ij")
    coarse = f(2 * I, 2 * J, 2 * K).reshape(-1)
    fine = oracle_mod.dual_extend(n, nc, kind, coarse)
    K, J, I = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    exact = f(I, J, K).reshape(-1)
    assert np.abs(fine - exact).max() <= 1e-13 * np.abs(exact).max()


# the reference's departures from normalised inverse-distance weights, as
# (kind, residues, corner index a + 2 b + 4 c)
REF_DEVIATIONS = {(0, (2, 0, 1), 4),        # src/DVR.cpp:1570: w1 where 1/sqrt(8) belongs
                  (0, (2, 1, 2), 1), (0, (2, 1, 2), 3)}  # :1615-1620: 1/sqrt(9) and 1/sqrt(6) swapped


@pytest.mark.parametrize("kind", [0, 1])
def test_triple_spacing_weights_are_inverse_distances(oracle_mod, kind):
    for r in itertools.product(range(3), repeat=3):
        w = oracle_mod.dual_weights(kind, *r)
        if r == (0, 0, 0):
            np.testing.assert_array_equal(w, np.eye(8)[0])
            continue
        moving = [d for d in range(3) if r[d]]
        ideal = np.zeros(8)
        for cidx in range(8):
            corner = [(cidx >> d) & 1 for d in range(3)]
            if any(corner[d] and not r[d] for d in range(3)):
                continue
            dist = np.sqrt(sum(((3 - r[d]) if corner[d] else r[d]) ** 2 for d in moving))
            ideal[cidx] = 1.0 / dist if len(moving) > 1 else (3 - dist)
        ideal /= ideal.sum()
        for cidx in range(8):
            if (kind, r, cidx) in REF_DEVIATIONS:
                assert abs(w[cidx] - ideal[cidx]) > 0.01
            else:
                assert abs(w[cidx] - ideal[cidx]) <= 3e-6, (kind, r, cidx, w, ideal)
        if not any((kind, r, c) in REF_DEVIATIONS for c in range(8)):
            # six-digit constants (1/sqrt(9) written 0.33333): sums within 4e-6 of 1
            assert abs(w.sum() - 1.0) <= 4e-6


@pytest.mark.parametrize("n", [32, 48, 64])
def test_triple_spacing_values(oracle_mod, n):
    nc, _ = oracle_mod.dual_coarse(n, 40.0)
    check = n - (nc * 3 // 2 - 1) * 2
    sh = check // 2
    rng = np.random.default_rng(n)
    coarse = rng.normal(size=nc ** 3)
    for kind in (0, 1):
        fine = oracle_mod.dual_extend(n, nc, kind, coarse).reshape(n, n, n)
        cz = coarse.reshape(nc, nc, nc)
        # exact on the coarse lattice, zero outside its hull
        np.testing.assert_array_equal(fine[sh:sh + 3 * nc - 2:3, sh:sh + 3 * nc - 2:3,
                                           sh:sh + 3 * nc - 2:3], cz)
        inside = np.zeros(n, bool)
        inside[sh:sh + 3 * nc - 2] = True
        outside = ~(inside[:, None, None] & inside[None, :, None] & inside[None, None, :])
        assert np.all(fine[outside] == 0.0)
        # a point between two coarse points along x: 2/3 and 1/3
        k0 = sh + 3
        np.testing.assert_allclose(fine[k0, k0, sh + 1], (2 * cz[1, 1, 0] + cz[1, 1, 1]) / 3,
                                   rtol=1e-15)
