"""What the mathematics fixes about the oracle, independent of the reference.

Each check is chosen so that a plausible slip in the restatement (a dropped
term, a wrong sign or index, a transposed operand) fails it.
"""
import numpy as np
import pytest


def dense_h(grid, V):
    """Dense Kronecker-sum Hamiltonian, x fastest (the construction of
    DVR::build_h, src/fulldiag.cpp:26-95)."""
    Tx, Ty, Tz = grid.T
    nx, ny, nz = grid.n
    Ix, Iy, Iz = np.eye(nx), np.eye(ny), np.eye(nz)
    # index g = i + nx*(j + ny*k) -> kron order (k, j, i)
    H = np.kron(Iz, np.kron(Iy, Tx)) + np.kron(Iz, np.kron(Ty, Ix)) + np.kron(Tz, np.kron(Iy, Ix))
    return H + np.diag(V)


def test_sine_dvr_spectrum_is_particle_in_a_box(oracle_mod):
    # eigenvalues of the sine-DVR T are exactly pi^2 n^2 / (2 L^2), L = (N+1) dx
    # (the reference uses pi to 15 digits, a relative error of 1e-15)
    for n, L in [(7, 5.0), (16, 20.0), (33, 36.0)]:
        x, T = oracle_mod.sine_dvr(n, L)
        w = np.linalg.eigvalsh(T)
        exact = np.pi ** 2 * np.arange(1, n + 1) ** 2 / (2 * L ** 2)
        np.testing.assert_allclose(w, exact, rtol=1e-12)
        np.testing.assert_allclose(T, T.T, rtol=0, atol=0)
        dx = L / (n + 1)
        np.testing.assert_allclose(x, -L / 2 + dx * np.arange(1, n + 1), rtol=0, atol=1e-14)


def test_sine_dvr_equals_fbr_transform(oracle_mod):
    n, L = 12, 9.0
    _, T = oracle_mod.sine_dvr(n, L)
    j = np.arange(1, n + 1)
    S = np.sqrt(2.0 / (n + 1)) * np.sin(np.pi * np.outer(j, j) / (n + 1))
    K = np.diag(np.pi ** 2 * j ** 2 / (2 * L ** 2))
    np.testing.assert_allclose(T, S @ K @ S.T, rtol=0, atol=1e-12 * np.abs(T).max())


def test_apply_h_equals_dense_kronecker(oracle_mod):
    g = oracle_mod.Grid((5, 6, 7), (6.0, 7.0, 8.0))
    rng = np.random.default_rng(3)
    V = rng.normal(size=g.ngp)
    X = rng.normal(size=(3, g.ngp))
    H = dense_h(g, V)
    Y = g.apply_h(V, X)
    np.testing.assert_allclose(Y, X @ H.T, rtol=0, atol=1e-12 * np.abs(H).max())
    np.testing.assert_allclose(g.diag(V), np.diag(H), rtol=0, atol=1e-14)


def test_davidson_equals_dense_eigensolve(oracle_mod):
    import workload
    xyz, _ = workload.water_cluster(6, 3.5, 1000)
    cl = oracle_mod.Cluster(xyz)
    g = oracle_mod.Grid((8, 8, 8), (16.0, 16.0, 16.0))
    V = cl.potential_grid(g)
    w = np.linalg.eigvalsh(dense_h(g, V))
    rng = np.random.default_rng(5)
    start = np.vstack([g.piab()] + [rng.uniform(-1, 1, g.ngp) for _ in range(2)])
    r = g.solve(V, 3, 12, 200, 10, 1, start)
    assert r["nconv"] == 3
    np.testing.assert_allclose(r["evals"], w[:3], rtol=0, atol=1e-12)


def test_water_projection_roundtrip(oracle_mod):
    import workload
    xyz, _ = workload.water_cluster(8, 4.5, 42)
    sites, conf = oracle_mod.water_sites(xyz)
    s = sites.reshape(-1, 4, 3)
    b = xyz.reshape(-1, 3, 3) * workload.ANGS2BOHR
    # rigid TTM geometry: input reproduced to the 1e-9 bohr precision of the
    # reference's tabulated monomer
    assert np.abs(s[:, :3] - b).max() < 5e-9
    # M site: 0.25 A from O along the bisector towards the H atoms
    bis = (s[:, 1] + s[:, 2]) / 2 - s[:, 0]
    bis /= np.linalg.norm(bis, axis=1)[:, None]
    np.testing.assert_allclose(s[:, 3], s[:, 0] + 0.25 * workload.ANGS2BOHR * bis, atol=1e-12)
    # centre of mass with masses 16, 1, 1
    com = (16 * b[:, 0] + b[:, 1] + b[:, 2]) / 18
    np.testing.assert_allclose(conf[:, :3], com, atol=1e-13)


def test_induced_dipoles_solve_linear_system(oracle_mod):
    """The SCF fixed point satisfies mu = alpha (E_pc + T mu) (a dense solve)."""
    import workload
    xyz, _ = workload.water_cluster(5, 3.5, 7)
    sites, _ = oracle_mod.water_sites(xyz)
    ind = oracle_mod.water_induced(sites)
    assert ind["converged"]
    a3 = workload.ANGS2BOHR ** 3
    alpha = np.tile([1.22 * a3, 0.28 * a3, 0.28 * a3], 5)
    pos = sites.reshape(-1, 4, 3)[:, :3].reshape(-1, 3)
    ns = pos.shape[0]

    def damp(r, ai, aj, a, dd):
        A = (ai * aj) ** (1 / 6)
        x = a * (r / A) ** 3
        return 1 - (1 + x) * np.exp(-x) if dd else 1 - np.exp(-x)

    M = np.zeros((3 * ns, 3 * ns))
    for i in range(ns):
        for j in range(ns):
            if i == j:
                continue
            R = pos[i] - pos[j]
            r = np.linalg.norm(R)
            d1, d2 = damp(r, alpha[i], alpha[j], 0.3, 1), damp(r, alpha[i], alpha[j], 0.3, 0)
            M[3 * i:3 * i + 3, 3 * j:3 * j + 3] = (3 * d1 * np.outer(R, R) - d2 * r * r * np.eye(3)) / r ** 5
    A = np.eye(3 * ns) - np.repeat(alpha, 3)[:, None] * M
    mu = np.linalg.solve(A, np.repeat(alpha, 3) * ind["epc"])
    np.testing.assert_allclose(ind["mu"], mu, rtol=0, atol=1e-9 * np.abs(mu).max())


def test_potential_invariant_under_rigid_motion(oracle_mod):
    import workload
    xyz, _ = workload.water_cluster(4, 3.0, 11)
    rng = np.random.default_rng(1)
    pts = rng.uniform(-6, 6, size=(300, 3))
    v0, _ = oracle_mod.Cluster(xyz).potential_points(pts)
    th = 0.7
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    t = np.array([0.3, -0.2, 0.5])
    xyz2 = (xyz.reshape(-1, 3) @ R.T + t).reshape(-1, 9)
    pts2 = pts @ R.T + t * workload.ANGS2BOHR
    v1, _ = oracle_mod.Cluster(xyz2).potential_points(pts2)
    np.testing.assert_allclose(v1, v0, rtol=0, atol=1e-9 * np.abs(v0).max())


def test_potential_far_field_is_point_dipoles(oracle_mod):
    """Far from a neutral water the charge term is the field of its point charges
    and the polarization term is -alpha/(2 R^4) per M site."""
    import workload
    xyz, _ = workload.water_cluster(1, 0.1, 3)
    cl = oracle_mod.Cluster(xyz)
    d = np.array([[60.0, 0, 0], [0, 70.0, 0], [0, 0, -80.0]])
    _, parts = cl.potential_points(d)
    s = cl.sites
    vpc = np.array([-sum(q / np.linalg.norm(p - s[i]) for q, i in zip(cl.q, cl.iq)) for p in d])
    np.testing.assert_allclose(parts[:, 0], vpc, rtol=1e-13)
    m = s[3]
    np.testing.assert_allclose(parts[:, 3], [-0.5 * 9.745 / np.linalg.norm(p - m) ** 4 for p in d],
                               rtol=1e-13)
    assert np.all(parts[:, 1] == 0.0)  # a lone water has no induced dipoles


def test_effective_distance_damping_is_c1_continuous(oracle_mod):
    """Reff(R; r0) and its first derivative are continuous at R = r0."""
    xyz = np.array([[0.0, 0.0, 0.0, 0.0, 0.7569503, -0.5858823, 0.0, -0.7569503, -0.5858823]])
    cl = oracle_mod.Cluster(xyz)
    m = cl.sites[3]
    r0 = 2.536
    eps = 1e-6
    dirn = np.array([0.0, 0.0, 1.0])
    pts = np.array([m + (r0 + k * eps) * dirn for k in (-2, -1, 0, 1, 2)])
    flags = (4, 2, 0, 1, 0, 0)
    _, parts = cl.potential_points(pts, flags)
    vp = parts[:, 3]
    d_left = (vp[2] - vp[0]) / (2 * eps)
    d_right = (vp[4] - vp[2]) / (2 * eps)
    assert abs(d_left - d_right) < 1e-4 * abs(d_left) + 1e-9


def test_diag_is_dense_diagonal(oracle_mod):
    g = oracle_mod.Grid((5, 4, 6), (9.0, 8.0, 10.0))
    V = np.random.default_rng(3).uniform(-1, 1, g.ngp)
    np.testing.assert_allclose(g.diag(V), np.diag(dense_h(g, V)), rtol=0, atol=1e-13)


def test_syev_matches_lapack(oracle_mod):
    rng = np.random.default_rng(9)
    for n in (1, 2, 7, 24):
        A = rng.normal(size=(n, n))
        A = A + A.T
        W, Vv = oracle_mod.syev(A)
        w = np.linalg.eigvalsh(A)
        np.testing.assert_allclose(W, w, rtol=0, atol=1e-12 * max(1.0, np.abs(w).max()))
        np.testing.assert_allclose(A @ Vv, Vv * W[None, :], rtol=0, atol=1e-11 * max(1.0, np.abs(w).max()))


def test_piab_start_is_the_central_half_box_ground_state(oracle_mod):
    """The start vector (DVR::ParticleInAnDBoxWf, src/DVR.cpp:1725-1805) on the
    symmetric sine-DVR grid: with Lg = x_last - x_first per axis it is the
    lowest particle-in-a-box state of the central half box |q| <= Lg / 4,
    prod_d cos(2 pi q_d / Lg_d) there and zero outside, x fastest.  A slip in
    the shift (the full box), the factor 2, the mask or the axis order fails
    this closed form."""
    for n, L in [((12, 11, 10), (30.0, 26.0, 22.0)), ((33, 32, 17), (36.0, 36.0, 20.0))]:
        g = oracle_mod.Grid(n, L)
        wf = g.piab()
        f = []
        for d in range(3):
            x = g.x[d]
            Lg = x[-1] - x[0]
            assert abs(x[0] + x[-1]) < 1e-13  # symmetric grid
            f.append(np.where(np.abs(x) <= Lg / 4 + 1e-12, np.cos(2 * np.pi * x / Lg), 0.0))
        expect = np.einsum("k,j,i->kji", f[2], f[1], f[0]).ravel()
        np.testing.assert_allclose(wf, expect, rtol=0, atol=1e-14)
        assert (wf >= -1e-15).all() and np.count_nonzero(np.abs(wf) > 1e-12) > 0
    # a grid with a centre point on every axis peaks at exactly 1 there
    g = oracle_mod.Grid((33, 33, 33), (36.0, 36.0, 36.0))
    wf = g.piab()
    assert abs(wf[16 + 33 * (16 + 33 * 16)] - 1.0) < 1e-15
    assert wf.max() <= 1.0 + 1e-15
