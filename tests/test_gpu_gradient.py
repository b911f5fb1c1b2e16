"""The electron side of the analytic gradient on the GPU (pisces_gradient_terms,
pisces_compute_gradient) against the reference's own gradient code
(grad_*.npz: ClusterAnion::GetAnalGrad with DVR::ComputeGradient's grid loop
over Potential::EvaluateGradient, FinalGradient and SubtractWWGradient, and its
WaterCluster for the water side), against central differences of the GPU
potential itself, and against the closed-form gradient of one damped charge."""
import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu

FLAGS = (3, 2, 0, 3, 0, 0)


def _ctx(pisces, f):
    n, L = tuple(int(v) for v in f["n"]), tuple(float(v) for v in f["L"])
    ctx = pisces.Context(n, L)
    sites, mu, _ = ctx.build_cluster_potential(f["xyz"], flags=FLAGS)
    return ctx, sites


@pytest.mark.parametrize("tag", ["c1", "c2"])
def test_gradient_matches_reference(pisces, tag):
    print("gradient", tag, flush=True)
    f = golden("grad_%s.npz" % tag)
    ctx, sites = _ctx(pisces, f)
    with ctx:
        nsites, npol = sites.shape[0], f["tmu"].size // 3
        gel, tmu, mcm = ctx.gradient_terms(f["psi"], nsites, npol)
        # the water model's SCF stops at a mean change of 1e-12 on both sides
        assert np.abs(gel - f["gel"]).max() <= 1e-10 * np.abs(f["gel"]).max()
        assert np.abs(tmu - f["tmu"]).max() <= 1e-10 * np.abs(f["tmu"]).max()
        assert np.abs(mcm.reshape(-1) - f["mcm"]).max() <= 1e-10 * np.abs(f["mcm"]).max()
        g = ctx.compute_gradient(f["psi"], f["gww"], f["dTx"], f["dTy"], f["dTz"], f["polgrad"],
                                 f["dEfield"])
    assert np.abs(g - f["gfinal"]).max() <= 1e-10 * np.abs(f["gfinal"]).max()


def test_gradient_is_the_derivative_of_the_gpu_potential(pisces, oracle_mod):
    """Central differences of sum psi^2 V over the GPU's own PolType 3 grid
    potential, one site displaced at a time (the permanent-charge field and
    VpolWaterWater held, InvA rebuilt from the moved sites), against the GPU
    gradient terms plus the dipole-tensor term of the reference's BuildSuperdT."""
    f = golden("grad_c1.npz")
    n, L = tuple(int(v) for v in f["n"]), tuple(float(v) for v in f["L"])
    cl = oracle_mod.Cluster(f["xyz"])
    alpha, ip, epc, vww = cl.pol3_lists()
    pp = oracle_mod.default_potpara()
    psi = f["psi"]
    w = psi * psi
    npol = ip.size
    mcm_ref = f["mcm"].reshape(3 * npol, 3 * npol)
    dT = [f[k].reshape(3 * npol, 3 * npol) for k in ("dTx", "dTy", "dTz")]
    h = 1e-4
    with pisces.Context(n, L) as ctx:
        ctx.build_potential_scp(FLAGS, pp, cl.sites, cl.q, cl.iq, alpha, ip, epc, vww)
        gel, _, mcm = ctx.gradient_terms(psi, cl.sites.shape[0], npol)
        scale = np.abs(gel).max()
        for s in (0, 1, 2, 3, 5, 10, 15, 23):
            for a in range(3):
                F = []
                for sign in (1, -1):
                    moved = cl.sites.copy()
                    moved[s, a] += sign * h
                    ctx.build_potential_scp(FLAGS, pp, moved, cl.q, cl.iq, alpha, ip, epc, vww)
                    F.append(w @ ctx.get_potential())
                fd = (F[0] - F[1]) / (2 * h)
                an = gel[3 * s + a]
                if s % 4 != 3:
                    j = 3 * (s // 4) + s % 4
                    an -= np.sum(dT[a][3 * j:3 * j + 3, :] * mcm[3 * j:3 * j + 3, :])
                assert abs(fd - an) <= 1e-6 * scale, (s, a, fd, an)
    assert np.abs(mcm - mcm_ref).max() <= 1e-10 * np.abs(mcm_ref).max()


def test_gradient_needs_a_poltype3_potential(pisces):
    f = golden("grad_c1.npz")
    n, L = tuple(int(v) for v in f["n"]), tuple(float(v) for v in f["L"])
    with pisces.Context(n, L) as ctx:
        ctx.build_cluster_potential(f["xyz"])  # Polarization 1
        with pytest.raises(pisces.PiscesError, match="PolType 3"):
            ctx.gradient_terms(f["psi"], 24, 18)


@pytest.mark.parametrize("nw, radius, seed", [(45, 8.0, 3000), (64, 9.0, 21)])
def test_gradient_terms_large_polarizable_sets(pisces, oracle_mod, nw, radius, seed):
    """(H2O)45 (405 polarizable dimensions, 26 column blocks: 351 blocks of
    S in several groups of the matrix-core kernel) and (H2O)64 (576, the
    library's limit) on a 16^3 grid with a smooth normalised psi, against the
    oracle's grid loop."""
    import workload
    print("gradient large", nw, flush=True)
    xyz, _ = workload.water_cluster(nw, radius, seed)
    cl = oracle_mod.Cluster(xyz)
    n, L = (16, 16, 16), (48.0, 48.0, 48.0)
    g = oracle_mod.Grid(n, L)
    X, Y, Z = np.meshgrid(g.x[0], g.x[1], g.x[2], indexing="ij")
    psi = np.exp(-(X ** 2 + Y ** 2 + Z ** 2) / 60.0).transpose(2, 1, 0).reshape(-1)
    psi /= np.linalg.norm(psi)
    gel_o, tmu_o, mcm_o = oracle_mod.gradient_grid(cl, g, psi)
    alpha, ip, epc, vww = cl.pol3_lists()
    with pisces.Context(n, L) as ctx:
        ctx.build_potential_scp(FLAGS, oracle_mod.default_potpara(), cl.sites, cl.q, cl.iq, alpha,
                                ip, epc, vww)
        gel, tmu, mcm = ctx.gradient_terms(psi, cl.sites.shape[0], ip.size)
    assert np.abs(gel - gel_o).max() <= 1e-10 * np.abs(gel_o).max()
    assert np.abs(tmu - tmu_o).max() <= 1e-10 * np.abs(tmu_o).max()
    assert np.abs(mcm - mcm_o).max() <= 1e-10 * np.abs(mcm_o).max()


def test_single_charge_gradient_closed_form(pisces):
    """One charge q on the M site of a quartet whose other sites sit 800 bohr
    away (the polarizable O site there, tiny alpha): the charge site's entry of
    sum psi^2 dV/dR is the derivative of the damped Coulomb term -q / Reff(d),
    d = |r - R|, Reff(d) = r0 (1/2 + rho^3 (1 - rho/2)) for rho = d / r0 < 1 and
    d beyond r0 (src/Potential.cpp:1191-1196), so
        dV/dR = -q Reff'(d) (r - R) / (d Reff(d)^2),  Reff' = 3 rho^2 - 2 rho^3 inside,
    summed with the weights psi^2 in numpy."""
    n, L = (24, 24, 24), (30.0, 30.0, 30.0)
    pp = pisces.default_potpara()
    r0, q = abs(pp[3]), 0.7
    R = np.array([0.37, -0.52, 0.21])
    far = np.array([800.0, 800.0, 800.0])
    sites = np.array([far, far + [1.0, 0, 0], far + [0, 1.0, 0], R])
    flags = (3, 2, 0, 3, 0, 0)
    with pisces.Context(n, L) as ctx:
        xs, _ = ctx.grid()
        X, Y, Z = np.meshgrid(xs[0], xs[1], xs[2], indexing="ij")
        psi = np.exp(-((X - 1.0) ** 2 + Y ** 2 + (Z + 0.5) ** 2) / 8.0)
        psi = (psi / np.linalg.norm(psi)).transpose(2, 1, 0).reshape(-1)
        ctx.build_potential_scp(flags, pp, sites, [q], [3], [1e-3], [0], np.zeros(3), 0.0)
        gel, _, _ = ctx.gradient_terms(psi, 4, 1)
    pts = np.stack([X.transpose(2, 1, 0).reshape(-1), Y.transpose(2, 1, 0).reshape(-1),
                    Z.transpose(2, 1, 0).reshape(-1)], axis=1)
    dr = pts - R
    d = np.linalg.norm(dr, axis=1)
    rho = d / r0
    inside = d < r0
    reff = np.where(inside, r0 * (0.5 + rho ** 3 * (1.0 - 0.5 * rho)), d)
    dreff = np.where(inside, 3.0 * rho ** 2 - 2.0 * rho ** 3, 1.0)
    dv = -q * (dreff / (d * reff ** 2))[:, None] * dr
    expect = (psi ** 2) @ dv
    assert inside.any()
    np.testing.assert_allclose(gel[9:12], expect, rtol=1e-11, atol=1e-14 * np.abs(expect).max())
