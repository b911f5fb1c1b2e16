"""The oracle's electron-side analytic gradient against the reference.

grad_*.npz hold the reference's ClusterAnion::GetAnalGrad
(src/ClusterAnion.cpp:322-430) with DVR::ComputeGradient's grid loop
(src/DVR.cpp:1099-1176) over its own Potential::EvaluateGradient,
FinalGradient and SubtractWWGradient (src/Gradients.cpp) and its WaterCluster
(CalcGradient, Calc_dInducedDipoles) and BuildSuperdT, for the PolType 3
ground state of C1 (16^3) and C2 (20^3).  Independently of any reference code,
the gradient must be the derivative of the potential: central differences of
sum_g psi^2 V(r_g) under a displacement of one site, with the permanent-charge
field held (as the electron part holds it), equal the electron terms plus the
dipole-tensor term -sum dT/dR : (sum psi^2 mu mu^T)."""
import numpy as np
import pytest

from conftest import golden


def _cluster_grid(oracle_mod, f):
    n, L = tuple(int(v) for v in f["n"]), tuple(float(v) for v in f["L"])
    return oracle_mod.Cluster(f["xyz"]), oracle_mod.Grid(n, L)


@pytest.mark.parametrize("tag", ["c1", "c2"])
def test_gradient_terms_match_reference(oracle_mod, tag):
    f = golden("grad_%s.npz" % tag)
    cl, g = _cluster_grid(oracle_mod, f)
    gel, tmu, mcm = oracle_mod.gradient_grid(cl, g, f["psi"])
    assert np.abs(gel - f["gel"]).max() <= 1e-12 * np.abs(f["gel"]).max()
    assert np.abs(tmu - f["tmu"]).max() <= 1e-12 * np.abs(f["tmu"]).max()
    assert np.abs(mcm.reshape(-1) - f["mcm"]).max() <= 1e-12 * np.abs(f["mcm"]).max()
    nsites, np_ = cl.sites.shape[0], f["tmu"].size // 3
    gf = oracle_mod.gradient_final(nsites, np_, gel, tmu, mcm, f["dTx"], f["dTy"], f["dTz"],
                                   f["dEfield"], f["gww"], f["polgrad"])
    assert np.abs(gf - f["gfinal"]).max() <= 1e-12 * np.abs(f["gfinal"]).max()


def test_gradient_is_the_derivative_of_the_potential(oracle_mod):
    f = golden("grad_c1.npz")
    cl, g = _cluster_grid(oracle_mod, f)
    psi = f["psi"]
    w = psi * psi
    flags, pp = f["flags"], f["potpara"]
    alpha, ip, epc, vww = cl.pol3_lists()
    X, Y, Z = np.meshgrid(g.x[0], g.x[1], g.x[2], indexing="ij")
    pts = np.stack([X.transpose(2, 1, 0).reshape(-1), Y.transpose(2, 1, 0).reshape(-1),
                    Z.transpose(2, 1, 0).reshape(-1)], axis=1)  # x fastest
    gel = f["gel"]
    np_ = ip.size
    mcm = f["mcm"].reshape(3 * np_, 3 * np_)
    dT = [f[k].reshape(3 * np_, 3 * np_) for k in ("dTx", "dTy", "dTz")]
    h = 1e-4
    base = cl.sites.copy()
    scale = np.abs(gel).max()
    for s in (0, 1, 2, 3, 4, 9, 14, 23):   # O, H, H and M sites of several waters
        for a in range(3):
            F = []
            for sign in (1, -1):
                cl.sites = base.copy()
                cl.sites[s, a] += sign * h
                v, _ = cl.potential_points_pol3(pts, flags, pp, pol=(alpha, ip, epc, vww))
                F.append(w @ v)
            cl.sites = base
            fd = (F[0] - F[1]) / (2 * h)
            an = gel[3 * s + a]
            if s % 4 != 3:  # a polarizable site: the dipole tensor moves with it
                j = 3 * (s // 4) + s % 4
                an -= np.sum(dT[a][3 * j:3 * j + 3, :] * mcm[3 * j:3 * j + 3, :])
            assert abs(fd - an) <= 1e-6 * scale, (s, a, fd, an)
