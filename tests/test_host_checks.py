"""Host-side logic of the library that runs without a GPU: the polarization
matrix inverse behind PolType 3 (both of its methods), the argument checks
of the PolType 3 potential build's binding and of the bandwidth probes."""
import numpy as np
import pytest

from conftest import golden


def _water_pol_sites():
    g = golden("potential_pol3_c1.npz")
    sites = g["sites"]
    return sites[g["ips"].astype(int)], g["alphas"]


@pytest.mark.parametrize("method", [0, 1, 2])
def test_inverse_of_the_polarization_matrix(pisces, oracle_mod, method):
    # InvA of the (H2O)6 polarizable sites by the oracle's ComputeInvA
    # restatement (pinned to the reference's own InvA through its SCF dipoles);
    # A = InvA^-1 goes through each method and must come back to InvA
    rpol, alpha = _water_pol_sites()
    inva = oracle_mod.pol3_inva(rpol, alpha)
    n = 3 * len(alpha)
    inva = inva[:n, :n]
    A = np.linalg.inv(inva)
    A = 0.5 * (A + A.T)
    out = pisces.sym_inverse(A, method)
    np.testing.assert_allclose(out, out.T, rtol=0, atol=0)
    np.testing.assert_allclose(out, inva, rtol=1e-9, atol=1e-9 * np.abs(inva).max())


def test_indefinite_matrix_takes_gauss_jordan(pisces):
    rng = np.random.default_rng(5)
    Q, _ = np.linalg.qr(rng.standard_normal((30, 30)))
    lam = np.concatenate([rng.uniform(0.5, 2.0, 20), -rng.uniform(0.5, 2.0, 10)])
    A = (Q * lam) @ Q.T
    A = 0.5 * (A + A.T)
    ref = np.linalg.inv(A)
    with pytest.raises(pisces.PiscesError):
        pisces.sym_inverse(A, 1)  # not positive definite: Cholesky alone fails
    for method in (0, 2):
        out = pisces.sym_inverse(A, method)
        np.testing.assert_allclose(out, ref, rtol=0, atol=1e-10 * np.abs(ref).max())


def test_zero_diagonal_needs_a_pivot(pisces):
    # a zero leading pivot: [[0, B], [B^T, 0]] inverts only with row exchanges
    B = np.array([[2.0, 1.0], [0.5, 3.0]])
    A = np.block([[np.zeros((2, 2)), B], [B.T, np.zeros((2, 2))]])
    out = pisces.sym_inverse(A, 0)
    np.testing.assert_allclose(out @ A, np.eye(4), rtol=0, atol=1e-14)


def test_singular_matrix_is_an_error(pisces):
    A = np.ones((4, 4))
    for method in (0, 1, 2):
        with pytest.raises(pisces.PiscesError):
            pisces.sym_inverse(A, method)


def _scp_args(g):
    return dict(flags=g["flags_scp"], potpara=g["potpara_scp"], sites=g["sites"], q=g["q"],
                iq=g["iq"], alpha=g["alphas"], ip=g["ips"], epc=g["epc"])


def test_scp_arrays_accepts_consistent_lists(pisces):
    g = golden("potential_pol3_c1.npz")
    f, pp, s, q, iq, alpha, ip, epc, ix, iqq, nion = pisces.scp_arrays(**_scp_args(g))
    assert s.size == 3 * len(g["sites"]) and nion == 0
    assert alpha.size == ip.size and epc.size == 3 * ip.size


@pytest.mark.parametrize("bad", ["alpha", "epc", "iq", "ip_range", "iq_range", "sites", "ions"])
def test_scp_arrays_rejects_mismatched_lists(pisces, bad):
    g = golden("potential_pol3_c1.npz")
    a = _scp_args(g)
    ions = None
    if bad == "alpha":
        a["alpha"] = a["alpha"][:-1]
    elif bad == "epc":
        a["epc"] = a["epc"][:-3]
    elif bad == "iq":
        a["iq"] = a["iq"][:-1]
    elif bad == "ip_range":
        a["ip"] = a["ip"].copy()
        a["ip"][0] = len(g["sites"])
    elif bad == "iq_range":
        a["iq"] = a["iq"].copy()
        a["iq"][0] = -1
    elif bad == "sites":
        a["sites"] = np.asarray(a["sites"]).reshape(-1)[:-1]
    else:
        ions = (np.zeros(4), np.ones(1))
    with pytest.raises(ValueError):
        pisces.scp_arrays(**a, ions=ions)


def test_dual_coarse_grid_matches_oracle(pisces, oracle_mod):
    """pisces_dual_grid_coarse (host arithmetic) against the oracle's coarse
    grid for odd and even n, and its refusals: non-cubic grids and grids too
    small to have a coarse one."""
    for n in range(5, 130, 7):
        for L in (30.0, 47.5):
            nc, Lc = pisces.dual_coarse((n,) * 3, (L,) * 3)
            nco, Lco = oracle_mod.dual_coarse(n, L)
            assert nc == (nco,) * 3
            assert Lc == (Lco,) * 3
    with pytest.raises(pisces.PiscesError):
        pisces.dual_coarse((32, 32, 33), (40.0,) * 3)
    with pytest.raises(pisces.PiscesError):
        pisces.dual_coarse((3, 3, 3), (40.0,) * 3)


def test_probe_argument_checks(pisces):
    """The copy and read-stream probes reject sizes that are not positive
    multiples of 16 bytes, non-positive repetition counts and a missing result
    pointer before they touch the device."""
    import ctypes
    L = pisces.lib()
    ms = ctypes.c_double(0)
    L.pisces_copy_probe.argtypes = [ctypes.c_int, ctypes.c_longlong, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_double)]
    L.pisces_stream_probe.argtypes = [ctypes.c_int, ctypes.c_longlong, ctypes.c_int, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_double)]
    for nbytes, reps in ((24, 1), (0, 1), (-16, 1), (1024, 0)):
        assert L.pisces_copy_probe(0, nbytes, reps, ctypes.byref(ms)) == -1
        for nt in (0, 1):
            assert L.pisces_stream_probe(0, nbytes, reps, nt, ctypes.byref(ms)) == -1
    assert L.pisces_copy_probe(0, 1024, 1, None) == -1
    assert L.pisces_stream_probe(0, 1024, 1, 0, None) == -1
