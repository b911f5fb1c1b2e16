"""ctypes binding of the CPU restatement (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, by the smoke check and by the CPU
baseline leg of bench.py, always as the checker or the timed CPU reference --
never by the product path.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DRIVER = os.path.join(HERE, "_ref", "ref_driver")

_lib = None

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)


def _p(a):
    return a.ctypes.data_as(_D)


def _pi(a):
    return a.ctypes.data_as(_I)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        _lib = ctypes.CDLL(LIB_PATH)
    return _lib


def set_threads(n):
    lib().or_set_threads(int(n))


def sine_dvr(n, L, mass=1.0, packed=False):
    """Grid and kinetic matrix, src/sine_dvr.cpp:45-131 on [-L/2, L/2]."""
    x = np.zeros(n)
    if packed:
        T = np.zeros(n * (n + 1) // 2)
        lib().or_sine_dvr_packed(n, ctypes.c_double(mass), ctypes.c_double(-0.5 * L),
                                 ctypes.c_double(0.5 * L), _p(x), _p(T))
    else:
        T = np.zeros((n, n))
        lib().or_sine_dvr(n, ctypes.c_double(mass), ctypes.c_double(-0.5 * L),
                          ctypes.c_double(0.5 * L), _p(x), _p(T))
    return x, T


def fft_kinetic(n, L, mass=1.0):
    """DVRType 3 (Tdiag): grid, the circulant kinetic matrix and the Fourier
    kinetic energies in FFT order."""
    x = np.zeros(n)
    T = np.zeros((n, n))
    ek = np.zeros(n)
    lib().or_fft_kinetic(n, ctypes.c_double(mass), ctypes.c_double(-0.5 * L),
                         ctypes.c_double(0.5 * L), _p(x), _p(T), _p(ek))
    return x, T, ek


class Grid:
    """Product grid (x fastest), as DVR::SetupDVR builds it: the sine DVR
    (dvr_type 0) or the FFT kinetic energy on the same points (3)."""

    def __init__(self, n, L, dvr_type=0):
        self.n = tuple(int(v) for v in n)
        self.L = tuple(float(v) for v in L)
        self.dvr_type = dvr_type
        self.x, self.T = [], []
        for a in range(3):
            if dvr_type == 3:
                x, T, _ = fft_kinetic(self.n[a], self.L[a])
            else:
                x, T = sine_dvr(self.n[a], self.L[a])
            self.x.append(x)
            self.T.append(np.ascontiguousarray(T))
        self.ngp = self.n[0] * self.n[1] * self.n[2]

    def apply_h(self, V, X):
        X = np.ascontiguousarray(X, dtype=np.float64)
        nvec = 1 if X.ndim == 1 else X.shape[0]
        Y = np.zeros_like(X)
        nx, ny, nz = self.n
        lib().or_apply_h(nx, ny, nz, _p(self.T[0]), _p(self.T[1]), _p(self.T[2]),
                         _p(np.ascontiguousarray(V)), _p(X), _p(Y), nvec)
        return Y

    def diag(self, V):
        D = np.zeros(self.ngp)
        nx, ny, nz = self.n
        lib().or_diag(nx, ny, nz, _p(self.T[0]), _p(self.T[1]), _p(self.T[2]),
                      _p(np.ascontiguousarray(V)), _p(D))
        return D

    def piab(self):
        wf = np.zeros(self.ngp)
        nx, ny, nz = self.n
        lib().or_piab_start(nx, ny, nz, _p(self.x[0]), _p(self.x[1]), _p(self.x[2]), _p(wf))
        return wf

    def solve(self, V, nroots, maxsub, maxmacro, ptol, corrflag, start):
        """Davidson (src/Davidson.cpp) on this grid; start: (nroots, ngp)."""
        B = np.zeros((maxsub, self.ngp))
        B[:nroots] = start[:nroots]
        ev = np.zeros(nroots)
        res = np.zeros(nroots)
        nm = ctypes.c_int(0)
        nmac = ctypes.c_int(0)
        nx, ny, nz = self.n
        nconv = lib().or_solve_grid(nx, ny, nz, _p(self.T[0]), _p(self.T[1]), _p(self.T[2]),
                                    _p(np.ascontiguousarray(V)), nroots, maxsub, maxmacro,
                                    ptol, corrflag, _p(B), _p(ev), ctypes.byref(nm),
                                    ctypes.byref(nmac), _p(res))
        return dict(nconv=nconv, evals=ev, vecs=B[:nroots].copy(), n_mtx=nm.value,
                    n_macro=nmac.value, resid=res)


def syev(A):
    n = A.shape[0]
    Af = np.asfortranarray(A, dtype=np.float64)
    W = np.zeros(n)
    Vv = np.zeros((n, n), order="F")
    lib().or_syev(n, _p(Af), n, _p(W), _p(Vv), n)
    return W, Vv


def water_sites(xyz_ang):
    xyz = np.ascontiguousarray(np.asarray(xyz_ang, dtype=np.float64).reshape(-1, 9))
    nw = xyz.shape[0]
    sites = np.zeros((nw, 12))
    conf = np.zeros((nw, 6))
    rc = lib().or_water_sites(nw, _p(xyz), _p(sites), _p(conf))
    if rc != 0:
        raise RuntimeError("water projection produced non-finite values")
    return sites.reshape(-1, 3), conf


def water_induced(sites):
    nw = sites.shape[0] // 4
    s = np.ascontiguousarray(sites, dtype=np.float64)
    mu = np.zeros(9 * nw)
    epc = np.zeros(9 * nw)
    ep = ctypes.c_double(0)
    it = ctypes.c_int(0)
    rc = lib().or_water_induced(nw, _p(s), _p(mu), _p(epc), ctypes.byref(ep), ctypes.byref(it))
    return dict(mu=mu, epc=epc, e_pol=ep.value, iters=it.value, converged=(rc == 0))


def water_energy(sites, mu):
    """Neutral-cluster energy: (E0, elec, pol, disp) in hartree."""
    nw = sites.shape[0] // 4
    out = np.zeros(4)
    lib().or_water_energy(nw, _p(np.ascontiguousarray(sites, dtype=np.float64)),
                          _p(np.ascontiguousarray(mu, dtype=np.float64)), _p(out))
    return out


def damp(r, ai, aj, a):
    """Charge-dipole and dipole-dipole damping factors of the water model."""
    r, ai, aj, a = (np.ascontiguousarray(v, dtype=np.float64) for v in (r, ai, aj, a))
    dc = np.zeros(r.size)
    dd = np.zeros(r.size)
    lib().or_damp(r.size, _p(r), _p(ai), _p(aj), _p(a), _p(dc), _p(dd))
    return dc, dd


def water_lists(sites, mu):
    nw = sites.shape[0] // 4
    q = np.zeros(3 * nw)
    iq = np.zeros(3 * nw, dtype=np.int32)
    d = np.zeros(9 * nw)
    idd = np.zeros(3 * nw, dtype=np.int32)
    lib().or_water_lists(nw, _p(np.ascontiguousarray(sites)), _p(np.ascontiguousarray(mu)),
                         _p(q), _pi(iq), _p(d), _pi(idd))
    return q, iq, d, idd


DEFAULT_FLAGS = (3, 2, 0, 1, 0, 0)


def default_potpara():
    """PotPara defaults for Potential 1-4, src/GetInput.cpp:53-61."""
    pp = np.zeros(11)
    pp[0] = 9.745
    pp[1] = 2.536
    pp[2] = 4.4837
    pp[3] = 1.7724
    pp[4] = 1.7724
    pp[5] = 4.4837
    pp[6] = 0.5 / np.sqrt(pp[0])
    pp[10] = 20.0
    return pp


class Cluster:
    """Sites and lists for the electron, in the order ClusterAnion builds them."""

    def __init__(self, xyz_ang, ions=None, ion_r=2.9):
        self.sites, self.conf = water_sites(xyz_ang)
        ind = water_induced(self.sites)
        self.mu, self.epc, self.e_pol = ind["mu"], ind["epc"], ind["e_pol"]
        self.scf_iters = ind["iters"]
        self.e0, self.e_elec, _, self.e_disp = water_energy(self.sites, self.mu)
        self.q, self.iq, self.d, self.id = water_lists(self.sites, self.mu)
        if ions is None:
            self.ion_xyz = np.zeros((0, 3))
            self.ion_q = np.zeros(0)
        else:
            self.ion_xyz = np.ascontiguousarray(np.asarray(ions[0], dtype=np.float64).reshape(-1, 3))
            self.ion_q = np.ascontiguousarray(np.asarray(ions[1], dtype=np.float64).reshape(-1))
        self.ion_r = float(ion_r)

    def _args(self, flags, potpara):
        f = np.ascontiguousarray(np.asarray(flags, dtype=np.int32))
        pp = np.ascontiguousarray(np.asarray(potpara, dtype=np.float64))
        return f, pp

    def potential_points(self, pts, flags=DEFAULT_FLAGS, potpara=None):
        f, pp = self._args(flags, default_potpara() if potpara is None else potpara)
        pts = np.ascontiguousarray(np.asarray(pts, dtype=np.float64).reshape(-1, 3))
        v = np.zeros(pts.shape[0])
        parts = np.zeros((pts.shape[0], 5))
        rc = lib().or_potential_points(
            _pi(f), _p(pp), self.sites.shape[0], _p(self.sites), len(self.q), _p(self.q),
            _pi(self.iq), len(self.id), _p(self.d), _pi(self.id), len(self.ion_q),
            _p(self.ion_xyz), _p(self.ion_q), ctypes.c_double(self.ion_r), pts.shape[0],
            _p(pts), _p(v), _p(parts))
        if rc != 0:
            raise ValueError("unsupported potential flags (rc=%d)" % rc)
        return v, parts

    def potential_grid(self, grid, flags=DEFAULT_FLAGS, potpara=None):
        f, pp = self._args(flags, default_potpara() if potpara is None else potpara)
        v = np.zeros(grid.ngp)
        nx, ny, nz = grid.n
        rc = lib().or_potential_grid(
            _pi(f), _p(pp), self.sites.shape[0], _p(self.sites), len(self.q), _p(self.q),
            _pi(self.iq), len(self.id), _p(self.d), _pi(self.id), len(self.ion_q),
            _p(self.ion_xyz), _p(self.ion_q), ctypes.c_double(self.ion_r), nx, ny, nz,
            _p(grid.x[0]), _p(grid.x[1]), _p(grid.x[2]), _p(v))
        if rc != 0:
            raise ValueError("unsupported potential flags (rc=%d)" % rc)
        return v


    # PolType 3: polarizable sites O,H1,H2 of every water, their permanent-
    # charge field and VpolWaterWater = the water model's polarization energy
    def pol3_lists(self):
        nw = self.sites.shape[0] // 4
        a3 = (1.0 / 0.52917720859) ** 3
        alpha = np.tile([1.22 * a3, 0.28 * a3, 0.28 * a3], nw)
        ip = np.array([4 * w + s for w in range(nw) for s in range(3)], dtype=np.int32)
        return alpha, ip, np.ascontiguousarray(self.epc), float(self.e_pol)

    def potential_points_pol3(self, pts, flags, potpara=None, pol=None):
        f, pp = self._args(flags, default_potpara() if potpara is None else potpara)
        alpha, ip, epc, vww = self.pol3_lists() if pol is None else pol
        pts = np.ascontiguousarray(np.asarray(pts, dtype=np.float64).reshape(-1, 3))
        v = np.zeros(pts.shape[0])
        parts = np.zeros((pts.shape[0], 5))
        rc = lib().or_potential_points_pol3(
            _pi(f), _p(pp), self.sites.shape[0], _p(self.sites), len(self.q), _p(self.q),
            _pi(self.iq), len(ip), _p(np.ascontiguousarray(alpha, dtype=np.float64)),
            _pi(np.ascontiguousarray(ip, dtype=np.int32)), _p(np.ascontiguousarray(epc)),
            ctypes.c_double(vww), pts.shape[0], _p(pts), _p(v), _p(parts))
        if rc != 0:
            raise ValueError("unsupported potential flags (rc=%d)" % rc)
        return v, parts

    def potential_grid_pol3(self, grid, flags, potpara=None):
        f, pp = self._args(flags, default_potpara() if potpara is None else potpara)
        alpha, ip, epc, vww = self.pol3_lists()
        v = np.zeros(grid.ngp)
        nx, ny, nz = grid.n
        rc = lib().or_potential_grid_pol3(
            _pi(f), _p(pp), self.sites.shape[0], _p(self.sites), len(self.q), _p(self.q),
            _pi(self.iq), len(ip), _p(alpha), _pi(ip), _p(epc), ctypes.c_double(vww),
            nx, ny, nz, _p(grid.x[0]), _p(grid.x[1]), _p(grid.x[2]), _p(v))
        if rc != 0:
            raise ValueError("unsupported potential flags (rc=%d)" % rc)
        return v


def gradient_grid(cl, grid, psi, flags=(3, 2, 0, 3, 0, 0), potpara=None):
    """Electron-side gradient reductions (gel, tmu, mcm) of cluster cl on grid
    for the wavefunction psi (the reference's ComputeGradient grid loop)."""
    f = np.ascontiguousarray(np.asarray(flags, dtype=np.int32))
    pp = np.ascontiguousarray(default_potpara() if potpara is None else potpara, dtype=np.float64)
    alpha, ip, epc, _ = cl.pol3_lists()
    ns = cl.sites.shape[0]
    n3 = 3 * ip.size
    gel, tmu, mcm = np.zeros(3 * ns), np.zeros(n3), np.zeros(n3 * n3)
    nx, ny, nz = grid.n
    rc = lib().or_gradient_grid(
        _pi(f), _p(pp), ns, _p(cl.sites), len(cl.q), _p(cl.q), _pi(cl.iq), ip.size,
        _p(np.ascontiguousarray(alpha)), _pi(np.ascontiguousarray(ip, dtype=np.int32)),
        _p(np.ascontiguousarray(epc)), nx, ny, nz, _p(grid.x[0]), _p(grid.x[1]), _p(grid.x[2]),
        _p(np.ascontiguousarray(psi, dtype=np.float64)), _p(gel), _p(tmu), _p(mcm))
    if rc != 0:
        raise ValueError("unsupported flags for the gradient (rc=%d)" % rc)
    return gel, tmu, mcm.reshape(n3, n3)


def gradient_final(nsites, np_, gel, tmu, mcm, dTx, dTy, dTz, dEfield, gww, polgrad):
    out = np.zeros(3 * nsites)
    a = [np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
         for v in (gel, tmu, mcm, dTx, dTy, dTz, dEfield, gww, polgrad)]
    lib().or_gradient_final(nsites, np_, *[_p(v) for v in a], _p(out))
    return out


def dual_coarse(n, L):
    nc = ctypes.c_int(0)
    Lc = ctypes.c_double(0)
    lib().or_dual_coarse(int(n), ctypes.c_double(L), ctypes.byref(nc), ctypes.byref(Lc))
    return nc.value, Lc.value


def dual_weights(kind, rx, ry, rz):
    w = np.zeros(8)
    lib().or_dual_weights(int(kind), int(rx), int(ry), int(rz), _p(w))
    return w


def dual_extend(n, nc, kind, coarse):
    fine = np.zeros(n ** 3)
    rc = lib().or_dual_extend(int(n), int(nc), int(kind),
                              _p(np.ascontiguousarray(coarse, dtype=np.float64)), _p(fine))
    if rc != 0:
        raise ValueError("coarse grid does not match")
    return fine


def pol3_inva(rpol, alpha):
    """InvA = (alpha^-1 - T)^-1 of the polarizable sites (row-major 3np x 3np)."""
    rp = np.ascontiguousarray(np.asarray(rpol, dtype=np.float64).reshape(-1))
    np_ = rp.size // 3
    out = np.zeros((3 * np_, 3 * np_))
    rc = lib().or_pol3_inva(np_, _p(rp), _p(np.ascontiguousarray(alpha, dtype=np.float64)),
                            _p(out))
    if rc != 0:
        raise ValueError("singular polarization matrix")
    return out


# ---------------------------------------------------------------- reference runs
def write_records(path, recs):
    with open(path, "wb") as f:
        for r in recs:
            a = np.ascontiguousarray(np.asarray(r, dtype=np.float64).reshape(-1))
            f.write(np.int64(a.size).tobytes())
            f.write(a.tobytes())


def read_records(path):
    out = []
    with open(path, "rb") as f:
        data = f.read()
    off = 0
    while off < len(data):
        n = int(np.frombuffer(data, dtype=np.int64, count=1, offset=off)[0])
        off += 8
        out.append(np.frombuffer(data, dtype=np.float64, count=n, offset=off).copy())
        off += 8 * n
    return out


def ref_available():
    return os.path.exists(REF_DRIVER)
