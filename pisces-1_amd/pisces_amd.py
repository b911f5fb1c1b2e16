"""Python binding of the MI355X PISCES library (lib/libpisces_hip.so) via ctypes.

The library is the product; this module only marshals numpy arrays through the
C ABI of include/pisces_hip.h.  It raises if the library is not built: there is
no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# PISCES_LIB names another build of the library (A/B timing of kernel variants)
LIB_PATH = os.environ.get("PISCES_LIB") or os.path.join(HERE, "lib", "libpisces_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "pisces_hip.h")

FP64, FP32, BF16X2 = 0, 1, 2
PRECISIONS = {"fp64": FP64, "fp32": FP32, "bf16x2": BF16X2}

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_lib = None


class PiscesError(RuntimeError):
    pass


class Stats(ctypes.Structure):
    _fields_ = [
        ("nconv", ctypes.c_int),
        ("n_matvec", ctypes.c_int),
        ("n_matvec_lo", ctypes.c_int),
        ("n_iter", ctypes.c_int),
        ("n_macro", ctypes.c_int),
        ("n_orth_data", ctypes.c_int),
        ("seconds", ctypes.c_double),
        ("resid", ctypes.c_double * 64),
    ]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PiscesError("pisces-1_amd/lib/libpisces_hip.so is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        L.pisces_version.restype = ctypes.c_char_p
        L.pisces_strerror.restype = ctypes.c_char_p
        L.pisces_last_error.restype = ctypes.c_char_p
        L.pisces_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), _I, _D, ctypes.c_int]
        L.pisces_destroy.argtypes = [ctypes.c_void_p]
        L.pisces_destroy.restype = None
        for name in ("pisces_grid", "pisces_set_potential", "pisces_get_potential",
                     "pisces_apply_h", "pisces_diag", "pisces_piab", "pisces_solve",
                     "pisces_build_potential", "pisces_build_cluster_potential",
                     "pisces_build_potential_scp",
                     "pisces_apply_h_device", "pisces_apply_h_ex", "pisces_time_hx", "pisces_time_precond",
                     "pisces_solve_block", "pisces_create_ex"):
            getattr(L, name).restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_D)


def _pi(a):
    return a.ctypes.data_as(_I)


def check(rc, what=""):
    if rc < 0:
        msg = lib().pisces_last_error().decode()
        raise PiscesError("%s failed: %s (%s)" % (what, lib().pisces_strerror(rc).decode(), msg))
    return rc


def exported_symbols():
    """Function names declared in include/pisces_hip.h."""
    import re
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(pisces_[a-z0-9_]+)\s*\(", txt)))


class DeviceArray:
    """Device buffer of float64 values allocated through the HIP runtime the
    library itself links (libamdhip64.so.7), for the device-pointer entry
    points.  torch's wheel bundles a different HIP runtime, so its tensors are
    not used here."""
    _hip = None

    @classmethod
    def hip(cls):
        if cls._hip is None:
            lib()  # loads the library and with it its HIP runtime
            h = ctypes.CDLL("libamdhip64.so.7")
            h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
            h.hipFree.argtypes = [ctypes.c_void_p]
            h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
            h.hipSetDevice.argtypes = [ctypes.c_int]
            for f in ("hipMalloc", "hipFree", "hipMemcpy", "hipSetDevice"):
                getattr(h, f).restype = ctypes.c_int
            cls._hip = h
        return cls._hip

    def __init__(self, shape, device=0):
        self.shape = tuple(shape)
        self.nbytes = 8 * int(np.prod(self.shape))
        self.ptr = ctypes.c_void_p()
        h = self.hip()
        if h.hipSetDevice(int(device)) or h.hipMalloc(ctypes.byref(self.ptr), self.nbytes):
            raise PiscesError("hipMalloc of %d bytes failed" % self.nbytes)

    @classmethod
    def from_host(cls, a, device=0):
        a = np.ascontiguousarray(a, dtype=np.float64)
        d = cls(a.shape, device)
        if cls.hip().hipMemcpy(d.ptr, a.ctypes.data, d.nbytes, 1):
            raise PiscesError("hipMemcpy host to device failed")
        return d

    def to_host(self):
        a = np.empty(self.shape)
        if self.hip().hipMemcpy(a.ctypes.data, self.ptr, self.nbytes, 2):
            raise PiscesError("hipMemcpy device to host failed")
        return a

    def free(self):
        if self.ptr:
            self.hip().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def dual_coarse(n, L):
    """The coarse grid of the dual-grid method: ((nc,)*3, (Lc,)*3)."""
    na = np.ascontiguousarray(n, dtype=np.int32)
    La = np.ascontiguousarray(L, dtype=np.float64)
    nc = np.zeros(3, dtype=np.int32)
    Lc = np.zeros(3)
    L_ = lib()
    L_.pisces_dual_grid_coarse.argtypes = [_I, _D, _I, _D]
    L_.pisces_dual_grid_coarse.restype = ctypes.c_int
    check(L_.pisces_dual_grid_coarse(_pi(na), _p(La), _pi(nc), _p(Lc)), "pisces_dual_grid_coarse")
    return tuple(int(v) for v in nc), tuple(float(v) for v in Lc)


def hx_spin_timeouts():
    L = lib()
    L.pisces_debug_hx_spin_timeouts.restype = ctypes.c_uint
    return int(L.pisces_debug_hx_spin_timeouts())


def sym_inverse(a, method=0):
    """Inverse of a symmetric matrix by the PolType 3 set-up's host routine
    (pisces_debug_sym_inverse): method 0 Cholesky else Gauss-Jordan, 1 Cholesky
    only, 2 Gauss-Jordan only.  Raises PiscesError where the routine fails."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.ndim != 2 or a.shape[0] != a.shape[1]:
        raise ValueError("a square matrix expected")
    out = np.empty_like(a)
    check(lib().pisces_debug_sym_inverse(a.shape[0], _p(a), int(method), _p(out)),
          "pisces_debug_sym_inverse")
    return out


def noalloc_scratch_bytes():
    L = lib()
    L.pisces_noalloc_scratch_bytes.restype = ctypes.c_size_t
    return int(L.pisces_noalloc_scratch_bytes())


def scp_arrays(flags, potpara, sites, q, iq, alpha, ip, epc, ions=None):
    """Contiguous, typed and size-checked arrays of pisces_build_potential_scp:
    sites (ns, 3); q and iq (charges and their site indices) of one length;
    alpha one polarizability per polarizable site ip; epc 3 per polarizable
    site; every index inside [0, ns).  A mismatch raises ValueError before the
    C side could read past a host array."""
    f = np.ascontiguousarray(flags, dtype=np.int32)
    pp = np.ascontiguousarray(potpara, dtype=np.float64)
    s = np.ascontiguousarray(sites, dtype=np.float64).reshape(-1)
    q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
    iq = np.ascontiguousarray(iq, dtype=np.int32).reshape(-1)
    alpha = np.ascontiguousarray(alpha, dtype=np.float64).reshape(-1)
    ip = np.ascontiguousarray(ip, dtype=np.int32).reshape(-1)
    epc = np.ascontiguousarray(epc, dtype=np.float64).reshape(-1)
    if f.size < 6:
        raise ValueError("flags: 6 values expected, got %d" % f.size)
    if s.size % 3:
        raise ValueError("sites: %d values, not a multiple of 3" % s.size)
    ns = s.size // 3
    if iq.size != q.size:
        raise ValueError("iq: %d indices for %d charges" % (iq.size, q.size))
    if alpha.size != ip.size:
        raise ValueError("alpha: %d polarizabilities for %d polarizable sites"
                         % (alpha.size, ip.size))
    if epc.size != 3 * ip.size:
        raise ValueError("epc: %d values for %d polarizable sites (3 each)"
                         % (epc.size, ip.size))
    for name, idx in (("iq", iq), ("ip", ip)):
        if idx.size and (idx.min() < 0 or idx.max() >= ns):
            raise ValueError("%s: site index outside [0, %d)" % (name, ns))
    if ions is None:
        ix, iqq, nion = np.zeros(3), np.zeros(1), 0
    else:
        ix = np.ascontiguousarray(ions[0], dtype=np.float64).reshape(-1)
        iqq = np.ascontiguousarray(ions[1], dtype=np.float64).reshape(-1)
        nion = iqq.size
        if ix.size != 3 * nion:
            raise ValueError("ions: %d coordinates for %d charges" % (ix.size, nion))
    return f, pp, s, q, iq, alpha, ip, epc, ix, iqq, nion


class Context:
    """One grid on one GPU (pisces_create; pisces_create_ex for the FFT
    kinetic energy, dvr_type 3)."""

    def __init__(self, n, L, device=0, dvr_type=0):
        self.n = tuple(int(v) for v in n)
        self.L = tuple(float(v) for v in L)
        self.ngp = self.n[0] * self.n[1] * self.n[2]
        h = ctypes.c_void_p()
        na = np.array(self.n, dtype=np.int32)
        La = np.array(self.L, dtype=np.float64)
        if dvr_type == 0:
            check(lib().pisces_create(ctypes.byref(h), _pi(na), _p(La), int(device)),
                  "pisces_create")
        else:
            check(lib().pisces_create_ex(ctypes.byref(h), _pi(na), _p(La), int(device),
                                         int(dvr_type)), "pisces_create_ex")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().pisces_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_hx_variant(self, v):
        """H*X kernels of this context: 1 matrix cores (default), 2 matrix cores
        in separate launches always, 3 split-BF16 H*X and preconditioner in one
        XCD-ordered launch each, 0 VALU."""
        L = lib()
        L.pisces_debug_set_hx_variant.argtypes = [ctypes.c_void_p, ctypes.c_int]
        check(L.pisces_debug_set_hx_variant(self.h, int(v)), "pisces_debug_set_hx_variant")

    def grid(self):
        xyz = np.zeros(sum(self.n))
        T = np.zeros(sum(v * v for v in self.n))
        check(lib().pisces_grid(self.h, _p(xyz), _p(T)), "pisces_grid")
        xs, Ts, ox, ot = [], [], 0, 0
        for v in self.n:
            xs.append(xyz[ox:ox + v].copy())
            Ts.append(T[ot:ot + v * v].reshape(v, v).copy())
            ox += v
            ot += v * v
        return xs, Ts

    def set_potential(self, V):
        V = np.ascontiguousarray(V, dtype=np.float64).reshape(-1)
        assert V.size == self.ngp
        check(lib().pisces_set_potential(self.h, _p(V)), "pisces_set_potential")

    def get_potential(self):
        V = np.zeros(self.ngp)
        check(lib().pisces_get_potential(self.h, _p(V)), "pisces_get_potential")
        return V

    def build_potential(self, flags, potpara, sites, q, iq, d, idd, ions=None, ion_r=2.9):
        f = np.ascontiguousarray(flags, dtype=np.int32)
        pp = np.ascontiguousarray(potpara, dtype=np.float64)
        s = np.ascontiguousarray(sites, dtype=np.float64).reshape(-1)
        q = np.ascontiguousarray(q, dtype=np.float64)
        iq = np.ascontiguousarray(iq, dtype=np.int32)
        d = np.ascontiguousarray(d, dtype=np.float64).reshape(-1)
        idd = np.ascontiguousarray(idd, dtype=np.int32)
        if ions is None:
            ix, iqq, nion = np.zeros(3), np.zeros(1), 0
        else:
            ix = np.ascontiguousarray(ions[0], dtype=np.float64).reshape(-1)
            iqq = np.ascontiguousarray(ions[1], dtype=np.float64).reshape(-1)
            nion = iqq.size
        check(lib().pisces_build_potential(
            self.h, _pi(f), _p(pp), s.size // 3, _p(s), q.size, _p(q), _pi(iq), idd.size, _p(d),
            _pi(idd), nion, _p(ix), _p(iqq), ctypes.c_double(ion_r)), "pisces_build_potential")

    def build_potential_scp(self, flags, potpara, sites, q, iq, alpha, ip, epc, vpol_ww,
                            ions=None, ion_r=2.9):
        """Potential with self-consistent polarizable sites (flags[3] == 3)."""
        f, pp, s, q, iq, alpha, ip, epc, ix, iqq, nion = scp_arrays(
            flags, potpara, sites, q, iq, alpha, ip, epc, ions)
        check(lib().pisces_build_potential_scp(
            self.h, _pi(f), _p(pp), s.size // 3, _p(s), q.size, _p(q), _pi(iq), ip.size,
            _p(alpha), _pi(ip), _p(epc), ctypes.c_double(vpol_ww), nion, _p(ix), _p(iqq),
            ctypes.c_double(ion_r)), "pisces_build_potential_scp")

    def build_cluster_potential(self, xyz_ang, flags=(3, 2, 0, 1, 0, 0), potpara=None,
                                ions=None, ion_r=2.9):
        """Water model (GPU dipole SCF) + grid potential; returns (sites, mu, rc)."""
        xyz = np.ascontiguousarray(xyz_ang, dtype=np.float64).reshape(-1)
        nw = xyz.size // 9
        f = np.ascontiguousarray(flags, dtype=np.int32)
        pp = np.ascontiguousarray(default_potpara() if potpara is None else potpara, dtype=np.float64)
        sites = np.zeros(12 * nw)
        mu = np.zeros(9 * nw)
        if ions is None:
            ix, iqq, nion = np.zeros(3), np.zeros(1), 0
        else:
            ix = np.ascontiguousarray(ions[0], dtype=np.float64).reshape(-1)
            iqq = np.ascontiguousarray(ions[1], dtype=np.float64).reshape(-1)
            nion = iqq.size
        rc = check(lib().pisces_build_cluster_potential(
            self.h, nw, _p(xyz), _pi(f), _p(pp), nion, _p(ix), _p(iqq), ctypes.c_double(ion_r),
            _p(sites), _p(mu)), "pisces_build_cluster_potential")
        return sites.reshape(-1, 3), mu, rc

    def apply_h(self, X):
        X = np.ascontiguousarray(X, dtype=np.float64)
        nvec = 1 if X.ndim == 1 else X.shape[0]
        Y = np.zeros_like(X)
        check(lib().pisces_apply_h(self.h, _p(X), _p(Y), nvec), "pisces_apply_h")
        return Y

    def apply_h_ex(self, X, precision):
        """H*X in a chosen precision (X, Y on the host as float64)."""
        X = np.ascontiguousarray(X, dtype=np.float64)
        nvec = 1 if X.ndim == 1 else X.shape[0]
        Y = np.zeros_like(X)
        check(lib().pisces_apply_h_ex(self.h, _p(X), _p(Y), nvec, int(precision)),
              "pisces_apply_h_ex")
        return Y

    def kinetic_precond(self, r, shift, scale=-1.0, floor=1e-3, precision=FP64):
        """t = scale (T - shift)^{-1} r through the 1-D eigenbases (pisces_kinetic_precond)."""
        r = np.ascontiguousarray(np.asarray(r, dtype=np.float64).reshape(-1, self.ngp))
        t = np.zeros_like(r)
        L = lib()
        L.pisces_kinetic_precond.argtypes = [ctypes.c_void_p, ctypes.c_int, _D, ctypes.c_double,
                                             ctypes.c_double, ctypes.c_double, ctypes.c_int, _D]
        check(L.pisces_kinetic_precond(self.h, r.shape[0], _p(r), float(shift), float(scale),
                                       float(floor), int(precision), _p(t)),
              "pisces_kinetic_precond")
        return t

    def gradient_terms(self, psi, nsites, npol):
        """Electron side of the analytic gradient (pisces_gradient_terms) for
        the context's PolType 3 potential: (gel[3 nsites], tmu[3 npol],
        mcm[3 npol, 3 npol])."""
        psi = np.ascontiguousarray(psi, dtype=np.float64).reshape(-1)
        assert psi.size == self.ngp
        gel, tmu, mcm = np.zeros(3 * nsites), np.zeros(3 * npol), np.zeros(9 * npol * npol)
        L = lib()
        L.pisces_gradient_terms.argtypes = [ctypes.c_void_p, _D, _D, _D, _D]
        L.pisces_gradient_terms.restype = ctypes.c_int
        check(L.pisces_gradient_terms(self.h, _p(psi), _p(gel), _p(tmu), _p(mcm)),
              "pisces_gradient_terms")
        return gel, tmu, mcm.reshape(3 * npol, 3 * npol)

    def compute_gradient(self, psi, gradient, dTx, dTy, dTz, polgrad, dEfield):
        """DVR::ComputeGradient (pisces_compute_gradient): returns gradient +
        the electron part - polgrad."""
        psi = np.ascontiguousarray(psi, dtype=np.float64).reshape(-1)
        g = np.array(gradient, dtype=np.float64).reshape(-1)
        a = [np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
             for v in (dTx, dTy, dTz, polgrad, dEfield)]
        L = lib()
        L.pisces_compute_gradient.argtypes = [ctypes.c_void_p, _D, ctypes.c_int, _D, _D, _D, _D,
                                              _D, _D]
        L.pisces_compute_gradient.restype = ctypes.c_int
        check(L.pisces_compute_gradient(self.h, _p(psi), g.size // 3, _p(g), *[_p(v) for v in a]),
              "pisces_compute_gradient")
        return g

    def dual_extend(self, nc, coarse, kind=0):
        """Coarse (nc^3) -> this grid: kind 0 wavefunction, 1 potential."""
        nca = np.ascontiguousarray(nc, dtype=np.int32)
        c = np.ascontiguousarray(coarse, dtype=np.float64).reshape(-1)
        fine = np.zeros(self.ngp)
        L = lib()
        L.pisces_dual_grid_extend.argtypes = [ctypes.c_void_p, _I, _D, ctypes.c_int, _D]
        L.pisces_dual_grid_extend.restype = ctypes.c_int
        check(L.pisces_dual_grid_extend(self.h, _pi(nca), _p(c), int(kind), _p(fine)),
              "pisces_dual_grid_extend")
        return fine

    def dual_potential(self, nc, coarse_v, rtol=20.0):
        """Polarization 6: the coarse potential carried over far from the sites."""
        nca = np.ascontiguousarray(nc, dtype=np.int32)
        c = np.ascontiguousarray(coarse_v, dtype=np.float64).reshape(-1)
        L = lib()
        L.pisces_dual_grid_potential.argtypes = [ctypes.c_void_p, _I, _D, ctypes.c_double]
        L.pisces_dual_grid_potential.restype = ctypes.c_int
        check(L.pisces_dual_grid_potential(self.h, _pi(nca), _p(c), float(rtol)),
              "pisces_dual_grid_potential")

    def kinetic_precond_noalloc(self, d_r, d_t, nvec, shift, scale=-1.0, floor=1e-3,
                                precision=FP64):
        """The allocation-free preconditioner (pisces_kinetic_precond_noalloc) on
        device arrays d_r, d_t (DeviceArray of nvec x ngp); returns the code."""
        L = lib()
        L.pisces_kinetic_precond_noalloc.argtypes = [
            ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
            ctypes.c_double, ctypes.c_int, ctypes.c_void_p]
        L.pisces_kinetic_precond_noalloc.restype = ctypes.c_int
        return L.pisces_kinetic_precond_noalloc(self.h, int(nvec), d_r.ptr, float(shift),
                                                float(scale), float(floor), int(precision),
                                                d_t.ptr)

    def davidson_rc(self, start, diag, matvec, maxsub=20, maxmacro=100, ptol=7, corrflag=1):
        """Reverse-communication Davidson (pisces_davidson): the caller's
        matvec(X) -> H X (rows = vectors) is called on every request, exactly as
        DVR::davdriver calls MatrixTimesVector.  Returns (nconv, evals, vecs, n_mtx)."""
        L = lib()
        L.pisces_davidson.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 7 + [
            _D, ctypes.POINTER(ctypes.c_int), _D, _D, _D, _D, ctypes.POINTER(ctypes.c_int)]
        nroots, N = start.shape
        B = np.zeros((maxsub, N))
        Z = np.zeros((maxsub, N))
        B[:nroots] = start
        diag = np.ascontiguousarray(diag, dtype=np.float64)
        ev = np.zeros(nroots)
        work = np.zeros(1)
        nconv = ctypes.c_int(0)
        inout = (ctypes.c_int * 3)()
        n_mtx = 0
        while True:
            ido = check(L.pisces_davidson(self.h, N, maxsub, nroots, maxmacro, ptol, corrflag, 0,
                                          _p(ev), ctypes.byref(nconv), _p(B), _p(Z), _p(diag),
                                          _p(work), inout), "pisces_davidson")
            if ido == 0:
                break
            b0, z0, nv = inout[0], inout[1], inout[2]
            Z[z0:z0 + nv] = matvec(B[b0:b0 + nv])
            n_mtx += nv
        return nconv.value, ev, B[:nroots].copy(), n_mtx

    def time_hx(self, nvec, precision, warmup=5, iters=50):
        ms = ctypes.c_double(0)
        check(lib().pisces_time_hx(self.h, int(nvec), int(precision), int(warmup), int(iters),
                                   ctypes.byref(ms)), "pisces_time_hx")
        return ms.value

    def time_precond(self, nvec, precision, warmup=5, iters=50):
        ms = ctypes.c_double(0)
        check(lib().pisces_time_precond(self.h, int(nvec), int(precision), int(warmup), int(iters),
                                        ctypes.byref(ms)), "pisces_time_precond")
        return ms.value

    def diag(self):
        D = np.zeros(self.ngp)
        check(lib().pisces_diag(self.h, _p(D)), "pisces_diag")
        return D

    def piab(self):
        wf = np.zeros(self.ngp)
        check(lib().pisces_piab(self.h, _p(wf)), "pisces_piab")
        return wf

    def solve(self, nroots, start, maxsub=20, maxmacro=100, ptol=7, corrflag=1,
              precision=FP64):
        vecs = np.ascontiguousarray(np.asarray(start, dtype=np.float64)[:nroots]).copy()
        assert vecs.shape == (nroots, self.ngp)
        ev = np.zeros(nroots)
        st = Stats()
        rc = check(lib().pisces_solve(self.h, nroots, maxsub, maxmacro, ptol, corrflag,
                                      int(precision), _p(vecs), _p(ev), ctypes.byref(st)),
                   "pisces_solve")
        return dict(nconv=rc, evals=ev, vecs=vecs, n_matvec=st.n_matvec,
                    n_matvec_lo=st.n_matvec_lo, n_iter=st.n_iter, n_macro=st.n_macro,
                    seconds=st.seconds, resid=np.array(st.resid[:nroots]))


PRECOND_DIAG, PRECOND_KINETIC = 1, 3


def _solve_block(self, nroots, start, block=None, maxsub=32, max_iter=2000, ptol=7,
                 precision=FP64, precond=PRECOND_KINETIC):
    """Block Davidson (pisces_solve_block)."""
    block = nroots if block is None else block
    vecs = np.ascontiguousarray(np.asarray(start, dtype=np.float64)[:nroots]).copy()
    assert vecs.shape == (nroots, self.ngp)
    ev = np.zeros(nroots)
    st = Stats()
    rc = check(lib().pisces_solve_block(self.h, nroots, block, maxsub, max_iter, ptol,
                                        int(precond), int(precision), _p(vecs), _p(ev),
                                        ctypes.byref(st)),
               "pisces_solve_block")
    return dict(nconv=rc, evals=ev, vecs=vecs, n_matvec=st.n_matvec,
                n_matvec_lo=st.n_matvec_lo, n_iter=st.n_iter, n_macro=st.n_macro,
                n_orth_data=st.n_orth_data, seconds=st.seconds, resid=np.array(st.resid[:nroots]))


Context.solve_block = _solve_block


def default_potpara():
    """PotPara defaults for Potential 1-4 (src/GetInput.cpp:53-61)."""
    pp = np.zeros(11)
    pp[:7] = [9.745, 2.536, 4.4837, 1.7724, 1.7724, 4.4837, 0.0]
    pp[6] = 0.5 / np.sqrt(pp[0])
    pp[10] = 20.0
    return pp


def start_vectors(ctx, nroots, seed=0):
    """Half-box particle-in-a-box vector plus seeded uniform [-1,1] vectors, as
    DVR::Diagonalize prepares them for StartVector=1 (src/DVR.cpp:408-453)."""
    rng = np.random.default_rng(seed)
    S = np.zeros((nroots, ctx.ngp))
    S[0] = ctx.piab()
    for r in range(1, nroots):
        S[r] = rng.uniform(-1, 1, size=ctx.ngp)
    return S


# ---------------------------------------------------------------- launch timing
PROF_KINDS = 11


def prof_enable(on=True):
    L = lib()
    L.pisces_prof_name.restype = ctypes.c_char_p
    L.pisces_prof_read.argtypes = [ctypes.c_int, _D, ctypes.POINTER(ctypes.c_longlong), _D]
    check(L.pisces_prof_enable(1 if on else 0), "pisces_prof_enable")


def prof_reset():
    check(lib().pisces_prof_reset(), "pisces_prof_reset")


def prof_read():
    """{family: (ms_total, launches, algorithmic_bytes_total)} of the timed launches."""
    L = lib()
    L.pisces_prof_name.restype = ctypes.c_char_p
    out = {}
    for k in range(PROF_KINDS):
        ms, n, b = ctypes.c_double(), ctypes.c_longlong(), ctypes.c_double()
        check(L.pisces_prof_read(k, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(b)),
              "pisces_prof_read")
        out[L.pisces_prof_name(k).decode()] = (ms.value, n.value, b.value)
    return out


def tuning():
    """Autotuned kernel configuration ({'gram_variant', 'gram_rows'})."""
    v, r = ctypes.c_int(), ctypes.c_int()
    check(lib().pisces_tuning_gram(ctypes.byref(v), ctypes.byref(r)), "pisces_tuning_gram")
    return {"gram_variant": v.value, "gram_rows": r.value}


# ---------------------------------------------------------------- server
class Request(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int * 3),
        ("L", ctypes.c_double * 3),
        ("nw", ctypes.c_int),
        ("xyz", _D),
        ("nion", ctypes.c_int),
        ("ion_xyz", _D),
        ("ion_q", _D),
        ("nroots", ctypes.c_int),
        ("warm", ctypes.c_int),
        ("chain", ctypes.c_longlong),
        ("want_vectors", ctypes.c_int),
    ]


class Result(ctypes.Structure):
    _fields_ = [
        ("id", ctypes.c_longlong),
        ("status", ctypes.c_int),
        ("nroots", ctypes.c_int),
        ("evals", ctypes.c_double * 8),
        ("resid", ctypes.c_double * 8),
        ("n_iter", ctypes.c_int),
        ("n_matvec", ctypes.c_int),
        ("latency_s", ctypes.c_double),
    ]


def _server_lib():
    L = lib()
    if not getattr(L, "_server_typed", False):
        L.pisces_server_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int,
                                           ctypes.c_size_t, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int, ctypes.c_int]
        L.pisces_server_destroy.argtypes = [ctypes.c_void_p]
        L.pisces_server_destroy.restype = None
        L.pisces_server_submit.argtypes = [ctypes.c_void_p, ctypes.POINTER(Request)]
        L.pisces_server_submit.restype = ctypes.c_longlong
        L.pisces_server_step.argtypes = [ctypes.c_void_p]
        L.pisces_server_poll.argtypes = [ctypes.c_void_p, ctypes.POINTER(Result), ctypes.c_int]
        L.pisces_server_footprint.argtypes = [ctypes.c_void_p, ctypes.POINTER(Request)]
        L.pisces_server_footprint.restype = ctypes.c_size_t
        L.pisces_server_vectors.argtypes = [ctypes.c_void_p, ctypes.c_longlong, _D]
        L.pisces_server_pending.argtypes = [ctypes.c_void_p]
        L.pisces_server_used_bytes.argtypes = [ctypes.c_void_p]
        L.pisces_server_used_bytes.restype = ctypes.c_size_t
        L._server_typed = True
    return L


class Server:
    """Continuous-batching solver service (pisces_server_*)."""

    def __init__(self, hbm_cap_bytes=200 << 30, max_slots=64, precision=BF16X2, maxsub=32,
                 ptol=7, device=0):
        L = _server_lib()
        h = ctypes.c_void_p()
        check(L.pisces_server_create(ctypes.byref(h), int(device), int(hbm_cap_bytes),
                                     int(max_slots), int(precision), int(maxsub), int(ptol)),
              "pisces_server_create")
        self.h = h
        self.shapes = {}

    def close(self):
        if getattr(self, "h", None):
            _server_lib().pisces_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _req(self, n, L, xyz, nroots, ions=None, chain=-1, warm=False, want_vectors=False):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1)
        r = Request()
        for a in range(3):
            r.n[a] = int(n[a])
            r.L[a] = float(L[a])
        r.nw = xyz.size // 9
        r.xyz = _p(xyz)
        keep = [xyz]
        if ions is not None:
            ix = np.ascontiguousarray(ions[0], dtype=np.float64).reshape(-1)
            iq = np.ascontiguousarray(ions[1], dtype=np.float64).reshape(-1)
            r.nion = iq.size
            r.ion_xyz = _p(ix)
            r.ion_q = _p(iq)
            keep += [ix, iq]
        r.nroots = int(nroots)
        r.warm = 1 if warm else 0
        r.chain = int(chain)
        r.want_vectors = 1 if want_vectors else 0
        return r, keep

    def submit(self, n, L, xyz, nroots, ions=None, chain=-1, warm=False, want_vectors=False):
        r, keep = self._req(n, L, xyz, nroots, ions, chain, warm, want_vectors)
        rid = _server_lib().pisces_server_submit(self.h, ctypes.byref(r))
        check(int(rid), "pisces_server_submit")
        self.shapes[int(rid)] = (tuple(n), int(nroots))
        return int(rid)

    def footprint(self, n, L, xyz, nroots):
        r, keep = self._req(n, L, xyz, nroots)
        return int(_server_lib().pisces_server_footprint(self.h, ctypes.byref(r)))

    def step(self):
        return check(_server_lib().pisces_server_step(self.h), "pisces_server_step")

    def pending(self):
        return _server_lib().pisces_server_pending(self.h)

    def used_bytes(self):
        return int(_server_lib().pisces_server_used_bytes(self.h))

    def poll(self, max_results=1024):
        buf = (Result * max_results)()
        k = check(_server_lib().pisces_server_poll(self.h, buf, max_results), "pisces_server_poll")
        out = []
        for i in range(k):
            r = buf[i]
            out.append(dict(id=r.id, status=r.status, evals=np.array(r.evals[:r.nroots]),
                            resid=np.array(r.resid[:r.nroots]), n_iter=r.n_iter,
                            n_matvec=r.n_matvec, latency_s=r.latency_s))
        return out

    def vectors(self, rid):
        n, nroots = self.shapes[rid]
        v = np.zeros((nroots, n[0] * n[1] * n[2]))
        check(_server_lib().pisces_server_vectors(self.h, int(rid), _p(v)), "pisces_server_vectors")
        return v

    def run_until_done(self, max_steps=100000):
        results = []
        for _ in range(max_steps):
            live = self.step()
            results += self.poll()
            if live == 0 and self.pending() == 0:
                break
        return results
