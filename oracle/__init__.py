"""ctypes binding of the CPU oracle (liboracle.so).

Test infrastructure only: imported by tests/, by ``__graft_entry__.smoke`` and
by bench.py's CPU-baseline leg, always as the checker.  The product package
never imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
MAXL = 8


class Params(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("nt", ctypes.c_int), ("levels", ctypes.c_int),
        ("bx", ctypes.c_int * MAXL), ("bt", ctypes.c_int * MAXL), ("ntest", ctypes.c_int * MAXL),
        ("sapx", ctypes.c_int * MAXL), ("sapt", ctypes.c_int * MAXL),
        ("m0", ctypes.c_double),
        ("nu1", ctypes.c_int), ("nu2", ctypes.c_int),
        ("sap_m", ctypes.c_int), ("sap_restarts", ctypes.c_int), ("sap_gmres_tol", ctypes.c_double),
        ("sap_tol", ctypes.c_double),
        ("kc_m", ctypes.c_int), ("kc_restarts", ctypes.c_int), ("kc_tol", ctypes.c_double),
        ("crs_m", ctypes.c_int), ("crs_restarts", ctypes.c_int), ("crs_tol", ctypes.c_double),
        ("out_m", ctypes.c_int), ("out_restarts", ctypes.c_int), ("out_tol", ctypes.c_double),
        ("tv_sap_iters", ctypes.c_int),
        ("seed", ctypes.c_uint),
        ("threads", ctypes.c_int),
        ("cycle", ctypes.c_int),
        ("nit", ctypes.c_int),
    ]


def build():
    """Compile liboracle.so in place (gcc; always available where tests run)."""
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        P = ctypes.POINTER(ctypes.c_double)
        vp = ctypes.c_void_p
        L.or_default_params.argtypes = [ctypes.POINTER(Params)]
        L.or_create.argtypes = [ctypes.POINTER(Params), P]
        L.or_create.restype = vp
        L.or_destroy.argtypes = [vp]
        L.or_nsites.argtypes = [vp, ctypes.c_int]
        L.or_dof.argtypes = [vp, ctypes.c_int]
        for f in ("or_dphi", "or_ddag"):
            getattr(L, f).argtypes = [vp, P, P]
        for f in ("or_dop", "or_prolong", "or_restrict"):
            getattr(L, f).argtypes = [vp, ctypes.c_int, P, P]
        L.or_sap.argtypes = [vp, ctypes.c_int, P, P, ctypes.c_int]
        L.or_setup.argtypes = [vp, ctypes.c_int]
        L.or_setup_from_tv.argtypes = [vp, ctypes.POINTER(P), ctypes.c_int]
        L.or_get_tv.argtypes = [vp, ctypes.c_int, ctypes.c_int, P]
        L.or_get_links.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, P]
        L.or_kcycle0.argtypes = [vp, P, P]
        L.or_vcycle.argtypes = [vp, ctypes.c_int, P, P]
        L.or_bicgstab.argtypes = [vp, P, P, ctypes.c_int, ctypes.c_double, P, ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_int)]
        L.or_cg.argtypes = [vp, P, P, ctypes.c_int, ctypes.c_double]
        L.or_fgmres_sap.argtypes = [vp, P, P, P, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.or_force.argtypes = [vp, P, P]
        L.or_force.restype = ctypes.c_double
        L.or_multigrid.argtypes = [vp, P, P, ctypes.c_int, ctypes.c_double, ctypes.POINTER(ctypes.c_double)]
        L.or_solve.argtypes = [vp, P, P, P, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                               ctypes.POINTER(ctypes.c_double)]
        L.or_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_long)]
        L.or_gmres_fine.argtypes = [vp, P, P, ctypes.c_int, ctypes.c_int, ctypes.c_double]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _c(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def default_params(nx, nt, rows, m0=None, **kw):
    """rows: list of (level, bx, bt, ntest, sapx, sapt) as in parameters.dat."""
    p = Params()
    lib().or_default_params(ctypes.byref(p))
    p.nx, p.nt = nx, nt
    p.levels = len(rows) + 1
    for (l, bx, bt, nv, sx, st) in rows:
        p.bx[l], p.bt[l], p.ntest[l], p.sapx[l], p.sapt[l] = bx, bt, nv, sx, st
    if m0 is not None:
        p.m0 = m0
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class Oracle:
    """One reference-semantics hierarchy on the CPU."""

    def __init__(self, params, U):
        self.p = params
        self.U = _c(U).reshape(-1)
        h = lib().or_create(ctypes.byref(params), _p(self.U))
        if not h:
            raise ValueError("oracle: inconsistent lattice / blocking")
        self.h = h
        self.L = params.levels

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_destroy(self.h)
            self.h = None

    def nsites(self, l):
        return lib().or_nsites(self.h, l)

    def dof(self, l):
        return lib().or_dof(self.h, l)

    def size(self, l):
        return self.nsites(l) * self.dof(l)

    def _un(self, fn, x, n_out, *pre):
        x = _c(x).reshape(-1)
        out = np.zeros(n_out, np.complex128)
        fn(self.h, *pre, _p(x), _p(out))
        return out

    def dphi(self, x):
        return self._un(lib().or_dphi, x, self.size(0))

    def ddag(self, x):
        return self._un(lib().or_ddag, x, self.size(0))

    def dop(self, l, x):
        return self._un(lib().or_dop, x, self.size(l), l)

    def prolong(self, l, xc):
        return self._un(lib().or_prolong, xc, self.size(l), l)

    def restrict(self, l, xf):
        return self._un(lib().or_restrict, xf, self.size(l + 1), l)

    def sap(self, l, v, x, nu):
        v = _c(v).reshape(-1)
        x = _c(x).reshape(-1).copy()
        conv = lib().or_sap(self.h, l, _p(v), _p(x), nu)
        return x, conv

    def setup(self, draw_order=0):
        lib().or_setup(self.h, draw_order)

    def setup_from_tv(self, tvs, smooth=True):
        arrs = [_c(t).reshape(-1) for t in tvs]
        ptrs = (ctypes.POINTER(ctypes.c_double) * len(arrs))(*[_p(a) for a in arrs])
        lib().or_setup_from_tv(self.h, ptrs, 1 if smooth else 0)
        self._keep = arrs

    def tv(self, l, i):
        out = np.zeros(self.size(l), np.complex128)
        lib().or_get_tv(self.h, l, i, _p(out))
        return out

    def links(self, l, which, mu=0):
        d = self.dof(l)
        out = np.zeros(self.nsites(l) * d * d, np.complex128)
        lib().or_get_links(self.h, l, which, mu, _p(out))
        return out.reshape(self.nsites(l), d, d)

    def kcycle0(self, eta):
        return self._un(lib().or_kcycle0, eta, self.size(0))

    def bicgstab(self, b, max_iter=10000, tol=1e-10, hist_cap=100000):
        b = _c(b).reshape(-1)
        x = np.zeros_like(b)
        h = np.zeros(hist_cap)
        nh = ctypes.c_int(0)
        it = lib().or_bicgstab(self.h, _p(b), _p(x), max_iter, tol, _p(h), hist_cap, ctypes.byref(nh))
        return x, it, h[: nh.value].copy()

    def cg(self, phi, max_iter=100000, tol=1e-10):
        phi = _c(phi).reshape(-1)
        x = np.zeros_like(phi)
        it = lib().or_cg(self.h, _p(phi), _p(x), max_iter, tol)
        return x, it

    def fgmres_sap(self, b, hist_cap=4096):
        b = _c(b).reshape(-1)
        x = np.zeros_like(b)
        h = np.zeros(hist_cap)
        nh = ctypes.c_int(0)
        it = lib().or_fgmres_sap(self.h, _p(b), _p(x), _p(h), hist_cap, ctypes.byref(nh))
        return x, it, h[: nh.value].copy()

    def force(self, phi):
        """(S, dS/dtheta [ns, 2]) for S = phi^+ (D D^+)^-1 phi (needs setup)."""
        phi = _c(phi).reshape(-1)
        F = np.zeros((self.size(0) // 2, 2))
        S = lib().or_force(self.h, _p(phi), F.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return S, F

    def multigrid(self, b, x0=None, it=100, tol=1e-10):
        """Stand-alone multigrid; returns x, cycles (-it if not converged), relres."""
        b = _c(b).reshape(-1)
        x = np.zeros_like(b) if x0 is None else _c(x0).reshape(-1).copy()
        rr = ctypes.c_double(0)
        cyc = lib().or_multigrid(self.h, _p(b), _p(x), it, tol, ctypes.byref(rr))
        return x, cyc, rr.value

    def vcycle(self, l, eta, psi=None):
        """V-cycle at level l from psi (zero when omitted)."""
        eta = _c(eta).reshape(-1)
        out = np.zeros(self.size(l), np.complex128) if psi is None else _c(psi).reshape(-1).copy()
        lib().or_vcycle(self.h, l, _p(eta), _p(out))
        return out

    def solve(self, b, hist_cap=4096):
        b = _c(b).reshape(-1)
        x = np.zeros_like(b)
        hist = np.zeros(hist_cap)
        nh = ctypes.c_int(0)
        rr = ctypes.c_double(0)
        it = lib().or_solve(self.h, _p(b), _p(x), _p(hist), hist_cap, ctypes.byref(nh), ctypes.byref(rr))
        return x, it, hist[: nh.value].copy(), rr.value

    def stats(self):
        out = (ctypes.c_long * (6 * self.L))()
        lib().or_stats(self.h, out)
        return np.array(out[:]).reshape(self.L, 6)

    def gmres_fine(self, b, m=20, restarts=50, tol=1e-10):
        b = _c(b).reshape(-1)
        x = np.zeros_like(b)
        it = lib().or_gmres_fine(self.h, _p(b), _p(x), m, restarts, tol)
        return x, it
