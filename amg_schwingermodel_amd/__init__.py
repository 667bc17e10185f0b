"""MI355X-native DD-alpha-AMG for the 2-D Schwinger-model Wilson-Dirac operator.

Python access to libamgs.so (include/amgs.h) through ctypes.  The library is
the product: HIP kernels for gfx950 plus a C++ host runtime.  This module only
marshals arrays; every numerical operation runs in libamgs on the GPU (the
generators and file readers run in libamgs on the host).  There is no CPU
fallback: if libamgs.so is missing or no HIP device is present, the calls
raise.

The class interface mirrors the reference's (Fabian2598/AMG_SchwingerModel):
``FGMRES_AMG(dim1, dim2, m, restarts, tol, GConf, m0)`` runs the set-up in its
constructor (include/amg.h:101-119) and ``fgmres(phi, x0, x)`` returns the
iteration count (src/fgmres.cpp:5-88).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libamgs.so")
MAXL = 6

AMGS_OK, AMGS_NOT_CONVERGED = 0, 1
PREC_FP64, PREC_MIXED, PREC_MIXED16 = 0, 1, 2
FORCE_DDAG_D, FORCE_D_DDAG = 0, 1


class AmgsError(RuntimeError):
    pass


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
        ("tv_seed", ctypes.c_uint32),
        ("precision", ctypes.c_int),
        ("max_rhs", ctypes.c_int),
        ("cycle", ctypes.c_int),
        ("nit", ctypes.c_int),
    ]


def build(arch="gfx950"):
    """Compile libamgs.so in place with hipcc (no GPU needed)."""
    jobs = str(max(1, min(8, os.cpu_count() or 1)))
    subprocess.run(["make", "-s", "-C", _HERE, f"ARCH={arch}", "-B", "-j", jobs], check=True)


def cli_path():
    """Path of the amgs_solve command-line driver (built with the library)."""
    return os.path.join(_HERE, "amgs_solve")


_lib = None

# exported symbols and their argument types (include/amgs.h)
_P = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_V = ctypes.c_void_p
SIGNATURES = {
    "amgs_default_params": ([ctypes.POINTER(Params)], ctypes.c_int),
    "amgs_read_parameters": ([ctypes.c_char_p, ctypes.POINTER(Params)], ctypes.c_int),
    "amgs_create": ([ctypes.POINTER(Params), ctypes.POINTER(_V)], ctypes.c_int),
    "amgs_create_multi": ([ctypes.POINTER(Params), ctypes.c_int, _P, ctypes.POINTER(_V)], ctypes.c_int),
    "amgs_num_configs": ([_V], ctypes.c_int),
    "amgs_destroy": ([_V], None),
    "amgs_last_error": ([_V], ctypes.c_char_p),
    "amgs_set_gauge": ([_V, _P], ctypes.c_int),
    "amgs_setup": ([_V], ctypes.c_int),
    "amgs_setup_from_tv": ([_V, ctypes.POINTER(_P), ctypes.c_int], ctypes.c_int),
    "amgs_solve": ([_V, ctypes.c_int, _P, _P, _I, _P, _P, ctypes.c_int, _I], ctypes.c_int),
    "amgs_solve_dev": ([_V, ctypes.c_int, _V, _V, _I, _P], ctypes.c_int),
    "amgs_apply_D": ([_V, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, _P], ctypes.c_int),
    "amgs_apply_D_dev": ([_V, ctypes.c_int, ctypes.c_int, ctypes.c_int, _V, _V], ctypes.c_int),
    "amgs_apply_D_batched_dev": ([_V, ctypes.c_int, ctypes.c_int, ctypes.c_int, _V, _V], ctypes.c_int),
    "amgs_autotune_batched_D": ([_V, ctypes.c_int, ctypes.c_int, ctypes.c_int, _I, _P], ctypes.c_int),
    "amgs_apply_P": ([_V, ctypes.c_int, ctypes.c_int, _P, _P], ctypes.c_int),
    "amgs_apply_Pt": ([_V, ctypes.c_int, ctypes.c_int, _P, _P], ctypes.c_int),
    "amgs_sap": ([_V, ctypes.c_int, ctypes.c_int, _P, _P, ctypes.c_int], ctypes.c_int),
    "amgs_sap_inner": ([_V, ctypes.c_int, _P, _P, ctypes.c_int], ctypes.c_int),
    "amgs_kcycle": ([_V, ctypes.c_int, _P, _P], ctypes.c_int),
    "amgs_sap_dev_noalloc": ([_V, ctypes.c_int, ctypes.c_int, _V, _V, ctypes.c_int], ctypes.c_int),
    "amgs_get_tv": ([_V, ctypes.c_int, ctypes.c_int, _P], ctypes.c_int),
    "amgs_get_links": ([_V, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P], ctypes.c_int),
    "amgs_gauge_metropolis": ([ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_uint64, _P, _P],
                              ctypes.c_int),
    "amgs_gaussian": ([ctypes.c_int64, ctypes.c_uint64, _P], ctypes.c_int),
    "amgs_read_conf": ([ctypes.c_char_p, ctypes.c_int, ctypes.c_int, _P], ctypes.c_int),
    "amgs_write_conf": ([ctypes.c_char_p, ctypes.c_int, ctypes.c_int, _P], ctypes.c_int),
    "amgs_read_rhs": ([ctypes.c_char_p, ctypes.c_int, ctypes.c_int, _P], ctypes.c_int),
    "amgs_write_rhs": ([ctypes.c_char_p, ctypes.c_int, ctypes.c_int, _P], ctypes.c_int),
    "amgs_write_residuals": ([ctypes.c_char_p, _P, ctypes.c_int], ctypes.c_int),
    "amgs_time_kernel": ([_V, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P], ctypes.c_int),
    "amgs_last_timings": ([_V, _P], ctypes.c_int),
    "amgs_bicgstab": ([_V, ctypes.c_int, _P, _P, ctypes.c_int, ctypes.c_double, _I, _P, _P, ctypes.c_int, _I],
                      ctypes.c_int),
    "amgs_cg_normal": ([_V, ctypes.c_int, _P, _P, ctypes.c_int, ctypes.c_double, _I, _P], ctypes.c_int),
    "amgs_fgmres_sap": ([_V, ctypes.c_int, _P, _P, _I, _P, _P, ctypes.c_int, _I], ctypes.c_int),
    "amgs_gmres": ([_V, ctypes.c_int, _P, _P, _I, _P, _P, ctypes.c_int, _I], ctypes.c_int),
    "amgs_force": ([_V, ctypes.c_int, ctypes.c_int, _P, _P, _P], ctypes.c_int),
    "amgs_multigrid": ([_V, ctypes.c_int, _P, _P, ctypes.c_int, ctypes.c_double, _I, _P], ctypes.c_int),
    "amgs_set_graph": ([_V, ctypes.c_int], ctypes.c_int),
    "amgs_autotune": ([_V, ctypes.c_int, _I, _P], ctypes.c_int),
    "amgs_get_tuning": ([_V, ctypes.c_int, _I, _I], ctypes.c_int),
    "amgs_set_tuning": ([_V, ctypes.c_int, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "amgs_solve_multi": ([ctypes.c_int, ctypes.POINTER(_V), _I, ctypes.POINTER(_V), ctypes.POINTER(_V), _I, _P, _I],
                         ctypes.c_int),
    "amgs_solve_multi_staggered": ([ctypes.c_int, ctypes.POINTER(_V), _I, ctypes.POINTER(_V), ctypes.POINTER(_V), _P,
                                    _I, _P, _I], ctypes.c_int),
    "amgs_kernel_timing": ([_V, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "amgs_kernel_stats": ([_V, ctypes.c_int, ctypes.POINTER(ctypes.c_longlong), _P, _P], ctypes.c_int),
    "amgs_kernel_stats_ex": ([_V, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_longlong), _P, _P, _P],
                             ctypes.c_int),
}


def lib():
    """Load libamgs.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise AmgsError(f"{LIB_PATH} is missing: run __graft_entry__.build() (hipcc, gfx950)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_P)


def _c(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def default_params(nx, nt, rows, m0=None, **kw):
    """rows: parameters.dat rows (level, bx, bt, ntest, sapx, sapt), block COUNTS."""
    p = Params()
    lib().amgs_default_params(ctypes.byref(p))
    p.nx, p.nt = nx, nt
    p.levels = len(rows) + 1
    for (l, bx, bt, nv, sx, st) in rows:
        p.bx[l], p.bt[l], p.ntest[l], p.sapx[l], p.sapt[l] = bx, bt, nv, sx, st
    if m0 is not None:
        p.m0 = m0
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def read_parameters(path, p=None):
    p = p or Params()
    if p.nx == 0:
        lib().amgs_default_params(ctypes.byref(p))
    rc = lib().amgs_read_parameters(path.encode(), ctypes.byref(p))
    if rc != AMGS_OK:
        raise AmgsError(f"amgs_read_parameters({path}) failed: {rc}")
    return p


def gauge_metropolis(nx, nt, beta, sweeps, seed):
    """Quenched U(1) field from a cold start; returns (U [nx*nt, 2], plaquette)."""
    U = np.zeros((nx * nt, 2), np.complex128)
    plaq = ctypes.c_double(0)
    rc = lib().amgs_gauge_metropolis(nx, nt, beta, sweeps, seed, _p(U), ctypes.byref(plaq))
    if rc != AMGS_OK:
        raise AmgsError(f"amgs_gauge_metropolis failed: {rc}")
    return U, plaq.value


def gaussian(n, seed):
    out = np.zeros(n, np.complex128)
    rc = lib().amgs_gaussian(n, seed, _p(out))
    if rc != AMGS_OK:
        raise AmgsError(f"amgs_gaussian failed: {rc}")
    return out


def solve_multi(solvers, b_ptrs, x_ptrs, nrhs=None, start_ms=None):
    """Concurrent solves of independent contexts (amgs_solve_multi): solver i
    solves nrhs[i] right-hand sides from device address b_ptrs[i] into
    x_ptrs[i] (e.g. torch tensors' data_ptr()).  With start_ms, solver i's
    first launch waits start_ms[i] milliseconds (amgs_solve_multi_staggered).
    Returns (status, [iters_i], [relres_i])."""
    k = len(solvers)
    nrhs = list(nrhs) if nrhs is not None else [1] * k
    tot = sum(nrhs)
    ctxs = (_V * k)(*[s.h for s in solvers])
    nr = (ctypes.c_int * k)(*nrhs)
    bp = (_V * k)(*b_ptrs)
    xp = (_V * k)(*x_ptrs)
    it = np.zeros(tot, np.int32)
    rr = np.zeros(tot)
    st = np.zeros(k, np.int32)
    if start_ms is None:
        rc = lib().amgs_solve_multi(k, ctxs, nr, bp, xp, it.ctypes.data_as(_I), _p(rr), st.ctypes.data_as(_I))
    else:
        sm = np.ascontiguousarray(start_ms, np.float64)
        if sm.shape != (k,):
            raise ValueError("start_ms needs one offset per solver")
        rc = lib().amgs_solve_multi_staggered(k, ctxs, nr, bp, xp, _p(sm), it.ctypes.data_as(_I), _p(rr),
                                              st.ctypes.data_as(_I))
    if rc < 0:
        raise AmgsError(f"amgs_solve_multi failed: {rc} (per-context status {st.tolist()})")
    off = np.cumsum([0] + nrhs)
    return st, [it[off[i]:off[i + 1]] for i in range(k)], [rr[off[i]:off[i + 1]] for i in range(k)]


class Solver:
    """One hierarchy on the GPU (one amgs context).  With m0s (a list of bare
    masses) the context holds len(m0s) gauge configurations of one lattice
    shape (amgs_create_multi); U is then [ncfg, nx*nt, 2] and every batch has
    one right-hand side per configuration."""

    def __init__(self, params, U=None, m0s=None):
        self.p = params
        h = _V()
        if m0s is None:
            self.ncfg = 1
            rc = lib().amgs_create(ctypes.byref(params), ctypes.byref(h))
        else:
            m = np.ascontiguousarray(m0s, dtype=np.float64)
            self.ncfg = len(m)
            rc = lib().amgs_create_multi(ctypes.byref(params), len(m), _p(m), ctypes.byref(h))
        if rc != AMGS_OK:
            raise AmgsError(f"amgs_create failed ({rc})")
        self.h = h
        self.L = params.levels
        self.shapes = [(params.nx * params.nt, 2)]
        for l in range(self.L - 1):
            self.shapes.append((params.bx[l] * params.bt[l], 2 * params.ntest[l]))
        if U is not None:
            self.set_gauge(U)

    def close(self):
        if getattr(self, "h", None):
            lib().amgs_destroy(self.h)
            self.h = None

    def __del__(self):
        # at interpreter exit the module's globals may already be gone: the
        # process then releases the device memory itself
        try:
            self.close()
        except TypeError:
            pass

    def _chk(self, rc, what, ok=(AMGS_OK,)):
        if rc not in ok:
            msg = lib().amgs_last_error(self.h)
            raise AmgsError(f"{what}: {rc} {msg.decode() if msg else ''}")
        return rc

    def size(self, l):
        return self.shapes[l][0] * self.shapes[l][1]

    def set_gauge(self, U):
        self.U = _c(U).reshape(-1)
        self._chk(lib().amgs_set_gauge(self.h, _p(self.U)), "amgs_set_gauge")

    def setup(self):
        self._chk(lib().amgs_setup(self.h), "amgs_setup")

    def autotune(self, nrhs=1):
        """Pick the level-0 SAP launch shape for batches of nrhs by timing; returns
        (workgroup target, sweep ms)."""
        best = ctypes.c_int(0)
        ms = ctypes.c_double(0)
        self._chk(lib().amgs_autotune(self.h, nrhs, ctypes.byref(best), ctypes.byref(ms)), "amgs_autotune")
        return best.value, ms.value

    def tuning(self, nrhs=1):
        """(level-0 SAP workgroup target, coarse workgroups per rhs) for batches of nrhs."""
        w = ctypes.c_int(0)
        g = ctypes.c_int(0)
        self._chk(lib().amgs_get_tuning(self.h, nrhs, ctypes.byref(w), ctypes.byref(g)), "amgs_get_tuning")
        return w.value, g.value

    def set_tuning(self, nrhs, sap_wgs=-1, coarse_grp=-1):
        """Fix launch shapes for batches of nrhs (> 0 sets, 0 default, < 0 keeps)."""
        self._chk(lib().amgs_set_tuning(self.h, nrhs, sap_wgs, coarse_grp), "amgs_set_tuning")

    def set_graph(self, enable=True):
        """Replay the level-0 preconditioner from a captured HIP graph."""
        self._chk(lib().amgs_set_graph(self.h, 1 if enable else 0), "amgs_set_graph")

    def setup_from_tv(self, tvs, smooth=True):
        arrs = [_c(t).reshape(-1) for t in tvs]
        ptrs = (_P * len(arrs))(*[_p(a) for a in arrs])
        self._chk(lib().amgs_setup_from_tv(self.h, ptrs, 1 if smooth else 0), "amgs_setup_from_tv")

    def solve(self, b, hist_cap=1000):
        """b: [nrhs, n0] (or [n0]); returns x, iterations, true relres, histories."""
        b = _c(b)
        single = b.ndim == 1
        b = b.reshape(-1, self.size(0))
        nrhs = b.shape[0]
        x = np.zeros_like(b)
        it = np.zeros(nrhs, np.int32)
        rr = np.zeros(nrhs)
        hist = np.zeros((nrhs, hist_cap))
        nh = np.zeros(nrhs, np.int32)
        rc = lib().amgs_solve(self.h, nrhs, _p(b), _p(x), it.ctypes.data_as(_I), _p(rr), _p(hist), hist_cap,
                              nh.ctypes.data_as(_I))
        self._chk(rc, "amgs_solve", ok=(AMGS_OK, AMGS_NOT_CONVERGED))
        hists = [hist[r, : nh[r]].copy() for r in range(nrhs)]
        if single:
            return x[0], int(it[0]), float(rr[0]), hists[0]
        return x, it, rr, hists

    def _apply(self, fn, nin, nout, x, *args, nrhs=1):
        x = _c(x).reshape(-1)
        out = np.zeros(nout * nrhs, np.complex128)
        self._chk(fn(self.h, *args, _p(x), _p(out)), fn.__name__)
        return out

    def apply_D(self, level, x, dagger=False, nrhs=1):
        n = self.size(level)
        return self._apply(lib().amgs_apply_D, n, n, x, level, 1 if dagger else 0, nrhs, nrhs=nrhs)

    def apply_D_batched(self, level, X, precision=PREC_FP64):
        """The coarse operator on a block of vectors X [nrhs, n_l] by MFMA
        block-GEMMs (device round trip through torch); returns [nrhs, n_l]."""
        import torch

        X = _c(X).reshape(-1, self.size(level))
        nrhs = X.shape[0]
        if precision == PREC_FP64:
            xd = torch.from_numpy(X.view(np.float64).copy()).to("cuda:0")
        else:
            xd = torch.from_numpy(X.astype(np.complex64).view(np.float32).copy()).to("cuda:0")
        od = torch.zeros_like(xd)
        torch.cuda.synchronize()
        self._chk(lib().amgs_apply_D_batched_dev(self.h, level, nrhs, precision, ctypes.c_void_p(xd.data_ptr()),
                                                 ctypes.c_void_p(od.data_ptr())), "amgs_apply_D_batched_dev")
        o = od.cpu().numpy()
        o = o.view(np.complex128) if precision == PREC_FP64 else o.view(np.complex64).astype(np.complex128)
        return o.reshape(nrhs, -1)

    def autotune_batched_D(self, level, nrhs, precision=PREC_FP64):
        """(wavefronts per workgroup, microseconds per apply) chosen for this level and batch size."""
        w = ctypes.c_int(0)
        us = ctypes.c_double(0)
        self._chk(lib().amgs_autotune_batched_D(self.h, level, nrhs, precision, ctypes.byref(w), ctypes.byref(us)),
                  "amgs_autotune_batched_D")
        return w.value, us.value

    def apply_P(self, level, xc, nrhs=1):
        return self._apply(lib().amgs_apply_P, self.size(level + 1), self.size(level), xc, level, nrhs, nrhs=nrhs)

    def apply_Pt(self, level, xf, nrhs=1):
        return self._apply(lib().amgs_apply_Pt, self.size(level), self.size(level + 1), xf, level, nrhs, nrhs=nrhs)

    def sap(self, level, v, x, nu, nrhs=1):
        v = _c(v).reshape(-1)
        x = _c(x).reshape(-1).copy()
        self._chk(lib().amgs_sap(self.h, level, nrhs, _p(v), _p(x), nu), "amgs_sap")
        return x

    def sap_inner(self, v, x, nu, nrhs=1):
        """Level-0 SAP in the inner precision of a mixed context (FP32, FP16
        Krylov basis under PREC_MIXED16), in the launch form the tuning selects."""
        v = _c(v).reshape(-1)
        x = _c(x).reshape(-1).copy()
        self._chk(lib().amgs_sap_inner(self.h, nrhs, _p(v), _p(x), nu), "amgs_sap_inner")
        return x

    def kcycle(self, eta, nrhs=1):
        return self._apply(lib().amgs_kcycle, self.size(0), self.size(0), eta, nrhs, nrhs=nrhs)

    def _krylov(self, fn, b, *args, hist_cap=0, extra_hist=True):
        b = _c(b)
        single = b.ndim == 1
        b = b.reshape(-1, self.size(0))
        nrhs = b.shape[0]
        x = np.zeros_like(b)
        it = np.zeros(nrhs, np.int32)
        rr = np.zeros(nrhs)
        hist = np.zeros((nrhs, max(hist_cap, 1)))
        nh = np.zeros(nrhs, np.int32)
        tail = (_p(hist), hist_cap, nh.ctypes.data_as(_I)) if extra_hist else ()
        rc = fn(self.h, nrhs, _p(b), _p(x), *args, it.ctypes.data_as(_I), _p(rr), *tail)
        self._chk(rc, fn.__name__, ok=(AMGS_OK, AMGS_NOT_CONVERGED))
        hists = [hist[r, : nh[r]].copy() for r in range(nrhs)]
        if single:
            return x[0], int(it[0]), float(rr[0]), hists[0]
        return x, it, rr, hists

    def bicgstab(self, b, max_iter=10000, tol=1e-10, hist_cap=10000):
        """Bi-CGstab on D from x0 = 0 (src/bi_cgstab.cpp:5-87): x, iterations, relres, |s| history."""
        return self._krylov(lib().amgs_bicgstab, b, max_iter, tol, hist_cap=hist_cap)

    def cg_normal(self, phi, max_iter=100000, tol=1e-10):
        """CG on D D^dagger from x0 = phi (src/conjugate_gradient.cpp:5-65): x, iterations, relres."""
        return self._krylov(lib().amgs_cg_normal, phi, max_iter, tol, extra_hist=False)[:3]

    def fgmres_sap(self, b, hist_cap=1000):
        """FGMRES with one SAP sweep as preconditioner (include/sap.h:194-225)."""
        return self._krylov(lib().amgs_fgmres_sap, b, hist_cap=hist_cap)

    def gmres(self, b, hist_cap=1000):
        """GMRES without preconditioning (include/fgmres.h:126-148)."""
        return self._krylov(lib().amgs_gmres, b, hist_cap=hist_cap)

    def force(self, phi, form="DdagD"):
        """Pseudofermion action and dS/dtheta [.., ns, 2] by the adjoint method
        (two solves): form "DdagD" is S = phi^+ (D^+ D)^-1 phi, "DDdag" is
        S = phi^+ (D D^+)^-1 phi."""
        forms = {"DdagD": FORCE_DDAG_D, "DDdag": FORCE_D_DDAG}
        if form not in forms:
            raise ValueError(f"form must be one of {sorted(forms)}")
        phi = _c(phi)
        single = phi.ndim == 1
        phi = phi.reshape(-1, self.size(0))
        nrhs = phi.shape[0]
        F = np.zeros((nrhs, self.size(0) // 2, 2))
        S = np.zeros(nrhs)
        self._chk(lib().amgs_force(self.h, nrhs, forms[form], _p(phi), F.ctypes.data_as(_P), _p(S)), "amgs_force",
                  ok=(AMGS_OK, AMGS_NOT_CONVERGED))
        return (float(S[0]), F[0]) if single else (S, F)

    def multigrid(self, b, x0=None, max_cycles=100, tol=1e-10):
        """Stand-alone multigrid (src/amg.cpp:181-231); returns x, cycles, relres."""
        b = _c(b)
        single = b.ndim == 1
        b = b.reshape(-1, self.size(0))
        nrhs = b.shape[0]
        x = np.zeros_like(b) if x0 is None else _c(x0).reshape(nrhs, -1).copy()
        cyc = np.zeros(nrhs, np.int32)
        rr = np.zeros(nrhs)
        rc = lib().amgs_multigrid(self.h, nrhs, _p(b), _p(x), max_cycles, tol, cyc.ctypes.data_as(_I), _p(rr))
        self._chk(rc, "amgs_multigrid", ok=(AMGS_OK, AMGS_NOT_CONVERGED))
        return (x[0], int(cyc[0]), float(rr[0])) if single else (x, cyc, rr)

    def tv(self, level, i):
        out = np.zeros(self.size(level), np.complex128)
        self._chk(lib().amgs_get_tv(self.h, level, i, _p(out)), "amgs_get_tv")
        return out

    def links(self, level, which, mu=0):
        ns, d = self.shapes[level]
        out = np.zeros(ns * d * d, np.complex128)
        self._chk(lib().amgs_get_links(self.h, level, which, mu, _p(out)), "amgs_get_links")
        return out.reshape(ns, d, d)

    def time_kernel(self, kernel, nrhs=1, reps=50):
        ms = ctypes.c_double(0)
        self._chk(lib().amgs_time_kernel(self.h, kernel, nrhs, reps, ctypes.byref(ms)), "amgs_time_kernel")
        return ms.value

    def kernel_timing(self, kclass, enable=True):
        """Record every launch of a kernel class (0: SAP Arnoldi step A,
        1: step B) in later solves with HIP events on the solver's stream."""
        self._chk(lib().amgs_kernel_timing(self.h, kclass, int(enable)), "amgs_kernel_timing")

    def kernel_stats(self, kclass):
        """(launches, total device ms, algorithmic bytes) recorded so far."""
        n = ctypes.c_longlong(0)
        ms = ctypes.c_double(0)
        b = ctypes.c_double(0)
        self._chk(lib().amgs_kernel_stats(self.h, kclass, ctypes.byref(n), ctypes.byref(ms), ctypes.byref(b)),
                  "amgs_kernel_stats")
        return n.value, ms.value, b.value

    def kernel_stats_ex(self, kclass, tag=-1):
        """(launches, device ms, algorithmic bytes, flops) recorded so far; tag >= 0
        selects the launches with that tag (SAP step A: the Arnoldi index J)."""
        n = ctypes.c_longlong(0)
        ms = ctypes.c_double(0)
        b = ctypes.c_double(0)
        f = ctypes.c_double(0)
        self._chk(lib().amgs_kernel_stats_ex(self.h, kclass, tag, ctypes.byref(n), ctypes.byref(ms), ctypes.byref(b),
                                             ctypes.byref(f)), "amgs_kernel_stats_ex")
        return n.value, ms.value, b.value, f.value

    def last_timings(self):
        a = np.zeros(3)
        lib().amgs_last_timings(self.h, _p(a))
        return a


class GaugeConf:
    """Gauge field holder (include/gauge_conf.h:12-49): Conf[n][mu], n = x*Nt + t."""

    def __init__(self, Nspace, Ntime):
        self.Nx, self.Nt = Nspace, Ntime
        self.Conf = np.ones((Nspace * Ntime, 2), np.complex128)

    def read_conf(self, name):
        rc = lib().amgs_read_conf(name.encode(), self.Nx, self.Nt, _p(self.Conf))
        if rc != AMGS_OK:
            raise AmgsError(f"read_conf({name}) failed: {rc}")

    def setGconf(self, conf):
        self.Conf = _c(conf).reshape(self.Nx * self.Nt, 2).copy()


class FGMRES_AMG:
    """FGMRES with the K-cycle preconditioner (include/amg.h:99-145).

    The constructor builds the hierarchy (set-up phase) on the GPU;
    ``fgmres(phi, x0, x)`` solves D x = phi from x0 = 0 and returns the
    reference's iteration count.  ``params`` carries parameters.dat's rows.
    """

    def __init__(self, dim1, dim2, m, restarts, tol, GConf, m0, params, precision=PREC_FP64):
        p = params
        p.out_m, p.out_restarts, p.out_tol, p.m0 = m, restarts, tol, m0
        p.precision = precision
        assert dim1 == GConf.Nx * GConf.Nt and dim2 == 2
        self.solver = Solver(p, GConf.Conf)
        self.solver.setup()
        self.residuals = []

    def fgmres(self, phi, x0, x, print_message=False, save_res=False):
        if np.any(np.asarray(x0) != 0):
            raise AmgsError("the GPU solve starts from x0 = 0 as the reference's driver does (src/tests.cpp:69)")
        sol, it, rr, hist = self.solver.solve(phi)
        x[...] = sol.reshape(np.shape(x))
        self.residuals = list(hist)
        if print_message:
            print(f"FGMRES finished with {it} iterations, true relative residual {rr:.3e}")
        return it
