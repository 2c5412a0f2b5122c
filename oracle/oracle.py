"""ctypes front end of the oracle (liboracle.so, built from nekbone_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the checker, and by bench.py's cpu_baseline leg.  The product path
(nekbone_amd) never imports this module.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

PRECOND_IDENTITY = 0
PRECOND_JACOBI = 1
PRECOND_HSMG = 2
MASK_REF = 0
MASK_DIRICHLET = 1

_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle liboracle`)")
        L = ctypes.CDLL(path)
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.c_int] * 4
        L.orc_free.argtypes = [ctypes.c_void_p]
        L.orc_npts.restype = ctypes.c_int64
        L.orc_npts.argtypes = [ctypes.c_void_p]
        for nm in ("orc_get_dxm1", "orc_get_c", "orc_get_g", "orc_set_g", "orc_get_dinv",
                   "orc_dssum", "orc_maskit", "orc_jacobi_diag_local"):
            getattr(L, nm).argtypes = [ctypes.c_void_p, _dp]
        L.orc_get_glo.argtypes = [ctypes.c_void_p, _i64p]
        L.orc_gs_ngroups.restype = ctypes.c_int64
        L.orc_gs_ngroups.argtypes = [ctypes.c_void_p]
        L.orc_gs_nshared.restype = ctypes.c_int64
        L.orc_gs_nshared.argtypes = [ctypes.c_void_p]
        L.orc_gs_map.argtypes = [ctypes.c_void_p, _i64p, _i32p]
        L.orc_mask_full_count.restype = ctypes.c_int64
        L.orc_mask_full_count.argtypes = [ctypes.c_void_p]
        L.orc_mask_ref_count.restype = ctypes.c_int64
        L.orc_mask_ref_count.argtypes = [ctypes.c_void_p]
        L.orc_mask_lists.argtypes = [ctypes.c_void_p, _i32p, _i32p]
        L.orc_set_mask_mode.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_reset_mask.argtypes = [ctypes.c_void_p]
        L.orc_ax_e.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
        L.orc_ax.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_glsc3.restype = ctypes.c_double
        L.orc_glsc3.argtypes = [_dp, _dp, _dp, ctypes.c_int64]
        L.orc_glsc3_exact.restype = ctypes.c_double
        L.orc_glsc3_exact.argtypes = [_dp, _dp, _dp, ctypes.c_int64]
        L.orc_set_exact_dots.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_setup_jacobi.argtypes = [ctypes.c_void_p]
        L.orc_cg.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int, _dp]
        L.orc_rand_uniform.argtypes = [_dp, ctypes.c_int64, ctypes.c_uint64]
        L.orc_rhs_from_raw.argtypes = [ctypes.c_void_p, _dp]
        L.orc_deformed_g.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_uint64]
        L.orc_zwgll.argtypes = [_dp, _dp, ctypes.c_int]
        L.orc_semhat.argtypes = [_dp, _dp, _dp, _dp, _dp, ctypes.c_int]
        L.orc_cubic.argtypes = [ctypes.POINTER(ctypes.c_int)] * 3 + [ctypes.c_int]
        L.orc_hsmg_setup.restype = ctypes.c_int
        L.orc_hsmg_setup.argtypes = [ctypes.c_void_p]
        L.orc_hsmg_solve.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_hsmg_levels.restype = ctypes.c_int
        L.orc_hsmg_levels.argtypes = [ctypes.c_void_p]
        L.orc_hsmg_set_coarse.restype = ctypes.c_int
        L.orc_hsmg_set_coarse.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int,
                                          ctypes.POINTER(ctypes.c_int32)]
        L.orc_hsmg_nh.restype = ctypes.c_int
        L.orc_hsmg_nh.argtypes = [ctypes.c_void_p, ctypes.c_int]
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(_dp)


def zwgll(np_):
    z = np.zeros(np_)
    w = np.zeros(np_)
    lib().orc_zwgll(_p(z), _p(w), np_)
    return z, w


def semhat(n):
    m = n + 1
    a, c, d = (np.zeros(m * m) for _ in range(3))
    b, z = np.zeros(m), np.zeros(m)
    lib().orc_semhat(_p(a), _p(b), _p(c), _p(d), _p(z), n)
    f = lambda v: v.reshape(m, m, order="F")
    return f(a), b, f(c), f(d), z


def cubic(np_):
    mx, my, mz = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    lib().orc_cubic(ctypes.byref(mx), ctypes.byref(my), ctypes.byref(mz), np_)
    return mx.value, my.value, mz.value


def rand_uniform(n, seed):
    f = np.empty(n)
    lib().orc_rand_uniform(_p(f), n, seed)
    return f


class Problem:
    """Oracle problem: box of nelx x nely x nelz elements, nx1 points per
    direction, set up as the reference driver does for one rank."""

    def __init__(self, nelx, nely, nelz, nx1):
        self._h = lib().orc_create(nelx, nely, nelz, nx1)
        if not self._h:
            raise ValueError("bad oracle problem size")
        self.nelx, self.nely, self.nelz, self.nx1 = nelx, nely, nelz, nx1
        self.nelt = nelx * nely * nelz
        self.n = lib().orc_npts(self._h)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_free(self._h)
            self._h = None

    def dxm1(self):
        d = np.empty(self.nx1 * self.nx1)
        lib().orc_get_dxm1(self._h, _p(d))
        return d

    def c(self):
        v = np.empty(self.n)
        lib().orc_get_c(self._h, _p(v))
        return v

    def g(self):
        v = np.empty(6 * self.n)
        lib().orc_get_g(self._h, _p(v))
        return v

    def set_g(self, g):
        g = np.ascontiguousarray(g, dtype=np.float64)
        assert g.size == 6 * self.n
        lib().orc_set_g(self._h, _p(g))

    def deformed_g(self, amp=0.1, seed=1):
        lib().orc_deformed_g(self._h, amp, seed)

    def glo(self):
        v = np.empty(self.n, dtype=np.int64)
        lib().orc_get_glo(self._h, v.ctypes.data_as(_i64p))
        return v

    def gs_map(self):
        ng = lib().orc_gs_ngroups(self._h)
        ns = lib().orc_gs_nshared(self._h)
        off = np.empty(ng + 1, dtype=np.int64)
        idx = np.empty(max(ns, 1), dtype=np.int32)
        lib().orc_gs_map(self._h, off.ctypes.data_as(_i64p), idx.ctypes.data_as(_i32p))
        return off, idx[:ns]

    def mask_lists(self):
        nf = lib().orc_mask_full_count(self._h)
        nr = lib().orc_mask_ref_count(self._h)
        full = np.empty(max(nf, 1), dtype=np.int32)
        ref = np.empty(max(nr, 1), dtype=np.int32)
        lib().orc_mask_lists(self._h, full.ctypes.data_as(_i32p), ref.ctypes.data_as(_i32p))
        return full[:nf], ref[:nr]

    def set_exact_dots(self, on):
        """cg with correctly rounded dot products (compensated sums of the
        reference's rounded terms) instead of the sequential sums."""
        lib().orc_set_exact_dots(self._h, int(on))

    def set_mask_mode(self, mode):
        lib().orc_set_mask_mode(self._h, mode)

    def reset_mask(self):
        lib().orc_reset_mask(self._h)

    def dssum(self, u):
        u = np.array(u, dtype=np.float64, copy=True)
        lib().orc_dssum(self._h, _p(u))
        return u

    def maskit(self, u):
        u = np.array(u, dtype=np.float64, copy=True)
        lib().orc_maskit(self._h, _p(u))
        return u

    def ax_e(self, e, u_e):
        u_e = np.ascontiguousarray(u_e, dtype=np.float64)
        w = np.empty(self.nx1 ** 3)
        lib().orc_ax_e(self._h, e, _p(w), _p(u_e))
        return w

    def ax(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        w = np.empty(self.n)
        lib().orc_ax(self._h, _p(w), _p(u))
        return w

    def jacobi_diag_local(self):
        v = np.empty(self.n)
        lib().orc_jacobi_diag_local(self._h, _p(v))
        return v

    def setup_jacobi(self):
        lib().orc_setup_jacobi(self._h)

    def dinv(self):
        v = np.empty(self.n)
        lib().orc_get_dinv(self._h, _p(v))
        return v

    def rhs_from_raw(self, fraw):
        f = np.array(fraw, dtype=np.float64, copy=True)
        lib().orc_rhs_from_raw(self._h, _p(f))
        return f

    def hsmg_setup(self):
        if lib().orc_hsmg_setup(self._h) != 0:
            raise ValueError("HSMG needs nx1 >= 4")

    def hsmg_set_coarse(self, ainv, cidx):
        ainv = np.ascontiguousarray(ainv, dtype=np.float64)
        cidx = np.ascontiguousarray(cidx, dtype=np.int32)
        nc = ainv.shape[0]
        if lib().orc_hsmg_set_coarse(self._h, ainv.ctypes.data_as(_dp), nc,
                                     cidx.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))) != 0:
            raise RuntimeError("orc_hsmg_set_coarse failed")

    def hsmg_levels(self):
        L = lib().orc_hsmg_levels(self._h)
        return [lib().orc_hsmg_nh(self._h, l) for l in range(1, L + 1)]

    def hsmg_solve(self, rhs):
        """z = M^-1 rhs (h1mg_solve); returns z and the rhs as the
        reference leaves it (Dirichlet points zeroed)."""
        r = np.array(rhs, dtype=np.float64, copy=True)
        z = np.zeros(self.n)
        lib().orc_hsmg_solve(self._h, _p(z), _p(r))
        return z, r

    def cg(self, f, niter, precond=PRECOND_IDENTITY):
        f = np.ascontiguousarray(f, dtype=np.float64)
        x = np.empty(self.n)
        r = np.empty(self.n)
        hist = np.empty(1 + 3 * niter)
        lib().orc_cg(self._h, _p(x), _p(f), _p(r), niter, precond, _p(hist))
        return x, r, hist


def glsc3(a, b, m):
    a, b, m = (np.ascontiguousarray(v, dtype=np.float64) for v in (a, b, m))
    return lib().orc_glsc3(_p(a), _p(b), _p(m), a.size)


def glsc3_exact(a, b, m):
    a, b, m = (np.ascontiguousarray(v, dtype=np.float64) for v in (a, b, m))
    return lib().orc_glsc3_exact(_p(a), _p(b), _p(m), a.size)
