"""ctypes binding of libnekbone_amd.so (the C ABI in include/nekbone_amd.h).

The product path: every compute call goes to the HIP library.  There is no
CPU fallback; if the library is not built, or no GPU is present, calls
raise NekboneError.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libnekbone_amd.so")

PRECOND_IDENTITY = 0
PRECOND_JACOBI = 1
PRECOND_HSMG = 2
MASK_REF = 0
MASK_DIRICHLET = 1
ARITH_EXACT = 0
ARITH_FMA = 1
MIXED_F32, MIXED_G16, MIXED_F16_MFMA = 0, 1, 2

KCLASS_AX, KCLASS_GS, KCLASS_POST, KCLASS_UPDATE, KCLASS_REDUCE = range(5)

_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)
_fp = ctypes.POINTER(ctypes.c_float)
_i32p = ctypes.POINTER(ctypes.c_int32)
_vp = ctypes.c_void_p

# name -> (restype, argtypes)
_SIGS = {
    "nekbone_last_error": (ctypes.c_char_p, []),
    "nekbone_init": (ctypes.c_int, [ctypes.c_int]),
    "nekbone_finalize": (ctypes.c_int, []),
    "nekbone_sync": (ctypes.c_int, []),
    "nekbone_problem_create_box": (ctypes.c_int, [ctypes.POINTER(_vp)] + [ctypes.c_int] * 4),
    "nekbone_problem_create": (ctypes.c_int, [ctypes.POINTER(_vp)] + [ctypes.c_int] * 5 +
                               [_i32p, _dp, _i64p, _dp, _dp]),
    "nekbone_problem_destroy": (ctypes.c_int, [_vp]),
    "nekbone_problem_npts": (ctypes.c_int64, [_vp]),
    "nekbone_problem_nx1": (ctypes.c_int, [_vp]),
    "nekbone_problem_nelt": (ctypes.c_int, [_vp]),
    "nekbone_set_geometry": (ctypes.c_int, [_vp, _dp]),
    "nekbone_set_geometry_deformed": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_uint64]),
    "nekbone_get_geometry": (ctypes.c_int, [_vp, _dp]),
    "nekbone_get_dxm1": (ctypes.c_int, [_vp, _dp]),
    "nekbone_get_c": (ctypes.c_int, [_vp, _dp]),
    "nekbone_get_glo": (ctypes.c_int, [_vp, _i64p]),
    "nekbone_gs_ngroups": (ctypes.c_int64, [_vp]),
    "nekbone_gs_nshared": (ctypes.c_int64, [_vp]),
    "nekbone_set_mask_mode": (ctypes.c_int, [_vp, ctypes.c_int]),
    "nekbone_reset_mask": (ctypes.c_int, [_vp]),
    "nekbone_set_arith": (ctypes.c_int, [_vp, ctypes.c_int]),
    "nekbone_set_fused": (ctypes.c_int, [_vp, ctypes.c_int]),
    "nekbone_set_gather": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    "nekbone_set_xring": (ctypes.c_int, [_vp, ctypes.c_int]),
    "nekbone_get_cg_form": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int),
                                           ctypes.POINTER(ctypes.c_int)]),
    "nekbone_setup_jacobi": (ctypes.c_int, [_vp]),
    "nekbone_setup_hsmg": (ctypes.c_int, [_vp]),
    "nekbone_hsmg_solve": (ctypes.c_int, [_vp, _dp, _dp]),
    "nekbone_set_hsmg_fp32": (ctypes.c_int, [_vp, ctypes.c_int]),
    "nekbone_set_hsmg_coarse": (ctypes.c_int, [_vp, ctypes.c_int]),
    "nekbone_set_mixed_inner": (ctypes.c_int, [_vp, ctypes.c_int]),
    "nekbone_set_mixed_refresh": (ctypes.c_int, [_vp, ctypes.c_int]),
    "nekbone_batch_set_graph": (ctypes.c_int, [_vp, ctypes.c_int]),
    "nekbone_get_dinv": (ctypes.c_int, [_vp, _dp]),
    "nekbone_ax": (ctypes.c_int, [_vp, _dp, _dp]),
    "nekbone_ax_d": (ctypes.c_int, [_vp, _vp, _vp]),
    "nekbone_ax_local": (ctypes.c_int, [_vp, _dp, _dp]),
    "nekbone_dssum": (ctypes.c_int, [_vp, _dp]),
    "nekbone_dssum_d": (ctypes.c_int, [_vp, _vp]),
    "nekbone_maskit": (ctypes.c_int, [_vp, _dp]),
    "nekbone_glsc3": (ctypes.c_int, [_vp, _dp, _dp, _dp, _dp]),
    "nekbone_cg": (ctypes.c_int, [_vp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int, _dp]),
    "nekbone_cg_d": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_int, _dp]),
    "nekbone_cg_tol_d": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_double, ctypes.c_int,
                                        ctypes.c_int, ctypes.POINTER(ctypes.c_int), _dp]),
    "nekbone_cg_mixed_d": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_double, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_double,
                                          ctypes.POINTER(ctypes.c_int),
                                          ctypes.POINTER(ctypes.c_int), _dp]),
    "nekbone_ax_local_f32": (ctypes.c_int, [_vp, _fp, _fp]),
    "nekbone_get_variant_f32": (ctypes.c_int, [_vp] + [ctypes.POINTER(ctypes.c_int)] * 3),
    "nekbone_set_variant_f32": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "nekbone_autotune_mixed": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "nekbone_set_mixed_g16": (ctypes.c_int, [_vp, ctypes.c_int]),
    "nekbone_alloc_vec": (ctypes.c_int, [_vp, ctypes.POINTER(_vp)]),
    "nekbone_free_vec": (ctypes.c_int, [_vp, _vp]),
    "nekbone_h2d": (ctypes.c_int, [_vp, _vp, _dp, ctypes.c_int64]),
    "nekbone_d2h": (ctypes.c_int, [_vp, _dp, _vp, ctypes.c_int64]),
    "nekbone_rand_uniform_d": (ctypes.c_int, [_vp, _vp, ctypes.c_int64, ctypes.c_uint64]),
    "nekbone_rhs_from_raw_d": (ctypes.c_int, [_vp, _vp]),
    "nekbone_set_f_raw": (ctypes.c_int, [ctypes.c_int64, _dp]),
    "nekbone_set_f_d": (ctypes.c_int, [_vp, _vp]),
    "nekbone_timing_enable": (ctypes.c_int, [_vp, ctypes.c_int]),
    "nekbone_timing_get": (ctypes.c_int, [_vp, ctypes.c_int, _i64p, _dp]),
    "nekbone_timing_reset": (ctypes.c_int, [_vp]),
    "nekbone_autotune": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "nekbone_set_graph": (ctypes.c_int, [_vp, ctypes.c_int]),
    "nekbone_get_variant": (ctypes.c_int, [_vp] + [ctypes.POINTER(ctypes.c_int)] * 3),
    "nekbone_set_variant": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "nekbone_bw_probe": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int, _dp]),
    "nekbone_cubic": (ctypes.c_int, [ctypes.c_int] + [ctypes.POINTER(ctypes.c_int)] * 3),
    "nekbone_batch_create": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                            ctypes.c_int]),
    "nekbone_batch_destroy": (ctypes.c_int, [_vp]),
    "nekbone_batch_size": (ctypes.c_int, [_vp]),
    "nekbone_batch_set_xu": (ctypes.c_int, [_vp, ctypes.c_int]),
    "nekbone_batch_autotune": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "nekbone_batch_set_queue": (ctypes.c_int, [_vp, ctypes.c_int]),
    "nekbone_batch_get_path": (ctypes.c_int, [_vp] + [ctypes.POINTER(ctypes.c_int)] * 3),
    "nekbone_batch_get_form": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int)]),
    "nekbone_batch_cg_tol_d": (ctypes.c_int, [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                              ctypes.POINTER(_vp), ctypes.c_double, ctypes.c_int,
                                              ctypes.c_int, ctypes.POINTER(ctypes.c_int), _dp]),
    "nekbone_glsc3_noalloc_d": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int64, _vp]),
    "nekbone_tuning_save": (ctypes.c_int, [ctypes.c_char_p]),
    "nekbone_tuning_load": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int)]),
    "glsc3_": (ctypes.c_double, [_dp, _dp, _dp, ctypes.POINTER(ctypes.c_int)]),
}

_LIB = None


class NekboneError(RuntimeError):
    pass


class NotConverged(NekboneError):
    """A tolerance solve stopped at maxiter without meeting tol
    (NEKBONE_NOT_CONVERGED); its outputs were written all the same."""
    pass


NOT_CONVERGED = 6


def lib():
    """Load the HIP library (fails loudly if it is not built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise NekboneError("libnekbone_amd.so is not built: run `python -m nekbone_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name, None)
            if f is None:
                continue
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _chk(rc, maxiter_ok=False):
    """Raise on a non-zero status; maxiter_ok accepts NEKBONE_NOT_CONVERGED
    (a tolerance solve that ran to maxiter on purpose)."""
    if rc == 0 or (maxiter_ok and rc == NOT_CONVERGED):
        return
    msg = lib().nekbone_last_error().decode(errors="replace")
    cls = NotConverged if rc == NOT_CONVERGED else NekboneError
    raise cls("nekbone error %d: %s" % (rc, msg))


def _p(a):
    return a.ctypes.data_as(_dp)


def _f64(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if n is not None and a.size != n:
        raise ValueError("expected %d values, got %d" % (n, a.size))
    return a


def init(device=0):
    _chk(lib().nekbone_init(device))


def sync():
    _chk(lib().nekbone_sync())


def glsc3_noalloc_d(a, b, mult, n, result):
    """sum a*b*mult over n points of device memory into result[0] (device),
    without allocating (nekbone_glsc3_noalloc_d); a, b, mult, result are
    DeviceVec objects or raw device addresses."""
    ptr = lambda v: v.ptr if isinstance(v, DeviceVec) else v
    _chk(lib().nekbone_glsc3_noalloc_d(ptr(a), ptr(b), ptr(mult), int(n), ptr(result)))


def tuning_save(path):
    """Write the element-kernel shapes the autotuner found (JSON)."""
    _chk(lib().nekbone_tuning_save(os.fsencode(path)))


def tuning_load(path):
    """Read a tuning file; returns the number of entries taken."""
    n = ctypes.c_int()
    _chk(lib().nekbone_tuning_load(os.fsencode(path), ctypes.byref(n)))
    return n.value


def cubic(np_):
    """Box factorization of np_ elements (cubic, src/driver.f:196-247)."""
    m = [ctypes.c_int() for _ in range(3)]
    _chk(lib().nekbone_cubic(int(np_), *[ctypes.byref(v) for v in m]))
    return tuple(v.value for v in m)


class Batch:
    """Problems solved together, one launch per kernel per order group."""

    def __init__(self, problems):
        self.problems = list(problems)
        arr = (_vp * len(self.problems))(*[p.handle for p in self.problems])
        h = _vp()
        _chk(lib().nekbone_batch_create(ctypes.byref(h), arr, len(self.problems)))
        self.handle = h.value

    def close(self):
        if getattr(self, "handle", None):
            lib().nekbone_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_xu(self, xu):
        _chk(lib().nekbone_batch_set_xu(self.handle, int(xu)))

    def set_graph(self, on):
        _chk(lib().nekbone_batch_set_graph(self.handle, int(bool(on))))

    def set_queue(self, tileset):
        """-1: one element-kernel launch per order group; k >= 0: the
        work-queue kernel with compiled tile set k."""
        _chk(lib().nekbone_batch_set_queue(self.handle, int(tileset)))

    def path(self):
        """(tile set or -1, x update in the element kernel, number of tile
        sets compiled)."""
        v = [ctypes.c_int() for _ in range(3)]
        _chk(lib().nekbone_batch_get_path(self.handle, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def gather_form(self):
        """Whether the last batched solve ran in the gather form."""
        g = ctypes.c_int(0)
        _chk(lib().nekbone_batch_get_form(self.handle, ctypes.byref(g)))
        return bool(g.value)

    def autotune(self, precond=None):
        precond = PRECOND_JACOBI if precond is None else precond
        ms = ctypes.c_double()
        _chk(lib().nekbone_batch_autotune(self.handle, precond, ctypes.byref(ms)))
        return ms.value

    def cg_tol_d(self, xs, fs, rs, tol, maxiter, precond=None, maxiter_ok=False):
        precond = PRECOND_JACOBI if precond is None else precond
        n = len(self.problems)
        ptrs = lambda vs: (_vp * n)(*[(v.ptr if v is not None else None) for v in vs])
        it = (ctypes.c_int * n)()
        rtr = (ctypes.c_double * n)()
        _chk(lib().nekbone_batch_cg_tol_d(self.handle, ptrs(xs), ptrs(fs), ptrs(rs), tol, maxiter,
                                          precond, it, rtr), maxiter_ok)
        return list(it), list(rtr)


def set_f_raw(n):
    """The reference's raw right-hand side (set_f, src/driver.f:88-95)
    for points 1..n, computed on the host."""
    raw = np.empty(int(n))
    _chk(lib().nekbone_set_f_raw(int(n), _p(raw)))
    return raw


def bw_probe(nbytes, reps=10):
    """Device copy of nbytes (half read, half written); returns ms/copy."""
    ms = ctypes.c_double()
    _chk(lib().nekbone_bw_probe(int(nbytes), int(reps), ctypes.byref(ms)))
    return ms.value


class DeviceVec:
    """n doubles of device memory owned by a problem."""

    def __init__(self, problem):
        self.problem = problem
        self.n = problem.n
        h = _vp()
        _chk(lib().nekbone_alloc_vec(problem.handle, ctypes.byref(h)))
        self.ptr = h.value

    def upload(self, a):
        a = _f64(a, self.n)
        _chk(lib().nekbone_h2d(self.problem.handle, self.ptr, _p(a), self.n))

    def download(self):
        out = np.empty(self.n)
        _chk(lib().nekbone_d2h(self.problem.handle, _p(out), self.ptr, self.n))
        return out

    def free(self):
        if self.ptr:
            _chk(lib().nekbone_free_vec(self.problem.handle, self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Problem:
    """One rank's share of a Nekbone box mesh, resident on the GPU.

    Mirrors the reference's set-up for one rank (src/driver.f:117-194 and
    the routines it calls).  Construct with a box size, or with
    ``Problem.from_arrays`` to take the reference's own set-up arrays.
    """

    def __init__(self, nelx, nely, nelz, nx1, _handle=None):
        self.handle = _handle
        if self.handle is None:
            h = _vp()
            _chk(lib().nekbone_problem_create_box(ctypes.byref(h), nelx, nely, nelz, nx1))
            self.handle = h.value
        self.nelx, self.nely, self.nelz, self.nx1 = nelx, nely, nelz, nx1
        self.nelt = lib().nekbone_problem_nelt(self.handle)
        self.n = lib().nekbone_problem_npts(self.handle)

    @classmethod
    def from_arrays(cls, nelt, nx1, nelx, nely, nelz, lglel, dxm1, glo_num, gxyz=None, c=None):
        lglel = np.ascontiguousarray(lglel, dtype=np.int32)
        dxm1 = _f64(dxm1, nx1 * nx1)
        glo = np.ascontiguousarray(glo_num, dtype=np.int64)
        g = None if gxyz is None else _f64(gxyz)
        cc = None if c is None else _f64(c)
        h = _vp()
        _chk(lib().nekbone_problem_create(
            ctypes.byref(h), nelt, nx1, nelx, nely, nelz, lglel.ctypes.data_as(_i32p), _p(dxm1),
            glo.ctypes.data_as(_i64p), None if g is None else _p(g), None if cc is None else _p(cc)))
        return cls(nelx, nely, nelz, nx1, _handle=h.value)

    def close(self):
        if self.handle:
            _chk(lib().nekbone_problem_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # set-up data -------------------------------------------------------
    def dxm1(self):
        d = np.empty(self.nx1 * self.nx1)
        _chk(lib().nekbone_get_dxm1(self.handle, _p(d)))
        return d

    def c(self):
        v = np.empty(self.n)
        _chk(lib().nekbone_get_c(self.handle, _p(v)))
        return v

    def glo(self):
        v = np.empty(self.n, dtype=np.int64)
        _chk(lib().nekbone_get_glo(self.handle, v.ctypes.data_as(_i64p)))
        return v

    def gs_counts(self):
        return (lib().nekbone_gs_ngroups(self.handle), lib().nekbone_gs_nshared(self.handle))

    def geometry(self):
        g = np.empty(6 * self.n)
        _chk(lib().nekbone_get_geometry(self.handle, _p(g)))
        return g

    def set_geometry(self, g):
        g = _f64(g, 6 * self.n)
        _chk(lib().nekbone_set_geometry(self.handle, _p(g)))

    def set_geometry_deformed(self, amp=0.1, seed=1):
        _chk(lib().nekbone_set_geometry_deformed(self.handle, amp, seed))

    def set_mask_mode(self, mode):
        _chk(lib().nekbone_set_mask_mode(self.handle, mode))

    def reset_mask(self):
        _chk(lib().nekbone_reset_mask(self.handle))

    def set_arith(self, arith):
        _chk(lib().nekbone_set_arith(self.handle, arith))

    def set_graph(self, on):
        _chk(lib().nekbone_set_graph(self.handle, int(on)))

    def set_fused(self, on):
        _chk(lib().nekbone_set_fused(self.handle, int(bool(on))))

    def set_xring(self, on):
        _chk(lib().nekbone_set_xring(self.handle, int(bool(on))))

    def cg_form(self):
        """(gather form, x-ring) of the last FP64 solve."""
        g, x = ctypes.c_int(0), ctypes.c_int(0)
        _chk(lib().nekbone_get_cg_form(self.handle, ctypes.byref(g), ctypes.byref(x)))
        return bool(g.value), bool(x.value)

    def set_gather(self, on=None):
        """Gather form of the fused CG (True/False; None keeps the setting).
        Returns whether it applies to the next solve."""
        act = ctypes.c_int(0)
        _chk(lib().nekbone_set_gather(self.handle, -1 if on is None else int(bool(on)),
                                      ctypes.byref(act)))
        return bool(act.value)

    def setup_jacobi(self):
        _chk(lib().nekbone_setup_jacobi(self.handle))

    def setup_hsmg(self):
        _chk(lib().nekbone_setup_hsmg(self.handle))

    def set_hsmg_fp32(self, on):
        """FP32 local FDM solves in the HSMG smoother (FP64 otherwise)."""
        _chk(lib().nekbone_set_hsmg_fp32(self.handle, int(bool(on))))

    def set_hsmg_coarse(self, mode):
        """HSMG coarse level: 0 the reference's (sum of the restricted
        residual), 1 an FP64 direct solve of the order-1 operator."""
        _chk(lib().nekbone_set_hsmg_coarse(self.handle, int(mode)))

    def hsmg_solve(self, r):
        """z = M^-1 r (h1mg_solve); returns z and r as the preconditioner
        leaves it (Dirichlet points zeroed)."""
        r = _f64(np.array(r, dtype=np.float64, copy=True), self.n)
        z = np.empty(self.n)
        _chk(lib().nekbone_hsmg_solve(self.handle, _p(z), _p(r)))
        return z, r

    def dinv(self):
        v = np.empty(self.n)
        _chk(lib().nekbone_get_dinv(self.handle, _p(v)))
        return v

    # operators (host arrays) -------------------------------------------
    def ax(self, u):
        u = _f64(u, self.n)
        w = np.empty(self.n)
        _chk(lib().nekbone_ax(self.handle, _p(w), _p(u)))
        return w

    def ax_local(self, u):
        u = _f64(u, self.n)
        w = np.empty(self.n)
        _chk(lib().nekbone_ax_local(self.handle, _p(w), _p(u)))
        return w

    def dssum(self, u):
        u = np.array(u, dtype=np.float64, copy=True)
        _chk(lib().nekbone_dssum(self.handle, _p(u)))
        return u

    def maskit(self, u):
        u = np.array(u, dtype=np.float64, copy=True)
        _chk(lib().nekbone_maskit(self.handle, _p(u)))
        return u

    def glsc3(self, a, b, m):
        a, b, m = (_f64(v, self.n) for v in (a, b, m))
        out = ctypes.c_double()
        _chk(lib().nekbone_glsc3(self.handle, _p(a), _p(b), _p(m), ctypes.byref(out)))
        return out.value

    def cg(self, f, niter, precond=PRECOND_IDENTITY):
        """PCG from x=0 (src/cg.f:2-84).  Returns x, r and the glsc3
        history (rnorm^2, then rtz1, pap, rtr per iteration)."""
        f = _f64(f, self.n)
        x = np.empty(self.n)
        r = np.empty(self.n)
        hist = np.empty(1 + 3 * niter)
        _chk(lib().nekbone_cg(self.handle, _p(x), _p(f), _p(r), niter, precond, _p(hist)))
        return x, r, hist

    # device-resident operations ------------------------------------------
    def vec(self):
        return DeviceVec(self)

    def rand_uniform_d(self, dv, seed):
        _chk(lib().nekbone_rand_uniform_d(self.handle, dv.ptr, dv.n, seed))

    def rhs_from_raw_d(self, dv):
        _chk(lib().nekbone_rhs_from_raw_d(self.handle, dv.ptr))

    def set_f_d(self, dv):
        """dv = set_f's right-hand side (src/driver.f:88-101)."""
        _chk(lib().nekbone_set_f_d(self.handle, dv.ptr))

    def cg_d(self, x, f, r, niter, precond=PRECOND_IDENTITY, want_hist=False):
        hist = np.empty(1 + 3 * niter) if want_hist else None
        _chk(lib().nekbone_cg_d(self.handle, x.ptr, f.ptr, r.ptr, niter, precond,
                                None if hist is None else _p(hist)))
        return hist

    def cg_tol_d(self, x, f, r, tol, maxiter, precond=PRECOND_IDENTITY, maxiter_ok=False):
        """CG to a relative residual on device vectors; returns (iterations,
        final rtr).  Raises NotConverged at maxiter unless maxiter_ok."""
        it = ctypes.c_int()
        rtr = ctypes.c_double()
        _chk(lib().nekbone_cg_tol_d(self.handle, x.ptr, f.ptr, r.ptr, tol, maxiter, precond,
                                    ctypes.byref(it), ctypes.byref(rtr)), maxiter_ok)
        return it.value, rtr.value

    def cg_mixed_d(self, x, f, r, tol, maxiter, precond=PRECOND_JACOBI, delta=0.0,
                   maxiter_ok=False):
        """Mixed-precision CG to a relative residual on device vectors;
        returns (inner iterations, FP64 residual updates, final rtr).
        Raises NotConverged at maxiter unless maxiter_ok."""
        it, nu = ctypes.c_int(), ctypes.c_int()
        rtr = ctypes.c_double()
        _chk(lib().nekbone_cg_mixed_d(self.handle, x.ptr, f.ptr, r.ptr, tol, maxiter, precond,
                                      delta, ctypes.byref(it), ctypes.byref(nu),
                                      ctypes.byref(rtr)), maxiter_ok)
        return it.value, nu.value, rtr.value

    def cg_mixed(self, f, tol, maxiter, precond=PRECOND_JACOBI, delta=0.0, maxiter_ok=False):
        """Host-array form: returns x, r (reference order), iterations,
        updates, final rtr."""
        fv, xv, rv = self.vec(), self.vec(), self.vec()
        try:
            fv.upload(f)
            it, nu, rtr = self.cg_mixed_d(xv, fv, rv, tol, maxiter, precond, delta, maxiter_ok)
            return xv.download(), rv.download(), it, nu, rtr
        finally:
            for v in (fv, xv, rv):
                v.free()

    def set_mixed_g16(self, on):
        _chk(lib().nekbone_set_mixed_g16(self.handle, int(on)))

    def set_mixed_inner(self, mode):
        """Inner operator of the mixed solver: MIXED_F32, MIXED_G16 or
        MIXED_F16_MFMA."""
        _chk(lib().nekbone_set_mixed_inner(self.handle, int(mode)))

    def set_mixed_refresh(self, fused):
        """1 (default): the FP64 residual refresh flushes x inside the
        element kernel; 0: in a pass of its own."""
        _chk(lib().nekbone_set_mixed_refresh(self.handle, int(fused)))

    def variant_f32(self):
        v = [ctypes.c_int() for _ in range(3)]
        _chk(lib().nekbone_get_variant_f32(self.handle, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def set_variant_f32(self, epb, pf, xu):
        _chk(lib().nekbone_set_variant_f32(self.handle, epb, pf, xu))

    def autotune_mixed(self, precond=PRECOND_JACOBI):
        ms = ctypes.c_double()
        _chk(lib().nekbone_autotune_mixed(self.handle, precond, ctypes.byref(ms)))
        return ms.value

    def ax_local_f32(self, u):
        u = np.ascontiguousarray(u, dtype=np.float32)
        assert u.size == self.n
        w = np.empty(self.n, dtype=np.float32)
        _chk(lib().nekbone_ax_local_f32(self.handle, w.ctypes.data_as(_fp),
                                        u.ctypes.data_as(_fp)))
        return w

    def ax_d(self, w, u):
        _chk(lib().nekbone_ax_d(self.handle, w.ptr, u.ptr))

    def autotune(self, precond=PRECOND_JACOBI):
        ms = ctypes.c_double()
        _chk(lib().nekbone_autotune(self.handle, precond, ctypes.byref(ms)))
        return ms.value

    def variant(self):
        v = [ctypes.c_int() for _ in range(3)]
        _chk(lib().nekbone_get_variant(self.handle, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def set_variant(self, epb, pf, xu):
        _chk(lib().nekbone_set_variant(self.handle, epb, pf, xu))

    # timing --------------------------------------------------------------
    def timing_enable(self, mask):
        _chk(lib().nekbone_timing_enable(self.handle, mask))

    def timing_reset(self):
        _chk(lib().nekbone_timing_reset(self.handle))

    def timing_get(self, kclass):
        n = ctypes.c_int64()
        ms = ctypes.c_double()
        _chk(lib().nekbone_timing_get(self.handle, kclass, ctypes.byref(n), ctypes.byref(ms)))
        return n.value, ms.value
