"""The allocation-free glsc3 (nekbone_glsc3_noalloc_d), through the C ABI,
against the entry points it is a variant of: the reference-named glsc3_
(src/math.f:603-617) and nekbone_glsc3.  Same kernels and block partials,
so the bar is bit-identical; the oracle's sequential sum bounds all three
to 1e-13 of sum |a b m|.

Two calls in a row, the second on a smaller input (fewer block partials
than the first left in the static scratch), then a call on a prefix of
the vectors, then the rejected arguments."""
import ctypes

import numpy as np
import pytest

from conftest import splitmix_uniform

pytestmark = pytest.mark.gpu


def _legacy_glsc3(nb, a, b, m):
    n = ctypes.c_int(a.size)
    p = lambda v: np.ascontiguousarray(v, dtype=np.float64).ctypes.data_as(
        ctypes.POINTER(ctypes.c_double))
    return float(nb.lib().glsc3_(p(a), p(b), p(m), ctypes.byref(n)))


def _noalloc(nb, P, a, b, m, n):
    av, bv, mv, out = P.vec(), P.vec(), P.vec(), P.vec()
    try:
        for v, h in ((av, a), (bv, b), (mv, m)):
            v.upload(h)
        out.upload(np.full(P.n, np.nan))
        nb.core.glsc3_noalloc_d(av, bv, mv, n, out)
        nb.sync()
        return float(out.download()[0])
    finally:
        for v in (av, bv, mv, out):
            v.free()


def test_glsc3_noalloc_matches_glsc3(nb, oracle):
    print("CASE glsc3_noalloc: 4x3x2 N=7 then 2x2x1 N=5", flush=True)
    P1 = nb.Problem(4, 3, 2, 8)   # 12,288 points: 6 block partials
    P2 = nb.Problem(2, 2, 1, 6)   # 864 points: 1 block partial
    assert P2.n < P1.n
    for P, seed in ((P1, 1), (P2, 7)):
        a, b = splitmix_uniform(P.n, seed), splitmix_uniform(P.n, seed + 1)
        m = 1.0 + 0.5 * splitmix_uniform(P.n, seed + 2)
        got = _noalloc(nb, P, a, b, m, P.n)
        assert got == P.glsc3(a, b, m)
        assert got == _legacy_glsc3(nb, a, b, m)
        assert abs(got - oracle.glsc3(a, b, m)) <= 1e-13 * np.sum(np.abs(a * b * m))
    # a prefix of the vectors (n below their length)
    a, b = splitmix_uniform(P1.n, 11), splitmix_uniform(P1.n, 12)
    m = np.ones(P1.n)
    k = 5000
    got = _noalloc(nb, P1, a, b, m, k)
    assert got == _legacy_glsc3(nb, a[:k], b[:k], m[:k])
    assert _noalloc(nb, P2, np.ones(P2.n), np.ones(P2.n), np.ones(P2.n), 0) == 0.0


def test_glsc3_noalloc_rejects_bad_arguments(nb):
    print("CASE glsc3_noalloc arguments", flush=True)
    P = nb.Problem(2, 2, 2, 4)
    v = P.vec()
    try:
        with pytest.raises(nb.NekboneError):
            nb.core.glsc3_noalloc_d(v, v, v, -1, v)
        with pytest.raises(nb.NekboneError):
            nb.core.glsc3_noalloc_d(v, v, v, 2 ** 31, v)
        with pytest.raises(nb.NekboneError):
            nb.core.glsc3_noalloc_d(None, v, v, 4, v)
    finally:
        v.free()
