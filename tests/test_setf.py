"""The reference's right-hand side set_f (src/driver.f:88-101) against
the compiled reference's own output (tests/golden/ref_setf.npz, written
by tests/golden/make_golden.py --setf): 8x8x8 elements at N=7, where the
32-bit product i*i overflows from i = 46,341 on and the flang-built
reference wraps it.  Bit-exact: the raw values come from the same C
library's sin and cos, and dssum and the product with c are exact
reproductions of the reference's order."""
import os

import numpy as np
import pytest

from conftest import GOLDEN


@pytest.fixture(scope="module")
def setf_fx():
    with np.load(os.path.join(GOLDEN, "ref_setf.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_set_f_raw_host_matches_reference(setf_fx, oracle):
    import nekbone_amd as nb
    Q = oracle.Problem(8, 8, 8, 8)
    raw = nb.set_f_raw(Q.n)
    f = Q.rhs_from_raw(raw)
    idx = setf_fx["setf_b888_n8_idx"]
    assert np.array_equal(f[idx], setf_fx["setf_b888_n8_f"])
    assert np.dot(f, f) == setf_fx["setf_b888_n8_sumsq"][0]
    # the overflow matters: a widening product gives other values
    i = np.arange(46341, 46400, dtype=np.float64)
    assert not np.array_equal(raw[46340:46399], np.sin(1e9 * np.cos(1e9 * (i * i))))


def test_set_f_raw_rejects_bad_arguments():
    import nekbone_amd as nb
    assert nb.lib().nekbone_set_f_raw(-1, None) != 0
    assert nb.lib().nekbone_set_f_raw(5, None) != 0
    assert nb.lib().nekbone_set_f_raw(0, None) == 0


@pytest.mark.gpu
def test_set_f_device_matches_reference(nb, setf_fx):
    P = nb.Problem(8, 8, 8, 8)
    f = P.vec()
    P.set_f_d(f)
    fh = f.download()
    idx = setf_fx["setf_b888_n8_idx"]
    assert np.array_equal(fh[idx], setf_fx["setf_b888_n8_f"])
    assert np.dot(fh, fh) == setf_fx["setf_b888_n8_sumsq"][0]
