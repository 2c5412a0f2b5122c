"""Host-side parts of the reference-named entry points that need no GPU:
semhat_ (src/semhat.f:2-37) against the compiled reference's dxm1, GLL
points and weights (bit-exact: the same Newton iteration and Fornberg
weights in the same floating-point order)."""
import ctypes

import numpy as np
import pytest

from conftest import SMALL_CASES, tag_of

_dp = ctypes.POINTER(ctypes.c_double)


@pytest.mark.parametrize("nel,nx1", SMALL_CASES)
def test_semhat_matches_reference(golden, nel, nx1):
    import nekbone_amd
    L = nekbone_amd.lib()
    tag = tag_of(nel, nx1)
    m = nx1
    a, c, d = np.empty(m * m), np.empty(m * m), np.empty(m * m)
    b, z, w = np.empty(m), np.empty(m), np.empty(2 * m)
    n = ctypes.c_int(nx1 - 1)
    L.semhat_(*[v.ctypes.data_as(_dp) for v in (a, b, c, d, z, w)], ctypes.byref(n))
    assert np.array_equal(d, golden[tag + "_dxm1"])
    assert np.array_equal(z, golden[tag + "_zgm1"])
    assert np.array_equal(b, golden[tag + "_wxm1"])
    # a = D^T B D and c = B D, the reference's own loops (src/semhat.f:26-34)
    D = d.reshape(m, m, order="F")
    np.testing.assert_allclose(a.reshape(m, m, order="F"), D.T @ np.diag(b) @ D, rtol=0,
                               atol=1e-13 * np.abs(a).max())
    np.testing.assert_allclose(c.reshape(m, m, order="F"), np.diag(b) @ D, rtol=0, atol=0)


def test_tuning_file_roundtrip(tmp_path):
    """nekbone_tuning_load takes the entries whose shape is compiled for
    its order and skips the others; nekbone_tuning_save writes them back
    (host code only, no GPU)."""
    import nekbone_amd as nb
    src = tmp_path / "in.json"
    src.write_text('{"format": "nekbone_amd tuning 1", "element_kernel": [\n'
                   '  {"nx1": 10, "precond": 1, "arith": 0, "epb": 1, "pf": 0, "xu": 1},\n'
                   '  {"nx1": 10, "precond": -1, "arith": -1, "epb": 2, "pf": 1, "xu": 1},\n'
                   '  {"nx1": 10, "precond": 0, "arith": 0, "epb": 7, "pf": 0, "xu": 1}\n]}\n')
    assert nb.tuning_load(str(src)) == 2
    out = tmp_path / "out.json"
    nb.tuning_save(str(out))
    text = out.read_text()
    assert '{"nx1": 10, "precond": 1, "arith": 0, "epb": 1, "pf": 0, "xu": 1}' in text
    assert '{"nx1": 10, "precond": -1, "arith": -1, "epb": 2, "pf": 1, "xu": 1}' in text
    assert '"epb": 7' not in text
    bad = tmp_path / "bad.json"
    bad.write_text("{}")
    with pytest.raises(nb.NekboneError):
        nb.tuning_load(str(bad))
