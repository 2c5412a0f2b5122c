"""Batched CG through the C ABI: every problem of a batch gets exactly the
values of its own single-problem solve (same workgroups, same partial
sums, same scalars), bit for bit, whatever the mix of orders and sizes."""
import numpy as np
import pytest

from conftest import splitmix_uniform

pytestmark = pytest.mark.gpu


def _case(msg):
    print("CASE", msg, flush=True)


def _problems(nb, oracle, specs):
    out = []
    for k, (E, nx1, deformed, mask) in enumerate(specs):
        P = nb.Problem(*nb.cubic(E), nx1)
        P.set_mask_mode(mask)
        if deformed:
            P.set_geometry_deformed(0.1, 10 + k)
        P.setup_jacobi()
        O = oracle.Problem(*nb.cubic(E), nx1)
        f = O.rhs_from_raw(splitmix_uniform(O.n, k))
        out.append((P, f))
    return out


SPECS = [(12, 6, True, 1), (27, 8, False, 1), (7, 10, True, 1), (20, 12, True, 1),
         (64, 6, False, 1), (9, 8, True, 0), (30, 10, False, 1), (5, 12, False, 1)]


@pytest.mark.parametrize("precond,xu", [(1, 1), (1, 0), (0, 1)])
def test_batch_matches_single_solves(nb, oracle, precond, xu):
    _case("batch precond=%d xu=%d" % (precond, xu))
    probs = _problems(nb, oracle, SPECS)
    B = nb.Batch([p for p, _ in probs])
    B.set_xu(xu)
    tol = 1e-8
    fs, xs, rs = [], [], []
    for P, f in probs:
        fv, xv, rv = P.vec(), P.vec(), P.vec()
        fv.upload(f)
        fs.append(fv), xs.append(xv), rs.append(rv)
    its, rtrs = B.cg_tol_d(xs, fs, rs, tol, 400, precond)
    for k, (P, f) in enumerate(probs):
        P.reset_mask()
        x1, r1 = P.vec(), P.vec()
        it1, rtr1 = P.cg_tol_d(x1, fs[k], r1, tol, 400, precond)
        assert its[k] == it1, k
        assert rtrs[k] == rtr1, k
        assert np.array_equal(xs[k].download(), x1.download()), k
        assert np.array_equal(rs[k].download(), r1.download()), k
        assert 0 < its[k] < 400
    B.close()


def test_batch_repeat_and_maxiter(nb, oracle):
    """A second solve reuses the captured graph; a maxiter below the
    chunk size and one that is not a multiple of it."""
    _case("batch repeat / maxiter")
    probs = _problems(nb, oracle, SPECS[:4])
    B = nb.Batch([p for p, _ in probs])
    fs = []
    for P, f in probs:
        v = P.vec()
        v.upload(f)
        fs.append(v)
    xs = [P.vec() for P, _ in probs]
    a = B.cg_tol_d(xs, fs, [None] * 4, 1e-8, 400)
    xa = [x.download() for x in xs]
    b = B.cg_tol_d(xs, fs, [None] * 4, 1e-8, 400)
    assert a == b
    assert all(np.array_equal(x.download(), y) for x, y in zip(xs, xa))
    for m in (3, 17):
        its, _ = B.cg_tol_d(xs, fs, [None] * 4, 1e-30, m, maxiter_ok=True)
        assert its == [m] * 4
        for k, (P, f) in enumerate(probs):
            x1, r1 = P.vec(), P.vec()
            it1, _ = P.cg_tol_d(x1, fs[k], r1, 1e-30, m, nb.PRECOND_JACOBI, maxiter_ok=True)
            assert np.array_equal(xs[k].download(), x1.download())
    ms = B.autotune()
    assert ms > 0
    B.close()
