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


@pytest.mark.parametrize("precond,xu,path", [(1, 1, 0), (1, 0, 0), (0, 1, 0), (1, 1, -1),
                                             (1, 1, 1), (1, 1, 2), (1, 0, 3)])
def test_batch_matches_single_solves(nb, oracle, precond, xu, path):
    """Every element-kernel path (per-group launches, the work-queue kernel
    with each tile set; an order outside the queue's, nx1 = 5, rides along
    per group) and both x-update placements."""
    _case("batch precond=%d xu=%d path=%d" % (precond, xu, path))
    probs = _problems(nb, oracle, SPECS + [(10, 5, True, 1)])
    B = nb.Batch([p for p, _ in probs])
    assert B.path()[0] == -1  # per-group launches by default
    B.set_queue(path)
    B.set_xu(xu)
    assert B.path()[:2] == (path, xu)
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


SPECS_GX = [(12, 6, True, 1), (27, 8, False, 1), (7, 10, True, 1), (20, 12, True, 1),
            (64, 6, False, 1), (9, 8, True, 1), (30, 10, False, 1), (5, 12, False, 1),
            (10, 5, True, 1)]


@pytest.mark.parametrize("precond,xu,path", [(1, 1, -1), (1, 0, -1), (0, 1, -1), (1, 1, 0),
                                             (1, 0, 2)])
def test_batch_gather_form_matches_single_solves(nb, oracle, precond, xu, path):
    """Every member with the full Dirichlet mask and a continuous
    right-hand side: the batch runs the gather form (no gather-scatter
    launch; the update kernel sums the shared copies), on every element-
    kernel path, and each member's values are those of its own single
    solve bit for bit; with one member on the reference's mask list the
    batch falls back to the gather-scatter form."""
    _case("batch gather form precond=%d xu=%d path=%d" % (precond, xu, path))
    probs = _problems(nb, oracle, SPECS_GX)
    B = nb.Batch([p for p, _ in probs])
    B.set_queue(path)
    B.set_xu(xu)
    tol = 1e-8
    fs, xs, rs = [], [], []
    for P, f in probs:
        fv, xv, rv = P.vec(), P.vec(), P.vec()
        fv.upload(f)
        fs.append(fv), xs.append(xv), rs.append(rv)
    its, rtrs = B.cg_tol_d(xs, fs, rs, tol, 400, precond)
    assert B.gather_form()
    for k, (P, f) in enumerate(probs):
        x1, r1 = P.vec(), P.vec()
        it1, rtr1 = P.cg_tol_d(x1, fs[k], r1, tol, 400, precond)
        assert P.cg_form() == (True, False)
        assert its[k] == it1, k
        assert rtrs[k] == rtr1, k
        assert np.array_equal(xs[k].download(), x1.download()), k
        assert np.array_equal(rs[k].download(), r1.download()), k
        assert 0 < its[k] < 400
    B.close()
    # one member on the reference's mask list: gather-scatter form
    probs = _problems(nb, oracle, SPECS_GX[:3] + [(9, 8, True, 0)])
    B = nb.Batch([p for p, _ in probs])
    fs = []
    for P, f in probs:
        v = P.vec()
        v.upload(f)
        fs.append(v)
    B.cg_tol_d([P.vec() for P, _ in probs], fs, [None] * 4, tol, 400, precond)
    assert not B.gather_form()
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
    tileset, xu, nsets = B.path()
    assert -1 <= tileset < nsets and xu in (0, 1)
    # after tuning, the chosen shapes are on the problems: still the values
    # of their single solves
    its, _ = B.cg_tol_d(xs, fs, [None] * 4, 1e-8, 400)
    for k, (P, f) in enumerate(probs):
        P.reset_mask()
        x1, r1 = P.vec(), P.vec()
        it1, _ = P.cg_tol_d(x1, fs[k], r1, 1e-8, 400, nb.PRECOND_JACOBI)
        assert its[k] == it1
        assert np.array_equal(xs[k].download(), x1.download())
    with pytest.raises(nb.NekboneError):
        B.set_queue(nsets)
    B.close()


def _true_rel_residual(P, f, x):
    """sqrt(glsc3(r,c,r) / glsc3(f0,c,f0)) for r = mask(f) - A x, with the
    library's FP64 operator (bit-exact against the reference's ax)."""
    f0 = P.maskit(f)
    r = f0 - P.ax(x)
    c = P.c()
    return float(np.sqrt(np.sum(r * c * r) / np.sum(f0 * c * f0)))


def test_batch_vs_oracle(nb, oracle):
    """Every problem of a mixed-order batch solved to 1e-8: the true
    residual of its x, recomputed by the oracle's FP64 operator, meets the
    tolerance; and for two problems the iterates after a fixed number of
    iterations (10 and 37) against the oracle's CG (x to 1e-12 of max|x|,
    the final rtr to 1e-12 of the oracle's history entry)."""
    _case("batch vs oracle")
    specs = SPECS
    probs = _problems(nb, oracle, specs)
    B = nb.Batch([p for p, _ in probs])
    tol = 1e-8
    fs, xs = [], []
    for P, f in probs:
        fv, xv = P.vec(), P.vec()
        fv.upload(f)
        fs.append(fv), xs.append(xv)
    its, _ = B.cg_tol_d(xs, fs, [None] * len(probs), tol, 400, nb.PRECOND_JACOBI)
    Os = []
    for k, ((E, nx1, deformed, mask), (P, f)) in enumerate(zip(specs, probs)):
        O = oracle.Problem(*nb.cubic(E), nx1)
        O.set_mask_mode(mask)
        if deformed:
            O.deformed_g(0.1, 10 + k)
        O.setup_jacobi()
        Os.append(O)
        O.reset_mask()
        x = xs[k].download()
        f0 = O.maskit(f)
        O.reset_mask()
        O.maskit(f)  # the reference mask: every later maskit applies the list
        r = f0 - O.ax(x)
        c = O.c()
        rel = float(np.sqrt(np.sum(r * c * r) / np.sum(f0 * c * f0)))
        assert rel <= tol * 1.01, (k, rel, its[k])
    for m in (10, 37):
        for P, _ in probs:
            P.reset_mask()
        its, rtrs = B.cg_tol_d(xs, fs, [None] * len(probs), 1e-30, m, nb.PRECOND_JACOBI,
                               maxiter_ok=True)
        for k in (0, 3):
            O = Os[k]
            O.reset_mask()
            xo, _, ho = O.cg(probs[k][1], m, oracle.PRECOND_JACOBI)
            x = xs[k].download()
            np.testing.assert_allclose(x, xo, rtol=0, atol=1e-12 * np.abs(xo).max())
            np.testing.assert_allclose(rtrs[k], ho[-1], rtol=1e-12)
    B.close()


def test_batch_full_size_workload(nb):
    """The batched workload at its full size: 64 problems, element counts
    log-uniform in [64, 8192] (box from cubic), orders drawn from
    {5, 7, 9, 11}, Dirichlet faces, RHS c * dssum(U[-1,1) seed k), Jacobi
    to 1e-8.  Every problem converges before maxiter, and the true
    residual of its x (recomputed with the FP64 operator) meets 1e-8."""
    _case("batch full size: 64 problems")
    import math
    u = (splitmix_uniform(128, 77) + 1.0) / 2.0
    probs, fs, xs = [], [], []
    for k in range(64):
        E = int(round(math.exp(math.log(64) + u[2 * k] * (math.log(8192) - math.log(64)))))
        nx1 = (6, 8, 10, 12)[min(3, int(4 * u[2 * k + 1]))]
        P = nb.Problem(*nb.cubic(E), nx1)
        P.set_mask_mode(nb.MASK_DIRICHLET)
        P.setup_jacobi()
        f, x = P.vec(), P.vec()
        P.rand_uniform_d(f, k)
        P.rhs_from_raw_d(f)
        probs.append(P), fs.append(f), xs.append(x)
    npts = sum(P.n for P in probs)
    print("batch full size: %d local points, E %d..%d" %
          (npts, min(P.nelt for P in probs), max(P.nelt for P in probs)), flush=True)
    assert max(P.nelt for P in probs) > 2048 and min(P.nelt for P in probs) < 256
    B = nb.Batch(probs)
    tol = 1e-8
    its, _ = B.cg_tol_d(xs, fs, [None] * 64, tol, 2000, nb.PRECOND_JACOBI)
    assert all(0 < it < 2000 for it in its)
    worst = 0.0
    for P, f, x in zip(probs, fs, xs):
        worst = max(worst, _true_rel_residual(P, f.download(), x.download()))
    print("batch full size: iterations %d..%d, worst true relative residual %.3e" %
          (min(its), max(its), worst), flush=True)
    assert worst <= tol * 1.01
    B.close()
    for P in probs:
        P.close()
