"""Mixed-precision CG (FP32 inner iterations with FP64 reliable updates)
and its FP32 element operator, through the C ABI, against the oracle's
FP64 operator.

Bars (there is no reference counterpart to be bit-compatible with):
  * FP32 element operator: within 2e-6 of the FP64 oracle, relative to
    max|w|, on the same (FP32-rounded) input - FP32 unit roundoff 6e-8
    times the ~30 rounding steps on a point's path.
  * Solve: the returned x meets the FP64 stopping rule
    glsc3(r,c,r) <= tol^2 glsc3(r0,c,r0) when the oracle recomputes
    r = mask(f - A x) in FP64, the reported final rtr is that residual,
    and x is as close to the converged FP64 solution as the FP64 solve
    to the same tolerance (within 10x, or 1e-13 of max|x|)."""
import numpy as np
import pytest

from conftest import splitmix_uniform

pytestmark = pytest.mark.gpu


def _case(msg):
    print("CASE", msg, flush=True)


def _pair(nb, oracle, nel, nx1, deformed, gseed=5, jacobi=True):
    P = nb.Problem(*nel, nx1)
    O = oracle.Problem(*nel, nx1)
    P.set_mask_mode(nb.MASK_DIRICHLET)
    O.set_mask_mode(oracle.MASK_DIRICHLET)
    if deformed:
        P.set_geometry_deformed(0.1, gseed)
        O.deformed_g(0.1, gseed)
    if jacobi:
        P.setup_jacobi()
        O.setup_jacobi()
    return P, O


@pytest.mark.parametrize("deformed", [False, True])
@pytest.mark.parametrize("nel,nx1", [((2, 2, 2), 8), ((3, 2, 2), 10), ((2, 3, 1), 12),
                                     ((1, 1, 1), 5), ((2, 1, 2), 16)])
def test_ax_local_f32_vs_fp64_oracle(nb, oracle, nel, nx1, deformed):
    _case("ax_local_f32 %s %d deformed=%s" % (nel, nx1, deformed))
    P, O = _pair(nb, oracle, nel, nx1, deformed, jacobi=False)
    u32 = splitmix_uniform(P.n, 11).astype(np.float32)
    w32 = P.ax_local_f32(u32)
    u = u32.astype(np.float64)
    n3 = nx1 ** 3
    w64 = np.concatenate([O.ax_e(e, u[e * n3:(e + 1) * n3]) for e in range(O.nelt)])
    scale = np.abs(w64).max()
    assert np.abs(w32.astype(np.float64) - w64).max() <= 2e-6 * scale


def _converged_oracle(O, f, precond):
    """The oracle's FP64 CG run to its attainable accuracy: iterations
    stop before rtr reaches zero (where beta = 0/0 would follow)."""
    x, r, h = O.cg(f, 1500, precond)
    rtr = h[3::3]
    bad = np.nonzero(~np.isfinite(rtr) | (rtr <= 1e-30 * h[0]))[0]
    if bad.size:
        x, r, h = O.cg(f, max(int(bad[0]) - 1, 1), precond)
    assert np.all(np.isfinite(x))
    return x


def _fp64_tol_solve(nb, P, f, tol, precond):
    fv, xv, rv = P.vec(), P.vec(), P.vec()
    fv.upload(f)
    it, _ = P.cg_tol_d(xv, fv, rv, tol, 5000, precond)
    x = xv.download()
    for v in (fv, xv, rv):
        v.free()
    return x, it


@pytest.mark.parametrize("precond", [1, 0])
@pytest.mark.parametrize("nel,nx1,deformed", [((4, 3, 2), 8, True), ((3, 3, 3), 10, True),
                                              ((4, 4, 2), 6, False)])
def test_cg_mixed_meets_fp64_criterion(nb, oracle, nel, nx1, deformed, precond):
    _case("cg_mixed %s %d deformed=%s precond=%d" % (nel, nx1, deformed, precond))
    tol = 1e-10
    P, O = _pair(nb, oracle, nel, nx1, deformed, jacobi=bool(precond))
    f = O.rhs_from_raw(splitmix_uniform(O.n, 2))
    x, r, it, nupd, rtr = P.cg_mixed(f, tol, 3000, precond)
    print("mixed: %d inner iterations, %d FP64 updates, final rtr %.3e" % (it, nupd, rtr),
          flush=True)
    assert nupd >= 1 and 0 < it < 3000
    c = O.c()
    f0 = O.maskit(f)
    rtrue = f0 - O.ax(x)
    rr = float(np.sum(rtrue * c * rtrue))
    rr0 = float(np.sum(f0 * c * f0))
    assert rr <= tol ** 2 * rr0 * 1.0001
    assert abs(rtr - rr) <= 1e-6 * rr
    np.testing.assert_allclose(r, rtrue, rtol=0, atol=1e-9 * np.abs(f0).max())
    # converged FP64 solution, and the FP64 solve to the same tolerance
    xs = _converged_oracle(O, f, precond)
    x64, it64 = _fp64_tol_solve(nb, P, f, tol, precond)
    err = np.abs(x - xs).max()
    err64 = np.abs(x64 - xs).max()
    assert err <= max(10 * err64, 1e-13 * np.abs(xs).max())
    print("iters mixed %d (+%d FP64 updates), fp64 %d; err %.2e vs %.2e" %
          (it, nupd, it64, err, err64))


def test_cg_mixed_maxiter_and_geometry_change(nb, oracle):
    _case("cg_mixed maxiter / geometry change")
    P, O = _pair(nb, oracle, (3, 2, 2), 8, True, gseed=1)
    f = O.rhs_from_raw(splitmix_uniform(O.n, 6))
    # a tolerance it cannot reach: stops at maxiter with x64 and r64 current
    x, r, it, nupd, rtr = P.cg_mixed(f, 1e-30, 25, 1, maxiter_ok=True)
    assert it == 25 and nupd >= 1
    f0 = O.maskit(f)
    rtrue = f0 - O.ax(x)
    assert abs(rtr - float(np.sum(rtrue * O.c() * rtrue))) <= 1e-6 * rtr
    # new geometry and diagonal: the FP32 copies must follow
    P.set_geometry_deformed(0.1, 9)
    O.deformed_g(0.1, 9)
    P.setup_jacobi()
    O.setup_jacobi()
    x, r, it, nupd, rtr = P.cg_mixed(f, 1e-10, 3000, 1)
    rtrue = f0 - O.ax(x)
    rr = float(np.sum(rtrue * O.c() * rtrue))
    assert rr <= 1e-20 * float(np.sum(f0 * O.c() * f0)) * 1.0001


def test_cg_mixed_rejects_bad_arguments(nb):
    _case("cg_mixed arguments")
    P = nb.Problem(2, 2, 2, 6)
    f = np.zeros(P.n)
    with pytest.raises(nb.NekboneError):
        P.cg_mixed(f, 1e-10, 10, nb.PRECOND_JACOBI)  # Jacobi not set up
    with pytest.raises(nb.NekboneError):
        P.cg_mixed(f, 1e-10, 10, nb.PRECOND_IDENTITY, delta=2.0)
    with pytest.raises(nb.NekboneError):
        P.cg_mixed(f, -1.0, 10, nb.PRECOND_IDENTITY)


@pytest.mark.parametrize("deformed", [False, True])
@pytest.mark.parametrize("nel,nx1", [((2, 2, 2), 8), ((3, 2, 2), 10), ((2, 3, 1), 12),
                                     ((1, 2, 1), 5), ((2, 1, 2), 16), ((2, 2, 1), 4)])
def test_ax_local_f16_mfma_vs_fp64_oracle(nb, oracle, nel, nx1, deformed):
    """FP16 matrix-core operator: within 2e-2 of max|w| of the FP64
    oracle (FP16 unit roundoff 4.9e-4 on the scaled input, on D, on the
    factors and on the scaled products, through sums of nx1 terms), and
    not equal to the FP32 operator (the FP16 path ran).  A wrong index or
    transposition in any of the six contractions gives errors of order
    max|w|."""
    _case("ax_local_f16_mfma %s %d deformed=%s" % (nel, nx1, deformed))
    P, O = _pair(nb, oracle, nel, nx1, deformed, jacobi=False)
    u32 = splitmix_uniform(P.n, 11).astype(np.float32)
    w32 = P.ax_local_f32(u32)
    P.set_mixed_inner(nb.MIXED_F16_MFMA)
    w16 = P.ax_local_f32(u32)
    u = u32.astype(np.float64)
    n3 = nx1 ** 3
    w64 = np.concatenate([O.ax_e(e, u[e * n3:(e + 1) * n3]) for e in range(O.nelt)])
    scale = np.abs(w64).max()
    err = np.abs(w16.astype(np.float64) - w64).max() / scale
    print("f16 mfma operator: max err %.2e of max|w|" % err)
    assert err <= 2e-2
    assert np.abs(w16 - w32).max() > 1e-5 * scale
    # a scaled input gives the scaled output (per-element scaling)
    w16s = P.ax_local_f32((u32 * np.float32(2.0 ** -20)))
    np.testing.assert_allclose(w16s.astype(np.float64) * 2.0 ** 20, w16, rtol=0,
                               atol=1e-6 * scale)


@pytest.mark.parametrize("nel,nx1", [((4, 3, 2), 8), ((3, 3, 3), 10)])
def test_cg_mixed_f16_mfma_meets_fp64_criterion(nb, oracle, nel, nx1):
    """The FP64 reliable updates make the FP16 inner operator reach the
    FP64 criterion all the same (more inner iterations)."""
    _case("cg_mixed f16 mfma %s %d" % (nel, nx1))
    tol = 1e-10
    P, O = _pair(nb, oracle, nel, nx1, True, jacobi=True)
    P.set_mixed_inner(nb.MIXED_F16_MFMA)
    f = O.rhs_from_raw(splitmix_uniform(O.n, 2))
    x, r, it, nupd, rtr = P.cg_mixed(f, tol, 5000, 1)
    c = O.c()
    f0 = O.maskit(f)
    rtrue = f0 - O.ax(x)
    rr = float(np.sum(rtrue * c * rtrue))
    rr0 = float(np.sum(f0 * c * f0))
    assert rr <= tol ** 2 * rr0 * 1.0001
    assert abs(rtr - rr) <= 1e-6 * rr
    P.set_mixed_inner(nb.MIXED_F32)
    _, _, it32, nupd32, _ = P.cg_mixed(f, tol, 5000, 1)
    print("inner iterations: f16 mfma %d (+%d updates), f32 %d (+%d)" % (it, nupd, it32, nupd32))


def test_mixed_inner_rejects_unknown_mode(nb):
    P = nb.Problem(2, 2, 2, 8)
    with pytest.raises(nb.NekboneError):
        P.set_mixed_inner(3)
    Q = nb.Problem(2, 2, 2, 3)
    with pytest.raises(nb.NekboneError):
        Q.set_mixed_inner(nb.MIXED_F16_MFMA)  # matrix-core kernels for nx1 >= 4


@pytest.mark.parametrize("precond", [1, 0])
@pytest.mark.parametrize("nel,nx1,deformed", [((3, 2, 2), 9, True), ((3, 1, 1), 5, True),
                                              ((2, 2, 3), 7, False), ((2, 2, 2), 12, True)])
def test_cg_mixed_flush_in_operator_matches_separate_pass(nb, oracle, nel, nx1, deformed, precond):
    """The FP64 residual refresh with the x flush inside the element
    kernel against the flush as a pass of its own: the same arithmetic on
    every point (x + float(x32 + alpha p32), the operator, the residual
    kernel), so the solves agree bit for bit; and both final rtr are the
    FP64 residual's.  (Gather-scatter form: in the gather form the fused
    refresh is the device-driven one, tested below.)"""
    _case("cg_mixed fused refresh %s %d deformed=%s precond=%d" % (nel, nx1, deformed, precond))
    P, O = _pair(nb, oracle, nel, nx1, deformed, jacobi=bool(precond))
    P.set_gather(False)
    f = O.rhs_from_raw(splitmix_uniform(O.n, 8))
    out = {}
    for fused in (1, 0):
        P.set_mixed_refresh(fused)
        out[fused] = P.cg_mixed(f, 1e-10, 3000, precond)
    (xa, ra, ita, nua, rtra), (xb, rb, itb, nub, rtrb) = out[1], out[0]
    assert (ita, nua, rtra) == (itb, nub, rtrb) and nua >= 1
    assert np.array_equal(xa, xb) and np.array_equal(ra, rb)
    f0 = O.maskit(f)
    for x, rtr in ((xa, rtra), (xb, rtrb)):
        rtrue = f0 - O.ax(x)
        assert abs(rtr - float(np.sum(rtrue * O.c() * rtrue))) <= 1e-6 * rtr


def test_mixed_refresh_rejects_bad_value(nb):
    _case("mixed refresh bad value")
    P = nb.Problem(2, 2, 2, 6)
    with pytest.raises(nb.NekboneError):
        P.set_mixed_refresh(2)


@pytest.mark.parametrize("deformed", [False, True])
@pytest.mark.parametrize("nel,nx1", [((2, 2, 2), 8), ((3, 2, 2), 10), ((2, 3, 1), 12),
                                     ((1, 1, 1), 5), ((2, 1, 2), 16)])
def test_ax_local_g16_vs_fp64_oracle(nb, oracle, nel, nx1, deformed):
    """FP16 geometric factors (one FP32 scale per element and factor, then
    each factor rounded to half precision: relative error <= 2^-11 above
    2^-14 of the largest, absolute 2^-25 of it below), FP32 contractions:
    within 5e-3 of max|w| of the FP64 oracle, and not equal to the FP32
    operator (the FP16 factors were read).  A wrong scale index or factor
    slot gives errors of order max|w|."""
    _case("ax_local_g16 %s %d deformed=%s" % (nel, nx1, deformed))
    P, O = _pair(nb, oracle, nel, nx1, deformed, jacobi=False)
    u32 = splitmix_uniform(P.n, 11).astype(np.float32)
    w32 = P.ax_local_f32(u32)
    P.set_mixed_inner(nb.MIXED_G16)
    w16 = P.ax_local_f32(u32)
    u = u32.astype(np.float64)
    n3 = nx1 ** 3
    w64 = np.concatenate([O.ax_e(e, u[e * n3:(e + 1) * n3]) for e in range(O.nelt)])
    scale = np.abs(w64).max()
    err = np.abs(w16.astype(np.float64) - w64).max() / scale
    print("g16 operator: max err %.2e of max|w|" % err)
    assert err <= 5e-3
    assert np.abs(w16 - w32).max() > 1e-7 * scale


@pytest.mark.parametrize("precond", [1, 0])
@pytest.mark.parametrize("nel,nx1", [((4, 3, 2), 8), ((3, 3, 3), 10)])
def test_cg_mixed_g16_meets_fp64_criterion(nb, oracle, nel, nx1, precond):
    """The FP16-factor inner operator (the fastest one at 32^3 N=9) on a
    deformed mesh: the returned x meets the FP64 criterion with r
    recomputed by the oracle, before maxiter, and the reported final rtr
    is that residual's."""
    _case("cg_mixed g16 %s %d precond=%d" % (nel, nx1, precond))
    tol = 1e-10
    P, O = _pair(nb, oracle, nel, nx1, True, jacobi=bool(precond))
    P.set_mixed_inner(nb.MIXED_G16)
    f = O.rhs_from_raw(splitmix_uniform(O.n, 2))
    x, r, it, nupd, rtr = P.cg_mixed(f, tol, 5000, precond)
    print("g16 mixed: %d inner iterations, %d FP64 updates" % (it, nupd), flush=True)
    assert nupd >= 1 and 0 < it < 5000
    c = O.c()
    f0 = O.maskit(f)
    rtrue = f0 - O.ax(x)
    rr = float(np.sum(rtrue * c * rtrue))
    rr0 = float(np.sum(f0 * c * f0))
    assert rr <= tol ** 2 * rr0 * 1.0001
    assert abs(rtr - rr) <= 1e-6 * rr
    xs = _converged_oracle(O, f, precond)
    err = np.abs(x - xs).max() / np.abs(xs).max()
    print("g16 mixed: max error vs converged FP64 solution %.2e of max|x|" % err)
    assert err <= 1e-8


@pytest.mark.parametrize("inner", ["f32", "g16"])
@pytest.mark.parametrize("precond", [1, 0])
def test_cg_mixed_gather_and_gs_forms(nb, oracle, inner, precond):
    """Inner iterations in the gather form (the update kernel sums the
    shared copies; the default on box meshes with the Dirichlet mask) and
    with the gather-scatter kernel: both meet the FP64 criterion with r
    recomputed by the oracle, and their inner iteration counts agree
    within 2 (the forms differ only in the rounding of the pap sum)."""
    _case("cg_mixed forms %s precond=%d" % (inner, precond))
    tol = 1e-10
    P, O = _pair(nb, oracle, (3, 3, 3), 10, True, jacobi=bool(precond))
    if inner == "g16":
        P.set_mixed_inner(nb.MIXED_G16)
    f = O.rhs_from_raw(splitmix_uniform(O.n, 2))
    c = O.c()
    f0 = O.maskit(f)
    rr0 = float(np.sum(f0 * c * f0))
    its = []
    for on in (True, False):
        assert P.set_gather(on) == on
        x, r, it, nupd, rtr = P.cg_mixed(f, tol, 5000, precond)
        print("gather=%s: %d inner iterations, %d FP64 updates" % (on, it, nupd), flush=True)
        assert nupd >= 1 and 0 < it < 5000
        rtrue = f0 - O.ax(x)
        rr = float(np.sum(rtrue * c * rtrue))
        assert rr <= tol ** 2 * rr0 * 1.0001
        assert abs(rtr - rr) <= 1e-6 * rr
        its.append(it)
    assert abs(its[0] - its[1]) <= 2


@pytest.mark.parametrize("inner", ["f32", "g16"])
@pytest.mark.parametrize("precond", [1, 0])
@pytest.mark.parametrize("nel,nx1", [((3, 3, 3), 10), ((4, 2, 3), 6)])
def test_cg_mixed_device_driven_refreshes(nb, oracle, inner, precond, nel, nx1):
    """Gather form: the refreshes decided on the device (the refresh
    kernels queued after every chunk run when due, the host checks one
    chunk behind) against the host-decided loop (set_mixed_refresh(0)):
    the same refresh rule, so the same inner iteration count within 1 and
    the same number of refreshes within 1 (the residual is the same per
    point, the copies summed in the same member order, but its dot
    products are summed over other blocks); both meet the FP64 criterion
    with r recomputed by the oracle, and each final rtr is that
    residual's.  Then the iteration limit: a tolerance that cannot be
    met stops after exactly maxiter inner iterations (not a multiple of
    the chunk) with x and r of the last refresh."""
    _case("cg_mixed device-driven %s %s %d precond=%d" % (inner, nel, nx1, precond))
    tol = 1e-10
    P, O = _pair(nb, oracle, nel, nx1, True, jacobi=bool(precond))
    if inner == "g16":
        P.set_mixed_inner(nb.MIXED_G16)
    assert P.set_gather(True)
    f = O.rhs_from_raw(splitmix_uniform(O.n, 3))
    c = O.c()
    f0 = O.maskit(f)
    rr0 = float(np.sum(f0 * c * f0))
    out = {}
    for fused in (1, 0):
        P.set_mixed_refresh(fused)
        x, r, it, nupd, rtr = P.cg_mixed(f, tol, 5000, precond)
        print("refresh=%d: %d inner iterations, %d FP64 updates" % (fused, it, nupd), flush=True)
        assert nupd >= 1 and 0 < it < 5000
        rtrue = f0 - O.ax(x)
        rr = float(np.sum(rtrue * c * rtrue))
        assert rr <= tol ** 2 * rr0 * 1.0001
        assert abs(rtr - rr) <= 1e-6 * rr
        np.testing.assert_allclose(r, rtrue, rtol=0, atol=1e-9 * np.abs(f0).max())
        out[fused] = (it, nupd)
    assert abs(out[1][0] - out[0][0]) <= 1 and abs(out[1][1] - out[0][1]) <= 1
    P.set_mixed_refresh(1)
    x, r, it, nupd, rtr = P.cg_mixed(f, 1e-30, 21, precond, maxiter_ok=True)
    assert it == 21 and nupd >= 1
    rtrue = f0 - O.ax(x)
    assert abs(rtr - float(np.sum(rtrue * c * rtrue))) <= 1e-6 * rtr
    with pytest.raises(nb.NotConverged):
        P.cg_mixed(f, 1e-30, 21, precond)


@pytest.mark.parametrize("inner", ["f32", "g16"])
@pytest.mark.parametrize("nx1", list(range(2, 17)))
def test_cg_mixed_every_order(nb, oracle, inner, nx1):
    """The mixed-precision CG at every order, FP32 and FP16 factors, in
    the gather form with device-driven refreshes (the default on these
    box meshes), Jacobi, deformed: the returned x meets the FP64 criterion
    with r recomputed by the oracle, before maxiter, and the reported final
    rtr is that residual's."""
    _case("cg_mixed order %d %s" % (nx1, inner))
    tol = 1e-10
    nel = (6, 5, 4) if nx1 <= 4 else (3, 2, 2) if nx1 <= 10 else (2, 2, 1)
    P, O = _pair(nb, oracle, nel, nx1, True, gseed=9, jacobi=True)
    if inner == "g16":
        P.set_mixed_inner(nb.MIXED_G16)
    f = O.rhs_from_raw(splitmix_uniform(O.n, 100 + nx1))
    x, r, it, nupd, rtr = P.cg_mixed(f, tol, 3000, 1)
    assert nupd >= 1 and 0 < it < 3000
    c = O.c()
    f0 = O.maskit(f)
    rtrue = f0 - O.ax(x)
    rr = float(np.sum(rtrue * c * rtrue))
    rr0 = float(np.sum(f0 * c * f0))
    assert rr <= tol ** 2 * rr0 * 1.0001
    assert abs(rtr - rr) <= 1e-6 * rr
