"""The oracle against the compiled reference: every fixture under
tests/golden/ref_small.npz was produced by the flang-compiled reference
(oracle/_ref), and the oracle must reproduce it bit for bit."""
import numpy as np
import pytest

from conftest import SMALL_CASES, splitmix_uniform, tag_of


def test_rng_stream_matches(oracle):
    for seed in (0, 3, 7, 2**63 + 5):
        assert np.array_equal(oracle.rand_uniform(1000, seed), splitmix_uniform(1000, seed))


@pytest.mark.parametrize("nel,nx1", SMALL_CASES)
def test_setup_bit_exact(oracle, golden, nel, nx1):
    t = tag_of(nel, nx1)
    P = oracle.Problem(*nel, nx1)
    assert np.array_equal(P.dxm1(), golden[t + "_dxm1"])
    assert np.array_equal(P.c(), golden[t + "_c"])
    assert np.array_equal(P.g(), golden[t + "_g"])
    assert np.array_equal(P.glo(), golden[t + "_glo"])
    z, w = oracle.zwgll(nx1)
    assert np.array_equal(z, golden[t + "_zgm1"])
    assert np.array_equal(w, golden[t + "_wxm1"])


@pytest.mark.parametrize("nel,nx1", SMALL_CASES)
def test_ax_and_mask_bit_exact(oracle, golden, nel, nx1):
    t = tag_of(nel, nx1)
    P = oracle.Problem(*nel, nx1)
    u = splitmix_uniform(P.n, 7)
    # first ax builds the mask list (src/cg.f:230-253), the second uses it
    assert np.array_equal(P.ax(u), golden[t + "_ax_w1"])
    assert np.array_equal(P.ax(u), golden[t + "_ax_w2"])
    full, ref = P.mask_lists()
    cm = golden[t + "_cmask"]
    # the reference records element 1's positions once per element
    assert cm.size == P.nelt * ref.size
    assert np.array_equal(np.unique(cm), ref)


@pytest.mark.parametrize("nel,nx1", SMALL_CASES)
def test_dssum_bit_exact(oracle, golden, nel, nx1):
    t = tag_of(nel, nx1)
    P = oracle.Problem(*nel, nx1)
    assert np.array_equal(P.dssum(splitmix_uniform(P.n, 7)), golden[t + "_dssum"])


@pytest.mark.parametrize("tag,nel,nx1,niter,ncalls,seed", [
    ("cg_b888_n8", (8, 8, 8), 8, 100, 2, 0),
    ("cg_b345_n6", (3, 4, 5), 6, 40, 1, 3),
])
def test_cg_history_bit_exact(oracle, golden, tag, nel, nx1, niter, ncalls, seed):
    """Full glsc3 history (rnorm^2, then rtz1, pap, rtr per iteration) of
    the driver's consecutive cg calls, identity preconditioner, the
    reference's mask behaviour."""
    P = oracle.Problem(*nel, nx1)
    f = P.rhs_from_raw(splitmix_uniform(P.n, seed))
    ref = golden[tag + "_hist"]
    for call in range(ncalls):
        x, r, h = P.cg(f, niter)
        assert np.array_equal(h, ref[call]), "call %d" % call
    if tag == "cg_b888_n8":
        assert np.array_equal(x[:4096], golden[tag + "_x2_head"])
        assert np.array_equal(r[:4096], golden[tag + "_r2_head"])
        assert np.dot(x, x) == golden[tag + "_x2_sumsq"][0]


HSMG_CASES = [((2, 2, 2), 8, "mg_b222_n8"), ((3, 2, 2), 10, "mg_b322_n10"),
              ((4, 4, 4), 8, "mg_b444_n8")]


@pytest.mark.parametrize("nel,nx1,tag", HSMG_CASES)
def test_hsmg_apply_vs_reference(oracle, golden, nel, nx1, tag):
    """One h1mg_solve on a raw (discontinuous) input, against the MGRID
    build of the reference.  The masked rhs it leaves behind is integer
    work (zeroing) and must match bit for bit.  z goes through the FDM
    eigenvectors, which the oracle computes with its own symmetric
    eigensolver instead of LAPACK's DSYGV: z agrees to 1e-13 of max|z|
    (measured 1.3e-15)."""
    P = oracle.Problem(*nel, nx1)
    P.hsmg_setup()
    z, r = P.hsmg_solve(splitmix_uniform(P.n, 11))
    assert np.array_equal(r, golden[tag + "_r"])
    zr = golden[tag + "_z"]
    assert np.abs(z - zr).max() <= 1e-13 * np.abs(zr).max()


def test_hsmg_levels_follow_reference_rule(oracle):
    # h1mg_setup_mg_nx, src/hsmg.f:45-97: orders 1, min(3, mgn2(nx1)), nx1-1
    # (two levels at nx1 = 4)
    expect = {4: [2, 4], 5: [2, 3, 5], 6: [2, 4, 6], 7: [2, 4, 7], 8: [2, 4, 8], 9: [2, 4, 9],
              10: [2, 4, 10]}
    for nx1, nh in expect.items():
        P = oracle.Problem(2, 2, 2, nx1)
        P.hsmg_setup()
        assert P.hsmg_levels() == nh, nx1


def test_hsmg_cg_history_vs_reference(oracle, golden):
    """CG with the HSMG preconditioner and the reference's mask, 4^3
    elements, N=7, 30 iterations.  The 1e-15 differences of z (see above)
    are amplified by the iteration, which with the reference's element-1
    mask list is not a symmetric one: the first ten iterations agree to
    1e-13 and all 30 to 5e-9 (measured 1.1e-9 at iteration 27, where
    rtr has fallen by 4e-9)."""
    P = oracle.Problem(4, 4, 4, 8)
    P.hsmg_setup()
    f = P.rhs_from_raw(splitmix_uniform(P.n, 5))
    _, _, h = P.cg(f, 30, oracle.PRECOND_HSMG)
    ref = golden["mgcg_b444_n8_hist"][0]
    rel = np.abs(h / ref - 1.0)
    assert rel[:31].max() <= 1e-13
    assert rel.max() <= 5e-9
