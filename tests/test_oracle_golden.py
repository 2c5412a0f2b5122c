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
