"""HSMG preconditioner on the device (through the C ABI) against the
MGRID build of the reference (fixtures) and the oracle.

Bars: the masked rhs h1mg_solve leaves behind is integer work and must
match bit for bit; z goes through the FDM eigenvectors, computed here and
in the oracle by a Jacobi eigensolver and in the reference by LAPACK's
DSYGV, so z agrees to 1e-12 of max|z| (the oracle itself agrees with the
reference to 1.3e-15).  CG histories: 1e-12 over the first ten
iterations; later iterations of the reference's masked (non-symmetric)
iteration amplify the 1e-15 differences, bounded at 5e-9 as for the
oracle."""
import numpy as np
import pytest

from conftest import splitmix_uniform

pytestmark = pytest.mark.gpu


def _case(msg):
    print("CASE", msg, flush=True)


@pytest.mark.parametrize("nel,nx1,tag", [((2, 2, 2), 8, "mg_b222_n8"),
                                         ((3, 2, 2), 10, "mg_b322_n10"),
                                         ((4, 4, 4), 8, "mg_b444_n8")])
def test_hsmg_apply_vs_reference(nb, oracle, golden, nel, nx1, tag):
    _case("hsmg apply %s" % tag)
    P = nb.Problem(*nel, nx1)
    P.setup_hsmg()
    u = splitmix_uniform(P.n, 11)
    z, r = P.hsmg_solve(u)
    assert np.array_equal(r, golden[tag + "_r"])
    zr = golden[tag + "_z"]
    assert np.abs(z - zr).max() <= 1e-12 * np.abs(zr).max()
    O = oracle.Problem(*nel, nx1)
    O.hsmg_setup()
    zo, ro = O.hsmg_solve(u)
    assert np.array_equal(r, ro)
    assert np.abs(z - zo).max() <= 1e-12 * np.abs(zo).max()


@pytest.mark.parametrize("nel,nx1", [((2, 3, 2), 4), ((3, 2, 2), 5), ((2, 2, 3), 6),
                                     ((1, 2, 2), 9), ((2, 2, 1), 12)])
def test_hsmg_apply_vs_oracle_other_orders(nb, oracle, nel, nx1):
    """Two levels at nx1 = 4, three-point middle level at nx1 = 5, a single
    element across x, and an order above the reference's table."""
    _case("hsmg apply %s %d" % (nel, nx1))
    P = nb.Problem(*nel, nx1)
    P.setup_hsmg()
    O = oracle.Problem(*nel, nx1)
    O.hsmg_setup()
    u = splitmix_uniform(P.n, 4)
    z, r = P.hsmg_solve(u)
    zo, ro = O.hsmg_solve(u)
    assert np.array_equal(r, ro)
    assert np.abs(z - zo).max() <= 1e-12 * np.abs(zo).max()


def test_hsmg_cg_history_vs_reference(nb, oracle, golden):
    _case("hsmg cg reference history")
    P = nb.Problem(4, 4, 4, 8)
    P.setup_hsmg()
    O = oracle.Problem(4, 4, 4, 8)
    f = O.rhs_from_raw(splitmix_uniform(O.n, 5))
    x, r, h = P.cg(f, 30, nb.PRECOND_HSMG)
    ref = golden["mgcg_b444_n8_hist"][0]
    rel = np.abs(h / ref - 1.0)
    assert rel[:31].max() <= 1e-12
    assert rel.max() <= 5e-9


@pytest.mark.parametrize("deformed", [False, True])
def test_hsmg_cg_dirichlet_vs_oracle(nb, oracle, deformed):
    """Dirichlet mask (a symmetric iteration): the history follows the
    oracle's to 1e-12 relative for 20 iterations and x to 1e-10 of
    max|x|."""
    _case("hsmg cg dirichlet deformed=%s" % deformed)
    nel, nx1 = (4, 3, 3), 8
    P = nb.Problem(*nel, nx1)
    O = oracle.Problem(*nel, nx1)
    P.set_mask_mode(nb.MASK_DIRICHLET)
    O.set_mask_mode(oracle.MASK_DIRICHLET)
    if deformed:
        P.set_geometry_deformed(0.1, 2)
        O.deformed_g(0.1, 2)
    P.setup_hsmg()
    O.hsmg_setup()
    f = O.rhs_from_raw(splitmix_uniform(O.n, 1))
    xo, ro, ho = O.cg(f, 20, oracle.PRECOND_HSMG)
    x, r, h = P.cg(f, 20, nb.PRECOND_HSMG)
    np.testing.assert_allclose(h, ho, rtol=1e-12, atol=0)
    np.testing.assert_allclose(x, xo, rtol=0, atol=1e-10 * np.abs(xo).max())


def test_hsmg_needs_setup_and_order(nb):
    _case("hsmg errors")
    P = nb.Problem(2, 2, 2, 3)
    with pytest.raises(nb.NekboneError):
        P.setup_hsmg()
    Q = nb.Problem(2, 2, 2, 6)
    with pytest.raises(nb.NekboneError):
        Q.cg(np.zeros(Q.n), 5, nb.PRECOND_HSMG)
