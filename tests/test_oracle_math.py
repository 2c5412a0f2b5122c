"""Pins of the oracle that do not depend on the reference: closed forms,
invariants, brute force and independent library routines.  Each check is
chosen so that a plausible slip (wrong index, transposed operand, dropped
term, wrong sign) fails it."""
import numpy as np
import pytest
from numpy.polynomial import legendre as L

from conftest import splitmix_uniform


@pytest.mark.parametrize("N", range(1, 16))
def test_gll_rule_closed_forms(oracle, N):
    z, w = oracle.zwgll(N + 1)
    # interior nodes are the roots of P_N' (independent: numpy's Legendre)
    roots = np.sort(L.Legendre.basis(N).deriv().roots().real) if N > 1 else np.array([])
    assert z[0] == -1.0 and z[-1] == 1.0
    np.testing.assert_allclose(z[1:-1], roots, rtol=0, atol=1e-14)
    # w_j = 2 / (N(N+1) P_N(z_j)^2)
    PN = L.legval(z, [0] * N + [1])
    np.testing.assert_allclose(w, 2.0 / (N * (N + 1) * PN ** 2), rtol=1e-13)
    assert abs(w.sum() - 2.0) < 1e-14
    # exact for degree <= 2N-1
    for k in range(2 * N):
        exact = 2.0 / (k + 1) if k % 2 == 0 else 0.0
        assert abs(np.dot(w, z ** k) - exact) < 1e-13


@pytest.mark.parametrize("N", range(1, 16))
def test_derivative_matrix_closed_forms(oracle, N):
    a, b, c, d, z = oracle.semhat(N)
    assert abs(d[0, 0] + N * (N + 1) / 4.0) < 1e-12 * N * N
    assert abs(d[N, N] - N * (N + 1) / 4.0) < 1e-12 * N * N
    # D applied to z^k gives k z^(k-1) for k <= N; D.1 = 0
    for k in range(N + 1):
        expect = k * z ** (k - 1) if k > 0 else np.zeros_like(z)
        np.testing.assert_allclose(d @ z ** k, expect, rtol=0, atol=1e-11 * max(1, N) ** 2)
    # off-diagonal closed form D_ij = P_N(z_i) / (P_N(z_j) (z_i - z_j))
    PN = L.legval(z, [0] * N + [1])
    for i in range(N + 1):
        for j in range(N + 1):
            if i != j:
                assert abs(d[i, j] - PN[i] / (PN[j] * (z[i] - z[j]))) < 1e-11 * N * N
    # A-hat = D^T B D is symmetric with null space 1; C = B D
    np.testing.assert_allclose(a, d.T @ np.diag(b) @ d, rtol=1e-14, atol=1e-13)
    np.testing.assert_allclose(a @ np.ones(N + 1), 0, atol=1e-12 * N * N)
    np.testing.assert_allclose(c, np.diag(b) @ d, rtol=0, atol=1e-15 * N * N)


@pytest.mark.parametrize("np_,expect", [(512, (8, 8, 8)), (32768, (32, 32, 32)),
                                        (4096, (16, 16, 16)), (12, (3, 2, 2)), (13, (13, 1, 1)),
                                        (64, (4, 4, 4)), (50, (5, 5, 2)), (1, (1, 1, 1))])
def test_cubic(oracle, np_, expect):
    assert oracle.cubic(np_) == expect


@pytest.mark.parametrize("nel,nx1", [((2, 2, 2), 4), ((3, 1, 2), 5), ((1, 1, 1), 3)])
def test_numbering_multiplicity_and_dssum_brute_force(oracle, nel, nx1):
    P = oracle.Problem(*nel, nx1)
    glo = P.glo()
    N = nx1 - 1
    nu = (N * nel[0] + 1) * (N * nel[1] + 1) * (N * nel[2] + 1)
    assert glo.min() == 1 and glo.max() == nu and np.unique(glo).size == nu
    mult = np.bincount(glo)[glo]
    np.testing.assert_array_equal(P.c(), 1.0 / mult)
    # integer-valued data: the sum is exact in any order
    u = np.round(splitmix_uniform(P.n, 5) * 100)
    np.testing.assert_array_equal(P.dssum(u), np.bincount(glo, weights=u)[glo])
    off, idx = P.gs_map()
    assert off.size - 1 == np.sum(np.bincount(glo) >= 2)
    # members of a group ascend, groups ordered by smallest member
    firsts = idx[off[:-1]]
    assert np.all(np.diff(firsts) > 0)
    for g in range(off.size - 1):
        assert np.all(np.diff(idx[off[g]:off[g + 1]]) > 0)


def test_dssum_known_answer_two_copies(oracle):
    """jl's gs known answer at one rank (src/jl/gs_test_old.c:78-88): two
    local copies of one id holding 0 and 1 both become 1."""
    P = oracle.Problem(2, 1, 1, 2)  # 2 linear elements sharing one face
    glo = P.glo()
    u = np.zeros(P.n)
    shared = np.where(np.bincount(glo)[glo] == 2)[0]
    first, second = shared[: shared.size // 2], shared[shared.size // 2:]
    u[second] = 1.0
    out = P.dssum(u)
    assert np.all(out[shared] == 1.0)
    assert np.all(out[np.bincount(glo)[glo] == 1] == 0.0)
    assert first.size == second.size == 4


def _dense_AL(oracle, P, e):
    """A_L of element e by applying ax_e to unit vectors (columns)."""
    m = P.nx1 ** 3
    A = np.empty((m, m))
    for j in range(m):
        v = np.zeros(m)
        v[j] = 1.0
        A[:, j] = P.ax_e(e, v)
    return A


def _kron_derivs(D):
    n = D.shape[0]
    I = np.eye(n)
    # local index l = i + n*j + n*n*k: D acts on i (r), j (s) or k (t)
    return np.kron(I, np.kron(I, D)), np.kron(I, np.kron(D, I)), np.kron(D, np.kron(I, I))


@pytest.mark.parametrize("nx1", [2, 3, 4, 6])
def test_ax_e_matches_dense_assembly_brick(oracle, nx1):
    P = oracle.Problem(1, 1, 1, nx1)
    D = P.dxm1().reshape(nx1, nx1, order="F")
    z, w = oracle.zwgll(nx1)
    W = np.diag(np.kron(w, np.kron(w, w)))
    Dr, Ds, Dt = _kron_derivs(D)
    A = Dr.T @ W @ Dr + Ds.T @ W @ Ds + Dt.T @ W @ Dt
    Ae = _dense_AL(oracle, P, 0)
    np.testing.assert_allclose(Ae, A, rtol=0, atol=1e-13 * np.abs(A).max())
    # separable closed form A = Ah(x)B(x)B + B(x)Ah(x)B + B(x)B(x)Ah
    a, b, c, d, zz = oracle.semhat(nx1 - 1)
    B = np.diag(b)
    sep = np.kron(B, np.kron(B, a)) + np.kron(B, np.kron(a, B)) + np.kron(a, np.kron(B, B))
    np.testing.assert_allclose(Ae, sep, rtol=0, atol=1e-13 * np.abs(sep).max())


@pytest.mark.parametrize("nx1", [3, 4, 5])
def test_ax_e_matches_dense_assembly_deformed(oracle, nx1):
    P = oracle.Problem(2, 2, 1, nx1)
    P.deformed_g(0.1, 9)
    g = P.g().reshape(-1, 6)
    D = P.dxm1().reshape(nx1, nx1, order="F")
    Ds = _kron_derivs(D)
    m = nx1 ** 3
    for e in range(P.nelt):
        ge = g[e * m:(e + 1) * m]
        Gi = {(0, 0): 0, (0, 1): 1, (0, 2): 2, (1, 1): 3, (1, 2): 4, (2, 2): 5}
        A = np.zeros((m, m))
        for a in range(3):
            for b_ in range(3):
                key = (min(a, b_), max(a, b_))
                A += Ds[a].T @ np.diag(ge[:, Gi[key]]) @ Ds[b_]
        Ae = _dense_AL(oracle, P, e)
        scale = np.abs(A).max()
        np.testing.assert_allclose(Ae, A, rtol=0, atol=1e-13 * scale)
        np.testing.assert_allclose(Ae, Ae.T, rtol=0, atol=1e-13 * scale)
        np.testing.assert_allclose(Ae @ np.ones(m), 0, atol=1e-13 * scale)
        # the matrix of a mapped Laplacian is positive semidefinite
        assert np.linalg.eigvalsh(0.5 * (Ae + Ae.T)).min() > -1e-12 * scale


def test_deformed_reduces_to_brick(oracle):
    P = oracle.Problem(3, 2, 2, 6)
    g0 = P.g()
    P.deformed_g(0.0, 123)
    assert np.array_equal(P.g(), g0)


def test_deformed_geometry_is_a_valid_map(oracle):
    """Six nonzero factors, positive Jacobian, G symmetric positive
    definite at every point."""
    P = oracle.Problem(3, 3, 3, 5)
    P.deformed_g(0.1, 1)
    g = P.g().reshape(-1, 6)
    assert np.all(g != 0.0)
    G = np.stack([np.stack([g[:, 0], g[:, 1], g[:, 2]], -1), np.stack([g[:, 1], g[:, 3], g[:, 4]], -1),
                  np.stack([g[:, 2], g[:, 4], g[:, 5]], -1)], -2)
    assert np.linalg.eigvalsh(G).min() > 0


def test_jacobi_diagonal_brute_force(oracle):
    """diag of the assembled operator: apply ax to Q e_g (all copies of one
    global node set to 1) and read the value at a copy."""
    for deformed in (False, True):
        P = oracle.Problem(2, 2, 1, 4)
        if deformed:
            P.deformed_g(0.1, 4)
        P.set_mask_mode(oracle.MASK_DIRICHLET)
        P.setup_jacobi()
        dinv = P.dinv()
        dl = P.jacobi_diag_local()
        glo = P.glo()
        full, _ = P.mask_lists()
        interior = np.setdiff1d(np.arange(P.n), full)
        for l in interior[:: max(1, interior.size // 40)]:
            v = (glo == glo[l]).astype(float)
            Av = P.ax(v)
            np.testing.assert_allclose(Av[l], 1.0 / dinv[l], rtol=1e-13)
        # local diagonal equals the diagonal of the dense element matrix
        m = P.nx1 ** 3
        Ae = _dense_AL(oracle, P, 1)
        np.testing.assert_allclose(dl[m:2 * m], np.diag(Ae), rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("precond", [0, 1])
def test_cg_converges_to_dense_solution(oracle, precond):
    """Dirichlet CG on a tiny mesh against numpy.linalg.solve of the
    assembled interior system K = Q^T A_L Q + 0.1 I."""
    P = oracle.Problem(2, 2, 1, 4)
    P.deformed_g(0.05, 2)
    P.set_mask_mode(oracle.MASK_DIRICHLET)
    if precond:
        P.setup_jacobi()
    glo = P.glo() - 1
    ng = glo.max() + 1
    full, _ = P.mask_lists()
    bnd = np.unique(glo[full])
    inner = np.setdiff1d(np.arange(ng), bnd)
    m = P.nx1 ** 3
    K = np.zeros((ng, ng))
    for e in range(P.nelt):
        Ae = _dense_AL(oracle, P, e)
        ids = glo[e * m:(e + 1) * m]
        K[np.ix_(ids, ids)] += Ae
    K += 0.1 * np.eye(ng)
    f = P.rhs_from_raw(splitmix_uniform(P.n, 1))
    fg = np.zeros(ng)
    fg[glo] = f
    xg = np.zeros(ng)
    xg[inner] = np.linalg.solve(K[np.ix_(inner, inner)], fg[inner])
    x, r, h = P.cg(f, 120, precond)
    np.testing.assert_allclose(x, xg[glo], rtol=0, atol=1e-10 * np.abs(xg).max())
    # glsc3 history: rtr decreases overall and ends tiny
    assert h[-1] < 1e-20 * h[0]


def test_rand_uniform_splitmix64_known_answer(oracle):
    """The synthetic-input generator is SplitMix64 (Steele, Lea, Flood,
    OOPSLA 2014; Vigna's reference splitmix64.c): value k of seed s is the
    (k+1)-th output of a generator whose state starts at s.  Published
    first outputs for state 0 pin the constants and the shift order; the
    mapping to U[-1,1) uses the top 53 bits."""
    known = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    expect = np.array([2.0 * ((v >> 11) * 2.0 ** -53) - 1.0 for v in known])
    assert np.array_equal(oracle.rand_uniform(3, 0), expect)
    # a seed shifts the state: value k of seed s equals value k+1 of seed
    # s - 0x9E3779B97F4A7C15 (mod 2^64)
    s = 12345
    prev = (s - 0x9E3779B97F4A7C15) % (1 << 64)
    assert np.array_equal(oracle.rand_uniform(4, s)[:3], oracle.rand_uniform(4, prev)[1:4])
    assert np.array_equal(splitmix_uniform(1000, 7), oracle.rand_uniform(1000, 7))


def _deformed_nodes(oracle, P, amp, seed):
    """Physical coordinates (n, 3) of every GLL node of the deformed box,
    computed here: vertex (a, b, c) of the width-2 grid moved by
    2 amp U (three values per vertex of the seeded generator, vertex
    index a + vx (b + vy c)), and each element mapped trilinearly from its
    8 vertices (lexicographic elements, x fastest)."""
    vx, vy, vz = P.nelx + 1, P.nely + 1, P.nelz + 1
    U = oracle.rand_uniform(3 * vx * vy * vz, seed).reshape(vz, vy, vx, 3)
    c, b, a = np.meshgrid(np.arange(vz), np.arange(vy), np.arange(vx), indexing="ij")
    V = 2.0 * np.stack([a, b, c], -1) + (2.0 * amp) * U          # V[c, b, a, d]
    nx = P.nx1
    z, _ = oracle.zwgll(nx)
    h = np.stack([0.5 * (1.0 - z), 0.5 * (1.0 + z)])            # h[corner, node]
    X = np.empty((P.nelt, nx, nx, nx, 3))                       # X[e, k, j, i, d]
    for e in range(P.nelt):
        ex, ey, ez = e % P.nelx, (e // P.nelx) % P.nely, e // (P.nelx * P.nely)
        C = V[ez:ez + 2, ey:ey + 2, ex:ex + 2]                   # C[cc, bb, aa, d]
        X[e] = np.einsum("ck,bj,ai,cbad->kjid", h, h, h, C)
    return X.reshape(-1, 3)


def _g_from_jacobian(oracle, P, X, transpose=False):
    """g (n, 6) from a Jacobian found by spectral differentiation of the
    node coordinates (exact: the map is trilinear) and numpy's inverse:
    G = w_i w_j w_k det(J) J^-1 J^-T, factors (rr, rs, rt, ss, st, tt)."""
    nx = P.nx1
    D = P.dxm1().reshape(nx, nx, order="F")
    _, w = oracle.zwgll(nx)
    Xe = X.reshape(P.nelt, nx, nx, nx, 3)
    dr = np.einsum("im,ekjmd->ekjid", D, Xe)
    ds = np.einsum("jm,ekmid->ekjid", D, Xe)
    dt = np.einsum("km,emjid->ekjid", D, Xe)
    J = np.stack([dr, ds, dt], -1)                              # J[..., d, a] = dx_d / dr_a
    if transpose:
        J = np.swapaxes(J, -1, -2)
    Ji = np.linalg.inv(J)
    W = np.einsum("k,j,i->kji", w, w, w)[None, :, :, :]
    G = (W * np.linalg.det(J))[..., None, None] * (Ji @ np.swapaxes(Ji, -1, -2))
    G = G.reshape(-1, 3, 3)
    return np.stack([G[:, 0, 0], G[:, 0, 1], G[:, 0, 2], G[:, 1, 1], G[:, 1, 2], G[:, 2, 2]], -1)


@pytest.mark.parametrize("nel,nx1", [((3, 2, 2), 3), ((2, 2, 2), 5), ((2, 1, 2), 8),
                                     ((1, 2, 1), 10)])
def test_deformed_g_against_numpy_jacobian(oracle, nel, nx1):
    """The oracle's deformed factors against an independent computation
    (node coordinates, spectral Jacobian, numpy inverse and determinant).
    A transposed Jacobian gives different factors."""
    P = oracle.Problem(*nel, nx1)
    P.deformed_g(0.1, 17)
    X = _deformed_nodes(oracle, P, 0.1, 17)
    g = P.g().reshape(-1, 6)
    gn = _g_from_jacobian(oracle, P, X)
    scale = np.abs(g).max()
    np.testing.assert_allclose(g, gn, rtol=0, atol=1e-12 * scale)
    assert np.abs(_g_from_jacobian(oracle, P, X, transpose=True) - g).max() > 1e-3 * scale


@pytest.mark.parametrize("nel,nx1", [((3, 2, 2), 3), ((2, 2, 2), 5), ((2, 1, 2), 8),
                                     ((1, 2, 1), 10), ((2, 2, 1), 12)])
def test_deformed_operator_annihilates_linear_functions(oracle, nel, nx1):
    """For u = a . x linear in the physical coordinates, the element
    operator's rows at element-interior nodes are zero: the exact row is
    the integral of grad(phi_i) . a over the element, which vanishes since
    phi_i is zero on the element's faces, and the GLL rule is exact here
    for N >= 2 (det(J) J^-1 = cof(J)^T has degree 2 in the differentiated
    coordinate and 1 in the others for a trilinear map, and the Piola
    identity makes its rows divergence-free).  A transposed Jacobian in
    the factors breaks this by orders of magnitude."""
    P = oracle.Problem(*nel, nx1)
    P.deformed_g(0.1, 23)
    X = _deformed_nodes(oracle, P, 0.1, 23)
    u = X @ np.array([0.3, -0.7, 1.1])
    n3 = nx1 ** 3
    inner = np.zeros((nx1, nx1, nx1), bool)
    inner[1:-1, 1:-1, 1:-1] = True
    inner = inner.reshape(-1)

    def rows(Q):
        w = np.concatenate([Q.ax_e(e, u[e * n3:(e + 1) * n3]) for e in range(Q.nelt)])
        return w.reshape(Q.nelt, n3)

    w = rows(P)
    scale = np.abs(w).max()
    assert scale > 1e-2
    assert np.abs(w[:, inner]).max() <= 1e-12 * scale
    # the same check with the factors of a transposed Jacobian (a planted
    # mistake, on the oracle) fails
    Q = oracle.Problem(*nel, nx1)
    Q.set_g(_g_from_jacobian(oracle, Q, X, transpose=True).reshape(-1))
    assert np.abs(rows(Q)[:, inner]).max() > 1e-4 * scale


def test_glsc3_exact_is_correctly_rounded(oracle):
    """orc_glsc3_exact against math.fsum (exactly rounded) of the same
    rounded terms (a b) m, on sums with heavy cancellation; the sequential
    glsc3 is measurably off on the same data."""
    import math
    rng = np.random.default_rng(5)
    worst_seq = 0.0
    for n in (1, 7, 1000, 200003):
        a = rng.standard_normal(n) * 10.0 ** rng.integers(-8, 8, n)
        b = rng.standard_normal(n)
        m = rng.choice([1.0, 0.5, 0.25, 0.125], n)
        t = (a * b) * m
        exact = math.fsum(t)
        got = oracle.glsc3_exact(a, b, m)
        assert abs(got - exact) <= 2.0 * np.spacing(abs(exact)) + 1e-300
        worst_seq = max(worst_seq, abs(oracle.glsc3(a, b, m) - exact) / np.sum(np.abs(t)))
    assert worst_seq > 1e-17
