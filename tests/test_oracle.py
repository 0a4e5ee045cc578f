"""Pins the oracle (oracle/lap_oracle.c): reference KATs, closed forms,
brute-force dense algebra, and bitwise agreement with the reference's own
CPU path compiled from its sources (oracle/_ref)."""
import numpy as np
import pytest
import scipy.linalg as sla

from lap_amd.problems import (CSR, augment_lower, augment_upper, grid_laplacian,
                              rhs_normal, rhs_uniform, split)

KINDS = ["none", "line_jacobi", "it_jacobi", "GS_std", "GS_it", "GS_it2", "ichol"]


from cases import CPU_PROBLEMS, PROBLEMS  # noqa: E402


@pytest.fixture(params=CPU_PROBLEMS)
def problem(request):
    return PROBLEMS[request.param]()


# ---------------------------------------------------------------- KATs

def test_kat_spmv(orc):
    """test_spmv_omp.c:17-37: alpha=-2, beta=1.5, y0=1."""
    A = CSR(5, np.array([0, 3, 4, 7, 9, 10], np.int32),
            np.array([0, 1, 2, 1, 2, 3, 4, 3, 4, 4], np.int32),
            np.array([1.0, 2, 3, 7, 4, 1, 2, 3, 1, 9]))
    y = orc.matvec(A, np.array([1.0, 2, 3, 4, 5]), np.ones(5), alpha=-2.0, beta=1.5)
    np.testing.assert_array_equal(y, [-26.5, -26.5, -50.5, -32.5, -88.5])


def test_kat_upper_solve(orc):
    """test_tri_solve_omp.c:63-85: strict upper, diagonal {1,7,4,3,9}, x=8."""
    U = CSR(5, np.array([0, 2, 2, 4, 5, 5], np.int32), np.array([1, 2, 3, 4, 4], np.int32),
            np.array([2.0, 3, 1, 2, 1]))
    y = orc.upper_solve(U, np.full(5, 8.0), np.array([1.0, 7, 4, 3, 9]))
    np.testing.assert_allclose(y, [2.825397, 1.142857, 0.962963, 2.370370, 0.888889],
                               atol=5e-7)


# --------------------------------------------------------- closed forms

def test_laplacian_kernel_is_constant(orc):
    A = grid_laplacian([6, 5, 4], 0.0)
    assert np.all(orc.matvec(A, np.ones(A.n)) == 0.0)


@pytest.mark.parametrize("dims", [[8, 6], [5, 4, 3]])
def test_grid_cosine_eigenvectors(orc, dims):
    """Neumann grid Laplacian eigenpairs: prod cos(pi k (x+1/2)/N),
    lambda = sum(2 - 2cos(pi k/N))."""
    A = grid_laplacian(dims, 0.0)
    k = [1, 2, 1][: len(dims)]
    grids = np.meshgrid(*[np.arange(d) for d in dims], indexing="ij")
    v = np.ones(dims)
    lam = 0.0
    for g, d, kk in zip(grids, dims, k):
        v = v * np.cos(np.pi * kk * (g + 0.5) / d)
        lam += 2 - 2 * np.cos(np.pi * kk / d)
    v = v.ravel(order="F")  # first grid dimension fastest
    np.testing.assert_allclose(orc.matvec(A, v), lam * v, atol=1e-13)


def test_matvec_dense(orc, problem):
    rng = np.random.default_rng(1)
    x, y0 = rng.standard_normal(problem.n), rng.standard_normal(problem.n)
    y = orc.matvec(problem, x, y0, alpha=0.75, beta=-1.25)
    np.testing.assert_allclose(y, 0.75 * problem.to_dense() @ x - 1.25 * y0, rtol=1e-12,
                               atol=1e-12)


def test_triangular_dense(orc, problem):
    rng = np.random.default_rng(2)
    x = rng.standard_normal(problem.n)
    L, d, U = split(problem)
    D = problem.to_dense()
    np.testing.assert_allclose(orc.lower_solve(L, x, d), sla.solve_triangular(np.tril(D), x,
                               lower=True), rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(orc.upper_solve(U, x, d), sla.solve_triangular(np.triu(D), x),
                               rtol=1e-11, atol=1e-12)
    # the augmented (diagonal inside) form gives the same numbers bitwise
    np.testing.assert_array_equal(orc.lower_solve(augment_lower(problem), x),
                                  orc.lower_solve(L, x, d))
    np.testing.assert_array_equal(orc.upper_solve(augment_upper(problem), x),
                                  orc.upper_solve(U, x, d))


def test_ichol_invariant(orc, problem):
    """(L L^T)_ij == A_ij on the pattern of A."""
    Lf, _ = orc.ichol_factor(problem)
    Ld = Lf.to_dense()
    LLt = Ld @ Ld.T
    D = problem.to_dense()
    mask = D != 0
    np.testing.assert_allclose(LLt[mask], D[mask], rtol=1e-12, atol=1e-12)


def test_ichol_exact_on_tree(orc):
    """No fill on a path graph: IC0 is the exact Cholesky factor, so IC0-PCG
    converges in one iteration."""
    A = grid_laplacian([40], 0.5)
    Lf, _ = orc.ichol_factor(A)
    np.testing.assert_allclose(Lf.to_dense(), np.linalg.cholesky(A.to_dense()), rtol=1e-13,
                               atol=1e-14)
    b = rhs_normal(A.n, 3)
    x, it, flag, _ = orc.cg(A, b, "ichol", tol=1e-12)
    assert (it, flag) == (1, 0)
    np.testing.assert_allclose(A.to_dense() @ x, b, atol=1e-12)


def test_prec_dense(orc, problem):
    """Dense restatements: Jacobi, SGS closed form (D+U)^-1 D (D+L)^-1 for
    GS_std with K=1, it_jacobi with K=1 == line_jacobi, IC0 apply."""
    rng = np.random.default_rng(5)
    r = rng.standard_normal(problem.n)
    D = problem.to_dense()
    dg = np.diag(np.diag(D))
    np.testing.assert_allclose(orc.prec_apply(problem, orc.make_prec(problem, "line_jacobi"), r),
                               r / np.diag(D), rtol=1e-15)
    sgs = np.linalg.solve(np.triu(D), dg @ np.linalg.solve(np.tril(D), r))
    np.testing.assert_allclose(orc.prec_apply(problem, orc.make_prec(problem, "GS_std"), r), sgs,
                               rtol=1e-10, atol=1e-12)
    np.testing.assert_array_equal(
        orc.prec_apply(problem, orc.make_prec(problem, "it_jacobi", k=1), r),
        orc.prec_apply(problem, orc.make_prec(problem, "line_jacobi"), r))
    Lf, _ = orc.ichol_factor(problem)
    Ld = Lf.to_dense()
    np.testing.assert_allclose(orc.prec_apply(problem, orc.make_prec(problem, "ichol"), r),
                               np.linalg.solve(Ld.T, np.linalg.solve(Ld, r)), rtol=1e-10,
                               atol=1e-12)


def test_cg_eigenvector_one_step(orc):
    A = grid_laplacian([8, 6], 0.0)
    grids = np.meshgrid(np.arange(8), np.arange(6), indexing="ij")
    v = (np.cos(np.pi * (grids[0] + 0.5) / 8) * np.cos(np.pi * 2 * (grids[1] + 0.5) / 6))
    v = v.ravel(order="F")
    x, it, flag, hist = orc.cg(A, v, "none", tol=1e-10)
    assert (it, flag) == (1, 0)
    lam = (2 - 2 * np.cos(np.pi / 8)) + (2 - 2 * np.cos(np.pi * 2 / 6))
    np.testing.assert_allclose(x, v / lam, rtol=1e-12)


def test_cg_true_residual(orc, problem):
    b = rhs_normal(problem.n, 9)
    for kind in ("line_jacobi", "GS_std", "ichol"):
        x, it, flag, hist = orc.cg(problem, b, kind, tol=1e-10)
        assert flag == 0
        r = b - problem.to_dense() @ x
        assert np.linalg.norm(r) <= 1e-10 * np.linalg.norm(b) * 1.5


def test_cg_maxit_flag(orc):
    A = grid_laplacian([30, 30], 1e-8)
    b = rhs_normal(A.n, 4)
    x, it, flag, hist = orc.cg(A, b, "line_jacobi", tol=1e-14, maxit=5)
    assert (it, flag) == (6, 1)  # cg.c: stops when iter > maxit
    assert hist.shape == (7,)


def test_omp_matches_serial(orc, problem):
    b = rhs_normal(problem.n, 11)
    x1, it1, _, h1 = orc.cg(problem, b, "line_jacobi", tol=1e-10)
    x2, it2, _, h2 = orc.cg(problem, b, "line_jacobi", tol=1e-10, omp=(4, 4))
    assert abs(it1 - it2) <= 1
    # the +1e-8*I shift leaves the constant vector as a near-null mode, so
    # rounding in b's mean moves x along it by ~1e-9; compare on 1-perp
    x1, x2 = x1 - x1.mean(), x2 - x2.mean()
    assert np.linalg.norm(x1 - x2) <= 1e-9 * np.linalg.norm(x1)


# ------------------------------------------- bitwise vs the reference build

def test_ref_kernels_bitwise(orc, ref, problem):
    import oracle as O
    rng = np.random.default_rng(3)
    x, y0 = rng.standard_normal(problem.n), rng.standard_normal(problem.n)
    A = problem
    y = y0.copy()
    al, be = np.array([-1.0]), np.array([1.0])
    ref.fn("csr_matvec")(A.n, A.nnz, O.ip(A.ia), O.ip(A.ja), O.dp(A.a), O.dp(x), O.dp(y),
                         O.dp(al), O.dp(be))
    np.testing.assert_array_equal(y, orc.matvec(A, x, y0, -1.0, 1.0))
    assert ref.fn("dot")(A.n, O.dp(x), O.dp(y0)) == orc.dot(x, y0)
    L, d, U = split(A)
    for nm, T, mine in (("lower_triangular_solve", L, orc.lower_solve),
                        ("upper_triangular_solve", U, orc.upper_solve)):
        out = np.zeros(A.n)
        ref.fn(nm)(A.n, T.nnz, O.ip(T.ia), O.ip(T.ja), O.dp(T.a), O.dp(d), O.dp(x), O.dp(out))
        np.testing.assert_array_equal(out, mine(T, x, d))


def test_ref_ichol_bitwise(orc, ref, problem):
    pd, keep = ref.make_pdata(problem, "ichol")
    Lf, Uf = orc.ichol_factor(problem)
    np.testing.assert_array_equal(keep["L"].a, Lf.a)
    np.testing.assert_array_equal(keep["U"].a, Uf.a)


@pytest.mark.parametrize("kind", KINDS)
def test_ref_prec_bitwise(orc, ref, problem, kind):
    r = np.random.default_rng(4).standard_normal(problem.n)
    got = orc.prec_apply(problem, orc.make_prec(problem, kind, m=2, k=2), r)
    want = ref.prec_apply(problem, ref.make_pdata(problem, kind, m=2, k=2), r)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("kind", ["none", "line_jacobi", "it_jacobi", "GS_std", "GS_it2", "ichol"])
def test_ref_cg_bitwise(orc, ref, problem, kind):
    b = rhs_uniform(problem.n, 8)
    x1, it1, f1, h1 = orc.cg(problem, b, kind, tol=1e-9, maxit=3000, m=1, k=2)
    x2, it2, f2, h2 = ref.cg(problem, b, kind, tol=1e-9, maxit=3000, m=1, k=2)
    assert (it1, f1) == (it2, f2)
    np.testing.assert_array_equal(h1, h2)
    np.testing.assert_array_equal(x1, x2)
