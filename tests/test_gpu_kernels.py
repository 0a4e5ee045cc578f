"""Kernel-level parity of the HIP library (called through the C ABI) with
the oracle.  Bar: bitwise wherever the reference's arithmetic is a per-row
sequential recurrence (SpMV rows up to 2048 nonzeros, triangular solves,
IC0, every preconditioner); the tolerance is stated where the GPU sums in a
different order (long-row SpMV, dot products)."""
import numpy as np
import pytest
import torch

from cases import PROBLEMS, nine_point, star_graph
from lap_amd import solver as S
from lap_amd.problems import CSR, augment_lower, augment_upper, chung_lu, grid_laplacian, split

pytestmark = pytest.mark.gpu

KINDS = ["none", "line_jacobi", "it_jacobi", "GS_std", "GS_it", "GS_it2", "ichol"]


@pytest.fixture(params=sorted(PROBLEMS))
def problem(request):
    print(f"[case] problem={request.param}", flush=True)
    return PROBLEMS[request.param]()


def dev(v):
    return S.to_dev(v)


def host(t):
    return t.cpu().numpy()


@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (-1.0, 1.0), (0.75, -1.25)])
def test_spmv_bitwise(orc, problem, alpha, beta):
    rng = np.random.default_rng(1)
    x, y0 = rng.standard_normal(problem.n), rng.standard_normal(problem.n)
    A = S.DeviceCSR.from_csr(problem)
    y = S.spmv(A, dev(x), dev(y0), alpha, beta)
    np.testing.assert_array_equal(host(y), orc.matvec(problem, x, y0, alpha, beta))


def test_spmv_kat():
    A = CSR(5, np.array([0, 3, 4, 7, 9, 10], np.int32),
            np.array([0, 1, 2, 1, 2, 3, 4, 3, 4, 4], np.int32),
            np.array([1.0, 2, 3, 7, 4, 1, 2, 3, 1, 9]))
    y = S.spmv(S.DeviceCSR.from_csr(A), dev(np.array([1.0, 2, 3, 4, 5])), dev(np.ones(5)),
               -2.0, 1.5)
    np.testing.assert_array_equal(host(y), [-26.5, -26.5, -50.5, -32.5, -88.5])


def test_spmv_long_rows(orc):
    """A hub row of 6000 nonzeros is summed by a whole workgroup (tree
    order): within a few ulps of sum |a x|; short rows stay bitwise."""
    H = star_graph(6000)
    rng = np.random.default_rng(2)
    x = rng.standard_normal(H.n)
    y = host(S.spmv(S.DeviceCSR.from_csr(H), dev(x)))
    want = orc.matvec(H, x)
    absrow = orc.matvec(CSR(H.n, H.ia, H.ja, np.abs(H.a)), np.abs(x))
    assert abs(y[0] - want[0]) <= 64 * np.finfo(float).eps * absrow[0]
    np.testing.assert_array_equal(y[1:], want[1:])


@pytest.mark.parametrize("rows,walk", [(64, 0), (512, 0), (1, 0), (256, 3), (16, 3), (512, 3)])
def test_spmv_blocking_invariant(orc, rows, walk):
    """Row-block size and walk order change scheduling, never the bits."""
    P = PROBLEMS["graph_big"]()
    x = np.random.default_rng(3).standard_normal(P.n)
    old = [S.set_option("spmv_rows", rows), S.set_option("spmv_walk", walk)]
    try:
        y = host(S.spmv(S.DeviceCSR.from_csr(P), dev(x)))
    finally:
        for name, v in zip(("spmv_rows", "spmv_walk"), old):
            S.set_option(name, v)
    np.testing.assert_array_equal(y, orc.matvec(P, x))


def test_dot():
    rng = np.random.default_rng(4)
    for n in (1, 63, 1000, 1 << 20):
        v, w = rng.standard_normal(n), rng.standard_normal(n)
        got = S.dot(dev(v), dev(w))
        want = float(np.dot(v, w))
        bound = 4 * np.log2(n + 1) * np.finfo(float).eps * float(np.abs(v) @ np.abs(w))
        assert abs(got - want) <= bound + 1e-300
        assert got == S.dot(dev(v), dev(w))  # deterministic


def test_elementwise():
    from lap_amd import lib as L
    import ctypes as C
    lib = L.load()
    v = np.array([2.0, 0.0, -4.0, 0.25, 9.0])
    dv, out = dev(v), torch.zeros(5, dtype=torch.float64, device="cuda")
    lib.hip_vector_reciprocal(5, S.ptr(dv), S.ptr(out))
    np.testing.assert_array_equal(host(out), [0.5, 0.0, -0.25, 4.0, 1.0 / 9.0])
    lib.hip_vector_sqrt(5, S.ptr(dv), S.ptr(out))
    np.testing.assert_array_equal(host(out), [np.sqrt(2.0), 0.0, 0.0, 0.5, 3.0])
    lib.hip_vec_vec(5, S.ptr(dv), S.ptr(dv), S.ptr(out))
    np.testing.assert_array_equal(host(out), v * v)
    lib.hip_vec_set(5, C.c_double(7.5), S.ptr(out))
    np.testing.assert_array_equal(host(out), np.full(5, 7.5))
    lib.hip_axpy(5, C.c_double(-2.0), S.ptr(dv), S.ptr(out))
    np.testing.assert_array_equal(host(out), 7.5 + -2.0 * v)
    lib.hip_scal(5, C.c_double(0.5), S.ptr(out))
    np.testing.assert_array_equal(host(out), (7.5 + -2.0 * v) * 0.5)
    lib.hip_vec_copy(5, S.ptr(dv), S.ptr(out))
    np.testing.assert_array_equal(host(out), v)
    lib.hip_vec_zero(5, S.ptr(out))
    assert not host(out).any()
    L.check()


def test_device_memory_roundtrip():
    from lap_amd import lib as L
    import ctypes as C
    lib = L.load()
    h = np.arange(1000, dtype=np.float64)
    back = np.zeros_like(h)
    d1 = lib.mallocForDevice(None, 1000, 8)
    d2 = lib.mallocForDevice(None, 1000, 8)
    assert d1 and d2
    hp = h.ctypes.data_as(C.c_void_p)
    bp = back.ctypes.data_as(C.c_void_p)
    lib.memcpyDevice(d1, hp, 1000, 8, b"H2D")
    lib.memcpyDevice(d2, d1, 1000, 8, b"D2D")  # the reference routes D2D as D2H
    lib.memcpyDevice(bp, d2, 1000, 8, b"D2H")
    lib.freeDevice(d1)
    lib.freeDevice(d2)
    L.check()
    np.testing.assert_array_equal(back, h)


def test_trisolve_bitwise(orc, problem):
    rng = np.random.default_rng(5)
    x = rng.standard_normal(problem.n)
    Ls, d, Us = split(problem)
    dx, dd = dev(x), dev(d)
    for T, lower, ref in ((Ls, True, orc.lower_solve), (Us, False, orc.upper_solve)):
        got = host(S.tri_solve(S.DeviceCSR.from_csr(T), dx, lower, diag=dd))
        np.testing.assert_array_equal(got, ref(T, x, d))
    for T, lower, ref in ((augment_lower(problem), True, orc.lower_solve),
                          (augment_upper(problem), False, orc.upper_solve)):
        got = host(S.tri_solve(S.DeviceCSR.from_csr(T), dx, lower))
        np.testing.assert_array_equal(got, ref(T, x))


def test_trisolve_kat():
    U = CSR(5, np.array([0, 2, 2, 4, 5, 5], np.int32), np.array([1, 2, 3, 4, 4], np.int32),
            np.array([2.0, 3, 1, 2, 1]))
    got = host(S.tri_solve(S.DeviceCSR.from_csr(U), dev(np.full(5, 8.0)), False,
                           diag=dev(np.array([1.0, 7, 4, 3, 9]))))
    np.testing.assert_allclose(got, [2.825397, 1.142857, 0.962963, 2.370370, 0.888889],
                               atol=5e-7)


def test_ichol_bitwise(orc, problem):
    A = S.DeviceCSR.from_csr(problem)
    vals = host(S.ichol_factor(A))
    Lf, _ = orc.ichol_factor(problem)
    rows = problem.row_ids()
    low = problem.ja <= rows
    np.testing.assert_array_equal(vals[low], Lf.a)           # factor in the lower triangle
    np.testing.assert_array_equal(vals[~low], problem.a[~low])  # upper triangle untouched


@pytest.mark.parametrize("kind", KINDS)
def test_prec_bitwise(orc, problem, kind):
    r = np.random.default_rng(6).standard_normal(problem.n)
    A = S.DeviceCSR.from_csr(problem)
    P = S.make_preconditioner(A, problem, kind, m=2, k=2)
    got = host(S.prec_apply(A, P, dev(r)))
    want = orc.prec_apply(problem, orc.make_prec(problem, kind, m=2, k=2), r)
    np.testing.assert_array_equal(got, want)


def test_spmv_rejects_bad_columns():
    """Out-of-range column indices are refused before any launch."""
    from lap_amd import lib as L
    bad = CSR(3, np.array([0, 1, 2, 3], np.int32), np.array([0, 5, 2], np.int32), np.ones(3))
    with pytest.raises(L.LapError):
        S.spmv(S.DeviceCSR.from_csr(bad), dev(np.ones(3)))


def test_trisolve_ignores_stale_diagonal_with_inrow_pivots(orc):
    """The reference's GPU driver hands the triangular solves a freed host
    pointer as `diagonal` (cg_driver2.c:246-248, GS.c:31/48) while the pivots
    sit inside L+D / D+U: the solve must ignore it, as rocSPARSE does."""
    import ctypes as C
    from lap_amd import lib as L
    lib = L.load()
    P = PROBLEMS["grid3d"]()
    x = np.random.default_rng(7).standard_normal(P.n)
    host_diag = np.full(P.n, 3.0)  # host memory: never to be read by a kernel
    hp = host_diag.ctypes.data_as(C.c_void_p)
    for T, lower, ref in ((augment_lower(P), True, orc.lower_solve),
                          (augment_upper(P), False, orc.upper_solve)):
        D = S.DeviceCSR.from_csr(T)
        out = torch.zeros(P.n, dtype=torch.float64, device="cuda")
        dx = dev(x)
        torch.cuda.synchronize()
        fn = lib.hip_lower_triangular_solve if lower else lib.hip_upper_triangular_solve
        fn(P.n, D.nnz, S.ptr(D.ia), S.ptr(D.ja), S.ptr(D.a), hp, S.ptr(dx), S.ptr(out))
        L.check()
        np.testing.assert_array_equal(host(out), ref(T, x))


def test_trisolve_refuses_host_diagonal_for_strict_factor():
    """A strict factor needs `diagonal`; a host pointer there is an error
    raised before any launch."""
    import ctypes as C
    from lap_amd import lib as L
    lib = L.load()
    P = PROBLEMS["grid3d"]()
    Ls, _, _ = split(P)
    D = S.DeviceCSR.from_csr(Ls)
    host_diag = np.full(P.n, 3.0)
    out = torch.zeros(P.n, dtype=torch.float64, device="cuda")
    dx = dev(np.ones(P.n))
    torch.cuda.synchronize()
    lib.hip_lower_triangular_solve(P.n, D.nnz, S.ptr(D.ia), S.ptr(D.ja), S.ptr(D.a),
                                   host_diag.ctypes.data_as(C.c_void_p), S.ptr(dx), S.ptr(out))
    with pytest.raises(L.LapError):
        L.check()


def test_entry_points_refuse_host_vectors():
    """Host pointers where device memory is required are refused, not read."""
    import ctypes as C
    from lap_amd import lib as L
    lib = L.load()
    h = np.ones(100)
    hp = h.ctypes.data_as(C.c_void_p)
    d = dev(np.ones(100))
    lib.hip_axpy(100, C.c_double(1.0), hp, S.ptr(d))
    with pytest.raises(L.LapError):
        L.check()
    lib.hip_dot(100, S.ptr(d), hp)
    with pytest.raises(L.LapError):
        L.check()


@pytest.fixture(params=[0, 3], ids=["levels", "chunked"])
def tri_mode(request):
    """Triangular solves for the duration of a test: one launch per level
    (mode 0) or the chunked persistent solve (mode 3, where the rows fit its
    records: the IC0 factor's triangles)."""
    prev = S.set_option("sptrsv_mode", request.param)
    yield request.param
    S.set_option("sptrsv_mode", prev)


@pytest.mark.parametrize("kind", ["GS_std", "ichol"])
def test_prec_tri_modes_bitwise(orc, problem, kind, tri_mode):
    print(f"[case] prec {kind} n={problem.n} sptrsv_mode={tri_mode}", flush=True)
    test_prec_bitwise(orc, problem, kind)


@pytest.mark.parametrize("kind", ["GS_std", "ichol"])
def test_prec_nine_point_bitwise(orc, kind, tri_mode):
    """The 9-point stencil's triangles hold rows of 5 entries: packed 8 wide
    for the level launches, and chunked (4 off-diagonal codes) in mode 3."""
    test_prec_bitwise(orc, nine_point(90, 70), kind)


LONG_ROWS = {"power_law": lambda: chung_lu(16384, seed=7), "star": lambda: star_graph(6000),
             # hub rows at the long-row boundaries: 1024 entries (one tree segment),
             # 1025 (two, the second holding one entry), 2049 (three; and two LDS
             # chunks of the sequential path, the second holding one entry)
             "star_1024": lambda: star_graph(1024), "star_1025": lambda: star_graph(1025),
             "star_2049": lambda: star_graph(2049)}


@pytest.mark.parametrize("name", sorted(LONG_ROWS))
def test_trisolve_long_rows_bitwise(orc, name):
    """Rows of more than 32 entries take a workgroup each (the power-law
    graph's U rows reach 5417 entries: three LDS chunks; the star's hub row
    5999); the subtractions keep the sequential order, so the solves, the
    IC0 factor and the IC0 preconditioner stay bitwise.  GS_std's backward
    sweeps start from r - A w, whose rows of more than 2048 entries are
    summed in tree order: there the result is within a few ulps of the
    terms' magnitudes."""
    A = LONG_ROWS[name]()
    print(f"[case] long rows {name} n={A.n} max row {np.diff(A.ia).max()}", flush=True)
    test_trisolve_bitwise(orc, A)
    test_ichol_bitwise(orc, A)
    test_prec_bitwise(orc, A, "ichol")
    r = np.random.default_rng(6).standard_normal(A.n)
    D = S.DeviceCSR.from_csr(A)
    for k in (1, 2):
        P = S.make_preconditioner(D, A, "GS_std", m=k, k=k)
        got = host(S.prec_apply(D, P, dev(r)))
        want = orc.prec_apply(A, orc.make_prec(A, "GS_std", m=k, k=k), r)
        if np.diff(A.ia).max() <= 2048:
            np.testing.assert_array_equal(got, want)
        else:
            np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * np.abs(want).max())


@pytest.fixture
def tree_long_rows():
    prev = S.set_option("sptrsv_long", 1)
    yield
    S.set_option("sptrsv_long", prev)


@pytest.mark.parametrize("name", sorted(LONG_ROWS))
def test_trisolve_long_rows_tree_tolerance(orc, name, tree_long_rows):
    """Option sptrsv_long 1 sums a long row's products in a fixed tree order,
    rows of more than 1024 entries split over several workgroups (the
    power-law graph's 5417-entry rows and the star's 5999-entry hub: 6
    segments, finished by the last to arrive on the row's ticket): not the
    sequential order, so compared within a tolerance -- each solve within
    1e-12 of the sequential solve relative to the solution's size, short rows
    untouched; IC0 PCG on the power-law graph within one iteration and 1e-10
    of the oracle's solution (many solves in a row: the tickets reset), and
    the mixed-precision IC0 solves to the mixed solver's bars."""
    A = LONG_ROWS[name]()
    print(f"[case] tree long rows {name} n={A.n}", flush=True)
    rng = np.random.default_rng(5)
    x = rng.standard_normal(A.n)
    Ls, d, Us = split(A)
    for T, lower, ref in ((augment_lower(A), True, orc.lower_solve),
                          (augment_upper(A), False, orc.upper_solve)):
        got = host(S.tri_solve(S.DeviceCSR.from_csr(T), dev(x), lower))
        want = ref(T, x)
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * np.abs(want).max())
    if name == "power_law":
        b = rng.standard_normal(A.n)
        b -= b.mean()
        D = S.DeviceCSR.from_csr(A)
        P = S.make_preconditioner(D, A, "ichol")
        res = S.pcg(D, dev(b), P, tol=1e-10, maxit=2000)
        xo, ito, _, _ = orc.cg(A, b, "ichol", tol=1e-10, maxit=2000)
        err = np.linalg.norm(host(res.x) - xo) / np.linalg.norm(xo)
        print(f"[diag] tree IC0 PCG: it {res.iterations} vs oracle {ito}, err {err:.2e}",
              flush=True)
        assert res.flag == 0 and abs(res.iterations - ito) <= 1
        assert err <= 1e-10
        # the mixed-precision inner solves take the same split rows with FP32
        # values (FP32 or FP64 inner vectors); refined to the same tolerance,
        # with the mixed solver's own bars: FP64 true residual within tol,
        # solution within 1e3 tol of the oracle's
        for mv in (0, 1):
            prev = S.set_option("mixed_vectors", mv)
            try:
                m = S.pcg_mixed(D, dev(b), P, tol=1e-10, maxit=4000)
            finally:
                S.set_option("mixed_vectors", prev)
            xm = host(m.x)
            true = np.linalg.norm(b - orc.matvec(A, xm)) / np.linalg.norm(b)
            merr = np.linalg.norm(xm - xo) / np.linalg.norm(xo)
            print(f"[diag] tree mixed IC0 (vectors {mv}): it {m.iterations}, true rel res "
                  f"{true:.2e}, err {merr:.2e}", flush=True)
            assert m.flag == 0 and true < 1e-10 * 1.01 and merr <= 1e3 * 1e-10
