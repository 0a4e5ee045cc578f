"""Value-indexed CSR in the PCG SpMV (option spmv_format 1, vi.hip) and
delta-indexed CSR (spmv_format 2: the column too, as a one-byte index into
the matrix's distinct column offsets col - row).

Bars: with the values stored as one-byte indices into a table of the
matrix's distinct doubles (and, format 2, the columns as indices into the
table of offsets), every PCG solve is bitwise the CSR solve (x, residual
history, iteration count) -- for Jacobi and for IC0 and GS_std, whose A
products use it too; a matrix with
255 distinct values is encoded, one with 256 is not (the solve then uses the
CSR values, again bitwise the same); format 2 falls back to format 1 when the
offsets do not fit (a random graph) or a row is longer than 64 entries.
"""
import numpy as np
import pytest
import torch

from cases import PROBLEMS, star_graph, weighted_graph
from lap_amd import solver as S
from lap_amd.problems import from_coo, grid_laplacian, rhs_normal

pytestmark = pytest.mark.gpu


def bits(t):
    return t.cpu().numpy().view(np.int64)


def graph_with_k_values(n, k, seed):
    """Random graph Laplacian whose off-diagonal weights take k distinct
    values; the diagonal (degree sums + shift) adds further values, so the
    total distinct count is reported by the caller."""
    rng = np.random.default_rng(seed)
    m = 6 * n
    r, c = rng.integers(0, n, m), rng.integers(0, n, m)
    keep = r != c
    r, c = np.minimum(r[keep], c[keep]), np.maximum(r[keep], c[keep])
    key = np.unique(r.astype(np.int64) * n + c)
    r, c = key // n, key % n
    levels = 1.0 + np.arange(k) / 8.0
    w = levels[rng.integers(0, k, r.shape[0])]
    deg = np.bincount(r, w, n) + np.bincount(c, w, n)
    rows = np.concatenate([c, np.arange(n)])
    cols = np.concatenate([r, np.arange(n)])
    return from_coo(n, rows, cols, np.concatenate([-w, deg + 1e-3]), symmetric=True)


def solve_both(Ah, kind, fmt=1):
    A = S.DeviceCSR.from_csr(Ah)
    P = S.make_preconditioner(A, Ah, kind)
    b = S.to_dev(rhs_normal(Ah.n, 21))
    try:
        S.set_option("spmv_format", 0)
        r0 = S.pcg(A, b, P, tol=1e-10, maxit=5000)
        S.set_option("spmv_format", fmt)
        r1 = S.pcg(A, b, P, tol=1e-10, maxit=5000)
    finally:
        S.set_option("spmv_format", 0)
    return r0, r1


def same(r0, r1):
    return (r0.iterations == r1.iterations and r0.flag == r1.flag
            and np.array_equal(bits(r0.x), bits(r1.x))
            and np.array_equal(r0.history.view(np.int64), r1.history.view(np.int64)))


@pytest.mark.parametrize("fmt", [1, 2])
@pytest.mark.parametrize("pname", ["grid3d_16", "grid2d"])
def test_value_indexed_jacobi_pcg_bitwise(pname, fmt):
    r0, r1 = solve_both(PROBLEMS[pname](), "line_jacobi", fmt)
    print(f"[diag] format {fmt} {pname}: {r1.iterations} it, csr {r0.gpu_ms:.2f} ms, "
          f"indexed {r1.gpu_ms:.2f} ms", flush=True)
    assert same(r0, r1)


@pytest.mark.parametrize("fmt", [1, 2])
@pytest.mark.parametrize("kind", ["ichol", "GS_std"])
def test_value_indexed_other_preconditioners_bitwise(kind, fmt):
    r0, r1 = solve_both(PROBLEMS["grid3d_16"](), kind, fmt=fmt)
    assert same(r0, r1)


@pytest.mark.parametrize("pname", ["graph", "star"])
def test_delta_indexed_falls_back(pname):
    """A random graph has thousands of distinct offsets, the star a row of
    6000 entries: format 2 is not built and the solve is still bitwise the
    CSR one."""
    Ah = PROBLEMS["graph"]() if pname == "graph" else star_graph(6000)
    r0, r1 = solve_both(Ah, "line_jacobi", fmt=2)
    assert same(r0, r1)


def test_delta_indexed_offset_limit():
    """A banded matrix with exactly 255 distinct offsets is encoded (rows of
    at most 64 entries), and its solve is bitwise the CSR solve."""
    n = 4000
    offs = np.concatenate([-np.arange(1, 128) * 3, [0], np.arange(1, 128) * 3])  # 255
    rng = np.random.default_rng(3)
    rows, cols = [], []
    for i in range(n):
        pick = offs[rng.choice(offs.shape[0], 24, replace=False)]
        pick = pick[(i + pick >= 0) & (i + pick < n)]
        rows.append(np.full(pick.shape[0], i))
        cols.append(i + pick)
    r, c = np.concatenate(rows), np.concatenate(cols)
    keep = r < c
    r, c = r[keep], c[keep]
    key = np.unique(r.astype(np.int64) * n + c)
    r, c = key // n, key % n
    deg = np.bincount(r, minlength=n) + np.bincount(c, minlength=n)
    Ah = from_coo(n, np.concatenate([c, np.arange(n)]), np.concatenate([r, np.arange(n)]),
                  np.concatenate([-np.ones(r.shape[0]), deg + 1e-2]), symmetric=True)
    d = np.unique(Ah.ja.astype(np.int64) - Ah.row_ids())
    print(f"[diag] offsets {d.shape[0]}, longest row {np.diff(Ah.ia).max()}", flush=True)
    assert d.shape[0] <= 255 and np.diff(Ah.ia).max() <= 64
    r0, r1 = solve_both(Ah, "line_jacobi", fmt=2)
    assert same(r0, r1)


def test_value_indexed_table_limit():
    # count the distinct values exactly, then build one matrix at the limit
    # and one above it
    for k in range(60, 0, -1):
        Ah = graph_with_k_values(3000, k, 5)
        if np.unique(Ah.a.view(np.int64)).shape[0] <= 255:
            break
    under = np.unique(Ah.a.view(np.int64)).shape[0]
    over = weighted_graph(3000, 8, 9)  # lognormal weights: thousands of values
    assert np.unique(over.a.view(np.int64)).shape[0] > 255
    print(f"[diag] vi table: {under} distinct values encoded", flush=True)
    for M in (Ah, over):
        r0, r1 = solve_both(M, "line_jacobi")
        assert same(r0, r1)


@pytest.mark.parametrize("fmt", [1, 2])
def test_value_indexed_spmv_timed_kernel(fmt):
    Ah = grid_laplacian([64, 64, 64], 1e-8)
    A = S.DeviceCSR.from_csr(Ah)
    x = torch.randn(A.n, dtype=torch.float64, device="cuda")
    y0, y1 = torch.empty_like(x), torch.empty_like(x)
    try:
        S.set_option("spmv_format", 0)
        t0 = S.spmv_timed(A, x, y0, 5, fused=True)
        S.set_option("spmv_format", fmt)
        t1 = S.spmv_timed(A, x, y1, 5, fused=True)
    finally:
        S.set_option("spmv_format", 0)
    assert t0 > 0 and t1 > 0
    assert torch.equal(y0, y1)


@pytest.mark.parametrize("vectors", [0, 1])
@pytest.mark.parametrize("fmt", [1, 2])
@pytest.mark.parametrize("kind", ["line_jacobi", "ichol"])
def test_value_indexed_mixed_precision_bitwise(kind, fmt, vectors):
    """The mixed-precision inner SpMV over the value-indexed (and, format 2,
    delta-indexed) FP32 copy gives the FP32-CSR solve's bits, with FP32 and
    with FP64 inner vectors."""
    Ah = PROBLEMS["grid3d_16"]()
    A = S.DeviceCSR.from_csr(Ah)
    P = S.make_preconditioner(A, Ah, kind)
    b = S.to_dev(rhs_normal(Ah.n, 22))
    prev = S.set_option("mixed_vectors", vectors)
    try:
        S.set_option("spmv_format", 0)
        m0 = S.pcg_mixed(A, b, P, tol=1e-10, maxit=5000)
        S.set_option("spmv_format", fmt)
        m1 = S.pcg_mixed(A, b, P, tol=1e-10, maxit=5000)
    finally:
        S.set_option("spmv_format", 0)
        S.set_option("mixed_vectors", prev)
    assert (m0.iterations, m0.refinements, m0.flag) == (m1.iterations, m1.refinements, m1.flag)
    assert np.array_equal(bits(m0.x), bits(m1.x))


def spmv_both(M, fmt):
    A = S.DeviceCSR.from_csr(M)
    x = torch.randn(M.n, dtype=torch.float64, device="cuda")
    y0, y1 = torch.empty_like(x), torch.empty_like(x)
    try:
        S.set_option("spmv_format", 0)
        S.spmv_timed(A, x, y0, 2, fused=True)
        S.set_option("spmv_format", fmt)
        S.spmv_timed(A, x, y1, 2, fused=True)
    finally:
        S.set_option("spmv_format", 0)
        A.forget()
    return y0, y1


@pytest.mark.parametrize("longest", [64, 65])
def test_delta_indexed_edge_rows(longest):
    """Empty rows, and one row of exactly 64 entries (encoded) or 65 (the
    form falls back): the SpMV is bitwise the CSR one either way."""
    G = grid_laplacian([40, 40], 1e-8)
    n = G.n
    rows = [G.ja[G.ia[i]:G.ia[i + 1]] for i in range(n)]
    vals = [G.a[G.ia[i]:G.ia[i + 1]] for i in range(n)]
    for i in (3, 500, n - 1):  # empty rows
        rows[i] = rows[i][:0]
        vals[i] = vals[i][:0]
    r = 800  # a long row: its stencil plus columns at distinct offsets
    extra = np.setdiff1d(np.arange(r - 2 * longest, r + 2 * longest, 3), rows[r])
    cols = np.sort(np.concatenate([rows[r], extra[:longest - rows[r].shape[0]]]))
    rows[r] = cols.astype(np.int32)
    vals[r] = -np.ones(cols.shape[0]) / (1 + np.arange(cols.shape[0]) % 3)
    ia = np.zeros(n + 1, dtype=np.int32)
    np.cumsum([c.shape[0] for c in rows], out=ia[1:])
    from lap_amd.problems import CSR
    M = CSR(n, ia, np.concatenate(rows).astype(np.int32), np.concatenate(vals))
    assert np.diff(M.ia).max() == longest
    y0, y1 = spmv_both(M, 2)
    assert torch.equal(y0, y1)
