"""Value-indexed CSR in the PCG SpMV (option spmv_format 1, vi.hip).

Bars: with the values stored as one-byte indices into a table of the
matrix's distinct doubles, every PCG solve is bitwise the CSR solve (x,
residual history, iteration count) -- for Jacobi in both three-launch
iteration variants and for IC0 and GS_std, whose A products use it too; a
matrix with 255 distinct values is encoded, one with 256 is not (the solve
then uses the CSR values, again bitwise the same).
"""
import numpy as np
import pytest
import torch

from cases import PROBLEMS, weighted_graph
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


def solve_both(Ah, kind, variant=1):
    A = S.DeviceCSR.from_csr(Ah)
    P = S.make_preconditioner(A, Ah, kind)
    b = S.to_dev(rhs_normal(Ah.n, 21))
    prev = S.set_option("jacobi_variant", variant)
    try:
        S.set_option("spmv_format", 0)
        r0 = S.pcg(A, b, P, tol=1e-10, maxit=5000)
        S.set_option("spmv_format", 1)
        r1 = S.pcg(A, b, P, tol=1e-10, maxit=5000)
    finally:
        S.set_option("spmv_format", 0)
        S.set_option("jacobi_variant", prev)
    return r0, r1


def same(r0, r1):
    return (r0.iterations == r1.iterations and r0.flag == r1.flag
            and np.array_equal(bits(r0.x), bits(r1.x))
            and np.array_equal(r0.history.view(np.int64), r1.history.view(np.int64)))


@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("pname", ["grid3d_16", "grid2d"])
def test_value_indexed_jacobi_pcg_bitwise(pname, variant):
    r0, r1 = solve_both(PROBLEMS[pname](), "line_jacobi", variant)
    print(f"[diag] vi {pname} v{variant}: {r1.iterations} it, csr {r0.gpu_ms:.2f} ms, "
          f"vi {r1.gpu_ms:.2f} ms", flush=True)
    assert same(r0, r1)


@pytest.mark.parametrize("kind", ["ichol", "GS_std"])
def test_value_indexed_other_preconditioners_bitwise(kind):
    r0, r1 = solve_both(PROBLEMS["grid3d_16"](), kind)
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


def test_value_indexed_spmv_timed_kernel():
    Ah = grid_laplacian([64, 64, 64], 1e-8)
    A = S.DeviceCSR.from_csr(Ah)
    x = torch.randn(A.n, dtype=torch.float64, device="cuda")
    y0, y1 = torch.empty_like(x), torch.empty_like(x)
    try:
        S.set_option("spmv_format", 0)
        t0 = S.spmv_timed(A, x, y0, 5, fused=True)
        S.set_option("spmv_format", 1)
        t1 = S.spmv_timed(A, x, y1, 5, fused=True)
    finally:
        S.set_option("spmv_format", 0)
    assert t0 > 0 and t1 > 0
    assert torch.equal(y0, y1)


@pytest.mark.parametrize("kind", ["line_jacobi", "ichol"])
def test_value_indexed_mixed_precision_bitwise(kind):
    """The mixed-precision inner SpMV over the value-indexed FP32 copy gives
    the FP32-CSR solve's bits."""
    Ah = PROBLEMS["grid3d_16"]()
    A = S.DeviceCSR.from_csr(Ah)
    P = S.make_preconditioner(A, Ah, kind)
    b = S.to_dev(rhs_normal(Ah.n, 22))
    try:
        S.set_option("spmv_format", 0)
        m0 = S.pcg_mixed(A, b, P, tol=1e-10, maxit=5000)
        S.set_option("spmv_format", 1)
        m1 = S.pcg_mixed(A, b, P, tol=1e-10, maxit=5000)
    finally:
        S.set_option("spmv_format", 0)
    assert (m0.iterations, m0.refinements, m0.flag) == (m1.iterations, m1.refinements, m1.flag)
    assert np.array_equal(bits(m0.x), bits(m1.x))
