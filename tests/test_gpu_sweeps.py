"""SpMV-only preconditioners with one launch per sweep (sweep.hip, option
fuse_sweeps): it_jacobi, GS_it and GS_it2.

Bars: the fused application is bitwise the unfused launch sequence (same
products, same CSR-order row sums) and value-equal to the oracle's
prec_apply (to 1e-12 where a row exceeds one LDS tile and is summed by a
workgroup tree), for several (M, K), on grids, an unstructured graph, a star (one
row of 6000 nonzeros: the workgroup-tree rows) and a matrix whose middle row
has 1500 nonzeros in each triangle (more than one LDS tile for L and U
together, less for each: the pair kernel's row-of-its-own path).  A PCG solve
with each preconditioner is bitwise the unfused solve (x and history).
"""
import numpy as np
import pytest

from cases import PROBLEMS, star_graph, weighted_graph
from lap_amd import solver as S
from lap_amd.problems import from_coo

pytestmark = pytest.mark.gpu

KINDS = ["it_jacobi", "GS_it", "GS_it2"]
MK = [(1, 1), (1, 2), (2, 1), (3, 3), (1, 5)]


def two_sided_hub(n=4000, hub=2000, side=1500, seed=3):
    """weighted_graph plus a hub joined to `side` nodes below and above it."""
    base = weighted_graph(n, 6, seed)
    rng = np.random.default_rng(seed)
    lo = rng.choice(np.arange(0, hub), side, replace=False)
    hi = rng.choice(np.arange(hub + 1, n), side, replace=False)
    nb = np.concatenate([lo, hi])
    rows = np.concatenate([base.row_ids(), np.full(nb.shape[0], hub), nb, nb, [hub]])
    cols = np.concatenate([base.ja, nb, np.full(nb.shape[0], hub), nb, [hub]])
    vals = np.concatenate([base.a, -np.ones(nb.shape[0]), -np.ones(nb.shape[0]),
                           np.ones(nb.shape[0]), [float(nb.shape[0])]])
    # duplicate (row, col) pairs summed so the matrix stays SPD and CSR-unique
    key = rows.astype(np.int64) * n + cols
    u, inv = np.unique(key, return_inverse=True)
    v = np.zeros(u.shape[0])
    np.add.at(v, inv, vals)
    return from_coo(n, u // n, u % n, v, symmetric=False)


CASES = {
    "grid3d_16": PROBLEMS["grid3d_16"],
    "graph_big": PROBLEMS["graph_big"],
    "star_6000": lambda: star_graph(6000),
    "two_sided_hub": two_sided_hub,
}


@pytest.fixture(params=sorted(CASES))
def problem(request):
    print(f"[case] sweeps problem={request.param}", flush=True)
    return CASES[request.param]()


def bits(t):
    return t.cpu().numpy().view(np.int64)


@pytest.mark.parametrize("kind", KINDS)
def test_fused_sweeps_bitwise_unfused_and_oracle(orc, problem, kind):
    A = S.DeviceCSR.from_csr(problem)
    r = S.to_dev(np.random.default_rng(4).standard_normal(problem.n))
    for m, k in MK:
        P = S.make_preconditioner(A, problem, kind, m=m, k=k)
        try:
            S.set_option("fuse_sweeps", 0)
            ref = S.prec_apply(A, P, r)
            S.set_option("fuse_sweeps", 1)
            got = S.prec_apply(A, P, r)
        finally:
            S.set_option("fuse_sweeps", 1)
        assert np.array_equal(bits(got), bits(ref)), (kind, m, k)
        want = orc.prec_apply(problem, orc.make_prec(problem, kind, m=m, k=k),
                              r.cpu().numpy())
        if np.diff(problem.ia).max() > 2048:
            # rows past one LDS tile are summed by a workgroup tree, not in
            # CSR order (as the plain SpMV does them)
            g = got.cpu().numpy()
            assert np.linalg.norm(g - want) <= 1e-12 * np.linalg.norm(want), (kind, m, k)
        else:
            np.testing.assert_array_equal(got.cpu().numpy(), want)


@pytest.mark.parametrize("kind,m,k", [("it_jacobi", 1, 2), ("GS_it", 1, 1), ("GS_it2", 1, 2),
                                      ("GS_it", 12, 1), ("GS_it2", 3, 15)])
def test_fused_pcg_bitwise_unfused(kind, m, k):
    """GS_it is not a symmetric preconditioner and PCG stagnates with it on
    this grid (M = 1 and M = 12 alike), which both paths must do identically
    (maxit 300 bounds the run)."""
    Ah = PROBLEMS["grid3d_16"]()
    A = S.DeviceCSR.from_csr(Ah)
    b = S.to_dev(np.random.default_rng(5).standard_normal(Ah.n))
    P = S.make_preconditioner(A, Ah, kind, m=m, k=k)
    try:
        S.set_option("fuse_sweeps", 0)
        r0 = S.pcg(A, b, P, tol=1e-10, maxit=300)
        S.set_option("fuse_sweeps", 1)
        r1 = S.pcg(A, b, P, tol=1e-10, maxit=300)
    finally:
        S.set_option("fuse_sweeps", 1)
    print(f"[diag] {kind} M={m} K={k}: {r1.iterations} iterations, "
          f"fused {r1.gpu_ms:.2f} ms vs unfused {r0.gpu_ms:.2f} ms", flush=True)
    assert r0.flag == r1.flag and r0.iterations == r1.iterations
    if kind != "GS_it":
        assert r1.flag == 0
    assert np.array_equal(bits(r0.x), bits(r1.x))
    assert np.array_equal(r0.history.view(np.int64), r1.history.view(np.int64))
