"""Batched Jacobi PCG in one persistent launch (lap_pcg_batched): every
system agrees with its own oracle solve (solution to 1e-9 relative with the
mean removed for the shifted Laplacians), for every kernel configuration and
the autotuned one; degenerate systems (1 row, zero right-hand side) are
handled.

Iteration counts: within +-1 of the oracle (the reference's serial path).
Where the oracle's residual is not monotone in its last 8 iterations (the
smallest grids, e.g. 8x8 + 1e-8 I, whose residual rises from 6e-10 to 3e-8
relative before the near-null mode is resolved) the exit iteration depends on
the dot products' summation order -- the reference's own serial and OpenMP
builds stop at 36 and 34 there -- so such systems are held to +-3 and to the
FP64 true-residual check instead.  A zero right-hand side stops at iteration 0
with x = 0 (the reference divides 0/0 there and runs to maxit)."""
import numpy as np
import pytest

from cases import weighted_graph
from lap_amd import solver as S
from lap_amd.problems import grid_laplacian, rhs_normal

pytestmark = pytest.mark.gpu


def make_batch(seed=5, count=24):
    rng = np.random.default_rng(seed)
    systems, rhs = [], []
    for k in range(count):
        side = int(rng.integers(8, 41))
        A = grid_laplacian([side, side], 1e-8)
        systems.append(A)
        rhs.append(rhs_normal(A.n, seed + k))
    systems.append(grid_laplacian([1], 1e-8))        # 1x1: b = 0 after mean removal
    rhs.append(np.zeros(1))
    systems.append(weighted_graph(700, 10, seed))     # unstructured, uneven rows
    rhs.append(rhs_normal(700, 99))
    return systems, rhs


def centred(v):
    return v - v.mean()


@pytest.mark.parametrize("block,tpr", [(256, 1), (512, 2), (1024, 4), (-1, 0), (0, 0)])
def test_batched_matches_per_system_oracle(orc, block, tpr):
    print(f"[case] batched block={block} tpr={tpr}", flush=True)
    systems, rhs = make_batch()
    B = S.Batch.from_systems(systems, rhs)
    if block == 0:
        print(f"[diag] autotuned: {S.batched_tune(B, 10)}", flush=True)
    x, it, fl, r0, rs, ms = S.pcg_batched(B, tol=1e-10, maxit=5000, block=block, tpr=tpr)
    xs = [v.cpu().numpy() for v in B.split(x)]
    worst = 0.0
    relaxed = []
    for k, (A, b) in enumerate(zip(systems, rhs)):
        if not b.any():
            assert it[k] == 0 and fl[k] == 0 and not xs[k].any()
            continue
        xo, ito, flo, ho = orc.cg(A, b, "line_jacobi", tol=1e-10, maxit=5000)
        assert fl[k] == flo == 0, (k, fl[k], flo)
        tail = ho[max(0, ito - 8):]
        bumpy = bool(np.any(np.diff(tail) > 0))
        assert abs(int(it[k]) - ito) <= (3 if bumpy else 1), (k, it[k], ito, bumpy)
        if bumpy:
            relaxed.append(k)
            true = np.linalg.norm(b - orc.matvec(A, xs[k].copy()))
            assert true <= 1e-10 * np.linalg.norm(b) * 1.5
        if np.linalg.norm(centred(xo)) > 0:
            err = np.linalg.norm(centred(xs[k]) - centred(xo)) / np.linalg.norm(centred(xo))
            worst = max(worst, err)
    print(f"[diag] batched {block}/{tpr}: {ms:.2f} ms, worst rel err {worst:.2e}, "
          f"systems with a non-monotone residual tail: {relaxed}", flush=True)
    assert worst <= 1e-9


def test_batched_tpr1_rows_are_reference_order(orc):
    """With one thread per row (or the lockstep mode's CSR-order SpMV) and
    maxit-bounded runs the iterates follow the oracle's (only the dot-product
    order differs)."""
    systems, rhs = make_batch(seed=9, count=6)
    B = S.Batch.from_systems(systems, rhs)
    for block, tpr in ((256, 1), (-1, 0)):  # persistent, lockstep
        x, it, fl, _, _, _ = S.pcg_batched(B, tol=1e-30, maxit=12, block=block, tpr=tpr)
        for k, (A, b) in enumerate(zip(systems[:-2], rhs[:-2])):
            xo, ito, _, _ = orc.cg(A, b, "line_jacobi", tol=1e-30, maxit=12)
            assert it[k] == ito == 13
            xk = B.split(x)[k].cpu().numpy()
            assert np.linalg.norm(xk - xo) <= 1e-10 * np.linalg.norm(xo), (block, k)
