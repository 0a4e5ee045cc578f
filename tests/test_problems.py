"""The power-law (Chung-Lu) workload generator: structure, the values its
definition fixes, and its direct CSR assembly against the general COO path."""
import numpy as np

from lap_amd.problems import chung_lu, from_coo, lower_coo


def test_chung_lu_structure_and_values():
    A = chung_lu(8192, seed=3)
    A.check()
    rows = A.row_ids()
    # rows strictly ascending, diagonal present once per row
    starts = A.ia[:-1]
    d = np.diff(A.ja.astype(np.int64))
    d[starts[1:] - 1] = 1  # row boundaries
    assert np.all(d > 0)
    diag = A.ja == rows
    assert np.array_equal(np.bincount(rows[diag], minlength=A.n), np.ones(A.n, np.int64))
    # symmetric, off-diagonals negative
    key = rows.astype(np.int64) * A.n + A.ja
    tkey = A.ja.astype(np.int64) * A.n + rows
    o, t = np.argsort(key), np.argsort(tkey)
    assert np.array_equal(key[o], tkey[t]) and np.array_equal(A.a[o], A.a[t])
    assert np.all(A.a[~diag] < 0)
    # diagonal = (1 + 1e-6) * weight sum; a vertex without edges gets 1
    wsum = np.bincount(rows[~diag], -A.a[~diag], A.n)
    dv = np.empty(A.n)
    dv[rows[diag]] = A.a[diag]
    has = wsum > 0
    np.testing.assert_allclose(dv[has], (1 + 1e-6) * wsum[has], rtol=1e-13)
    assert np.all(dv[~has] == 1.0)


def test_chung_lu_degrees_follow_weights():
    A = chung_lu(1 << 16, seed=5)
    deg = np.diff(A.ia) - 1
    # n*16/2 sampled pairs, minus self-loops and duplicates
    assert 12.0 < deg.mean() < 16.0
    # expected degree ~ i^(-1/1.1): the first vertices are the hubs
    assert deg[:16].min() > 50 * np.median(deg)
    assert deg[: A.n // 100].mean() > 10 * deg[A.n // 2:].mean()


def test_chung_lu_cap():
    """Capping equalises the hubs' expected degrees: their realised degrees
    bunch together, where without the cap they spread over a factor > 2."""
    top = lambda A: np.sort(np.diff(A.ia) - 1)[-8:]
    capped = top(chung_lu(1 << 16, cap=200.0, seed=5))
    free = top(chung_lu(1 << 16, cap=1e12, seed=5))
    assert capped[-1] < 1.15 * capped[0]
    assert free[-1] > 2 * free[0]


def test_chung_lu_direct_assembly_is_the_coo_path():
    A = chung_lu(4096, seed=9)
    r, c, v = lower_coo(A)
    B = from_coo(A.n, r, c, v, symmetric=True)
    assert np.array_equal(A.ia, B.ia) and np.array_equal(A.ja, B.ja)
    assert np.array_equal(A.a.view(np.int64), B.a.view(np.int64))


def test_chung_lu_deterministic():
    A, B = chung_lu(2048, seed=4), chung_lu(2048, seed=4)
    assert np.array_equal(A.ja, B.ja) and np.array_equal(A.a, B.a)
    C = chung_lu(2048, seed=5)
    assert not (A.ja.shape == C.ja.shape and np.array_equal(A.ja, C.ja))
