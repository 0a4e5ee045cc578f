"""Synthetic SPD graph-Laplacian systems and the reference's setup transforms.

Everything here is host-side input preparation (numpy), untimed, like the
reference's io_utils.c and driver setup:

* ``grid_laplacian``   -- d-dimensional 2d+1-point grid graph Laplacian in CSR
  (rows sorted, diagonal = degree + shift), the matrices the reference is
  benchmarked on.
* ``split``            -- A = L + D + U (io_utils.c:268-363).
* ``augment_lower`` / ``augment_upper`` -- L+D with the diagonal last and D+U
  with the diagonal first (cg_driver2.c:95-144 for GS_std on the GPU and
  :146-196 for ichol on the CPU).
* ``rhs_normal`` / ``rhs_uniform`` -- seeded, mean-subtracted right-hand sides.
* ``write_mm`` / ``read_mm`` -- Matrix Market coordinate files with the
  semantics of io_utils.c:15-74 and :141-235 (symmetric files are mirrored,
  rows sorted, ``pattern`` means values of 1.0), and RHS ``array`` files as
  read by io_utils.c:239-264.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class CSR:
    """Square CSR matrix with int32 row pointers / column indices."""

    n: int
    ia: np.ndarray  # int32[n+1]
    ja: np.ndarray  # int32[nnz]
    a: np.ndarray  # float64[nnz]

    @property
    def nnz(self) -> int:
        return int(self.ia[-1])

    def check(self) -> None:
        if self.ia.dtype != np.int32 or self.ja.dtype != np.int32:
            raise TypeError("CSR indices must be int32")
        if self.ia.shape != (self.n + 1,) or self.ia[0] != 0:
            raise ValueError("bad row pointer array")
        if self.ja.shape[0] != self.nnz or self.a.shape[0] != self.nnz:
            raise ValueError("ja/a length differs from ia[n]")

    def to_dense(self) -> np.ndarray:
        out = np.zeros((self.n, self.n))
        rows = np.repeat(np.arange(self.n), np.diff(self.ia))
        np.add.at(out, (rows, self.ja), self.a)
        return out

    def row_ids(self) -> np.ndarray:
        return np.repeat(np.arange(self.n, dtype=np.int32), np.diff(self.ia))


def grid_laplacian(dims, shift: float = 0.0) -> CSR:
    """Graph Laplacian of a len(dims)-dimensional grid, natural ordering with
    the first dimension fastest; off-diagonals -1, diagonal = degree + shift.
    Columns of every row are ascending (the order coo_to_csr produces)."""
    dims = [int(d) for d in dims]
    n = int(np.prod(dims))
    strides = np.cumprod([1] + dims[:-1]).astype(np.int64)
    coords = []
    idx = np.arange(n, dtype=np.int64)
    for d, s in zip(dims, strides):
        coords.append((idx // s) % d)
    # neighbour offsets in ascending column order: -s_k..., -s_0, 0, +s_0, ..., +s_k
    offs = [-s for s in strides[::-1]] + [0] + list(strides)
    valid = []
    for k in range(len(dims) - 1, -1, -1):
        valid.append(coords[k] > 0)
    valid.append(np.ones(n, dtype=bool))
    for k in range(len(dims)):
        valid.append(coords[k] < dims[k] - 1)
    mask = np.stack(valid, axis=1)  # (n, 2d+1)
    del valid
    degree = mask.sum(axis=1) - 1
    cols = idx[:, None] + np.asarray(offs, dtype=np.int64)[None, :]
    vals = np.full(mask.shape, -1.0)
    centre = len(dims)
    vals[:, centre] = degree.astype(np.float64) + shift
    ja = cols[mask].astype(np.int32)
    del cols
    a = vals[mask]
    del vals
    ia = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(mask.sum(axis=1), out=ia[1:])
    return CSR(n, ia.astype(np.int32), ja, a)


def rhs_normal(n: int, seed: int = 12345) -> np.ndarray:
    b = np.random.default_rng(seed).standard_normal(n)
    return b - b.mean()


def rhs_uniform(n: int, seed: int = 12345) -> np.ndarray:
    b = np.random.default_rng(seed).uniform(-1.0, 1.0, n)
    return b - b.mean()


@dataclass
class EdgeGraph:
    """Undirected weighted graph: edges (src[e], dst[e]) with src < dst."""
    n: int
    src: np.ndarray  # int32
    dst: np.ndarray  # int32
    w: np.ndarray    # float64
    points: np.ndarray | None = None

    @property
    def m(self) -> int:
        return int(self.src.shape[0])


def random_geometric_graph(n: int, seed: int = 12345, radius: float | None = None) -> EdgeGraph:
    """n points U[0,1]^2, an edge between points closer than r =
    sqrt(10/(pi n)) (mean degree ~10), weight exp(-d^2/sigma^2), sigma = r/2."""
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(seed)
    pts = rng.random((n, 2))
    r = float(np.sqrt(10.0 / (np.pi * n))) if radius is None else float(radius)
    pairs = cKDTree(pts).query_pairs(r, output_type="ndarray")
    if pairs.shape[0]:
        pairs = pairs[np.lexsort((pairs[:, 1], pairs[:, 0]))]
    src = pairs[:, 0].astype(np.int32)
    dst = pairs[:, 1].astype(np.int32)
    d2 = ((pts[src] - pts[dst]) ** 2).sum(axis=1)
    sigma = r / 2.0
    return EdgeGraph(n, src, dst, np.exp(-d2 / sigma**2), pts)


def graph_csr(g: EdgeGraph, w: np.ndarray, shift: float) -> CSR:
    """Host CSR of L(w) = D - W + shift*I (row sums in edge order; for
    reference and dense checks, not bitwise with the device assembly)."""
    deg = np.bincount(g.src, w, g.n) + np.bincount(g.dst, w, g.n)
    rows = np.concatenate([g.dst, np.arange(g.n)])
    cols = np.concatenate([g.src, np.arange(g.n)])
    vals = np.concatenate([-w, deg + shift])
    return from_coo(g.n, rows, cols, vals, symmetric=True)


def graph_dense(g: EdgeGraph, w: np.ndarray, shift: float) -> np.ndarray:
    """Dense L(w) = sum_e w_e (u_s - u_d)(u_s - u_d)^T + shift*I."""
    Ld = np.zeros((g.n, g.n))
    np.add.at(Ld, (g.src, g.dst), -w)
    np.add.at(Ld, (g.dst, g.src), -w)
    np.add.at(Ld, (g.src, g.src), w)
    np.add.at(Ld, (g.dst, g.dst), w)
    Ld[np.diag_indices(g.n)] += shift
    return Ld


def diagonal(A: CSR) -> np.ndarray:
    rows = A.row_ids()
    d = np.zeros(A.n)
    on = A.ja == rows
    d[rows[on]] = A.a[on]
    return d


def _select(A: CSR, keep: np.ndarray) -> CSR:
    rows = A.row_ids()[keep]
    ia = np.zeros(A.n + 1, dtype=np.int64)
    np.cumsum(np.bincount(rows, minlength=A.n), out=ia[1:])
    return CSR(A.n, ia.astype(np.int32), A.ja[keep].copy(), A.a[keep].copy())


def split(A: CSR):
    """A = L + D + U with strict L, strict U (io_utils.c:268-363)."""
    rows = A.row_ids()
    return _select(A, A.ja < rows), diagonal(A), _select(A, A.ja > rows)


def augment_lower(A: CSR) -> CSR:
    """L+D, rows sorted so the diagonal is the last entry."""
    return _select(A, A.ja <= A.row_ids())


def augment_upper(A: CSR) -> CSR:
    """D+U, rows sorted so the diagonal is the first entry."""
    return _select(A, A.ja >= A.row_ids())


def transpose(A: CSR) -> CSR:
    rows = A.row_ids()
    order = np.lexsort((rows, A.ja))
    ia = np.zeros(A.n + 1, dtype=np.int64)
    np.cumsum(np.bincount(A.ja, minlength=A.n), out=ia[1:])
    return CSR(A.n, ia.astype(np.int32), rows[order].astype(np.int32), A.a[order].copy())


def from_coo(n: int, rows, cols, vals, symmetric: bool) -> CSR:
    """COO -> CSR with the semantics of coo_to_csr (io_utils.c:141-235):
    symmetric entries are mirrored, each row is sorted by column; equal
    (row, column) pairs keep the fill order (entry i, then its mirror,
    before entry i+1), as the reference's merge-sorting qsort does."""
    rows = np.asarray(rows, dtype=np.int64)
    cols = np.asarray(cols, dtype=np.int64)
    vals = np.asarray(vals, dtype=np.float64)
    seq = 2 * np.arange(rows.shape[0], dtype=np.int64)  # the reference's fill order
    if symmetric:
        off = rows != cols
        rows, cols, vals, seq = (np.concatenate([rows, cols[off]]),
                                 np.concatenate([cols, rows[off]]),
                                 np.concatenate([vals, vals[off]]),
                                 np.concatenate([seq, seq[off] + 1]))
    order = np.lexsort((seq, cols, rows))
    rows, cols, vals = rows[order], cols[order], vals[order]
    ia = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.bincount(rows, minlength=n), out=ia[1:])
    if ia[-1] >= 2**31:
        raise ValueError("nnz exceeds int32 CSR")
    return CSR(n, ia.astype(np.int32), cols.astype(np.int32), vals)


def lower_coo(A: CSR):
    """Lower triangle (with diagonal) as 0-based COO, the symmetric Matrix
    Market content that split() expects (io_utils.c:317-321)."""
    rows = A.row_ids()
    keep = A.ja <= rows
    return rows[keep], A.ja[keep], A.a[keep]


def write_mm(path: str, A: CSR) -> None:
    r, c, v = lower_coo(A)
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real symmetric\n")
        f.write(f"{A.n} {A.n} {r.shape[0]}\n")
        np.savetxt(f, np.column_stack([r + 1, c + 1, v]), fmt=["%d", "%d", "%.17g"])


def write_adjacency(path: str, n: int, src, dst, w=None) -> None:
    """Adjacency of an undirected graph as Matrix Market coordinate
    symmetric, one line per edge with row > column, no diagonal: the input of
    the reference's Laplacian driver (read_adjacency_file, io_utils.c:80-130).
    Pattern when w is None, else real weights (%.17g)."""
    r = np.maximum(src, dst).astype(np.int64) + 1
    c = np.minimum(src, dst).astype(np.int64) + 1
    with open(path, "w") as f:
        kind = "pattern" if w is None else "real"
        f.write(f"%%MatrixMarket matrix coordinate {kind} symmetric\n")
        f.write(f"{n} {n} {r.shape[0]}\n")
        if w is None:
            f.write("".join(f"{i} {j}\n" for i, j in zip(r, c)))
        else:
            f.write("".join(f"{i} {j} {v:.17g}\n" for i, j, v in zip(r, c, w)))


def grid_edges(dims):
    """Edges (i < j) of the 2D/3D grid graph behind grid_laplacian."""
    dims = list(dims)
    n = int(np.prod(dims))
    idx = np.arange(n).reshape(dims[::-1])
    src, dst = [], []
    for ax in range(len(dims)):
        a = np.take(idx, range(idx.shape[ax] - 1), axis=ax).ravel()
        b = np.take(idx, range(1, idx.shape[ax]), axis=ax).ravel()
        src.append(a)
        dst.append(b)
    return n, np.concatenate(src).astype(np.int32), np.concatenate(dst).astype(np.int32)


def write_rhs(path: str, b: np.ndarray) -> None:
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix array real general\n")
        f.write(f"{b.shape[0]} 1\n")
        np.savetxt(f, b, fmt="%.17g")


def read_mm(path: str) -> CSR:
    with open(path) as f:
        header = f.readline()
        line = f.readline()
        while line.startswith("%"):
            line = f.readline()
        n, m, nz = (int(t) for t in line.split()[:3])
        if n != m:
            raise ValueError("matrix is not square")
        pattern = "pattern" in header
        data = np.loadtxt(f, ndmin=2, max_rows=nz)
    rows = data[:, 0].astype(np.int64) - 1
    cols = data[:, 1].astype(np.int64) - 1
    vals = np.ones(rows.shape[0]) if pattern else data[:, 2]
    return from_coo(n, rows, cols, vals, "symmetric" in header)


def read_rhs(path: str, n: int) -> np.ndarray:
    with open(path) as f:
        line = f.readline()
        while line.startswith("%"):
            line = f.readline()
        return np.loadtxt(f, max_rows=n, usecols=0, ndmin=1)



def chung_lu(n: int, gamma: float = 2.1, mean_degree: float = 16.0, cap: float = 20000.0,
             seed: int = 12345, diag_factor: float = 1.0 + 1e-6) -> CSR:
    """Weighted Chung-Lu power-law graph Laplacian, the mixed-precision
    workload on a graph with very uneven rows.

    Expected degrees w_i = c * i^(-1/(gamma-1)) (i = 1..n), c chosen for a
    mean of `mean_degree`, then capped at `cap`; n*mean_degree/2 edges whose
    endpoints are both drawn in proportion to w (the fast Chung-Lu sampler);
    self-loops dropped, duplicate edges merged, the result symmetric.  Edge
    weights are lognormal(0, 1) and the diagonal is diag_factor times the
    row's weight sum (a 1e-6 relative shift makes the M-matrix strictly
    diagonally dominant).  A vertex that received no edge gets diagonal 1.
    Rows are sorted by column."""
    rng = np.random.default_rng(seed)
    w = np.arange(1, n + 1, dtype=np.float64) ** (-1.0 / (gamma - 1.0))
    w *= mean_degree / w.mean()
    np.minimum(w, cap, out=w)
    m = int(n * mean_degree / 2)
    # m independent endpoint pairs: each side's vertex counts are
    # multinomial(m, w/sum w), and a random pairing of the two sides makes
    # the pairs independent
    prob = w / w.sum()
    verts = np.arange(n, dtype=np.int64)
    src = np.repeat(verts, rng.multinomial(m, prob))
    dst = rng.permutation(np.repeat(verts, rng.multinomial(m, prob)))
    keep = src != dst
    lo = np.minimum(src[keep], dst[keep]).astype(np.int64)
    hi = np.maximum(src[keep], dst[keep]).astype(np.int64)
    del src, dst, keep
    key = np.unique(lo * n + hi)  # sorted by (lo, hi): the upper triangle in CSR order
    lo, hi = key // n, key % n
    del key
    ew = rng.lognormal(0.0, 1.0, lo.shape[0])
    sums = np.bincount(lo, ew, n) + np.bincount(hi, ew, n)
    cl = np.bincount(hi, minlength=n)  # entries left of the diagonal per row
    cu = np.bincount(lo, minlength=n)  # entries right of it
    ia = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(cl + cu + 1, out=ia[1:])
    if ia[-1] >= 2**31:
        raise ValueError("nnz exceeds int32 CSR")
    nnz = int(ia[-1])
    ja = np.empty(nnz, dtype=np.int32)
    a = np.empty(nnz, dtype=np.float64)
    rows = np.arange(n)
    dpos = ia[:-1] + cl
    ja[dpos] = rows
    a[dpos] = np.where(sums > 0, diag_factor * sums, 1.0)
    # upper entries, already in (row, column) order
    ustart = np.zeros(n, dtype=np.int64)
    np.cumsum(cu[:-1], out=ustart[1:])
    upos = dpos[lo] + 1 + (np.arange(lo.shape[0]) - ustart[lo])
    ja[upos] = hi
    a[upos] = -ew
    # lower entries: the transpose, ordered by (hi, lo)
    order = np.argsort(hi * n + lo)  # keys are distinct: a unique order
    lr, lc = hi[order], lo[order]
    lstart = np.zeros(n, dtype=np.int64)
    np.cumsum(cl[:-1], out=lstart[1:])
    lpos = ia[lr] + (np.arange(lr.shape[0]) - lstart[lr])
    ja[lpos] = lc
    a[lpos] = -ew[order]
    return CSR(n, ia.astype(np.int32), ja, a)
