"""Small seeded test matrices shared by the CPU and GPU suites."""
import numpy as np

from lap_amd.problems import from_coo, grid_laplacian


def weighted_graph(n, deg, seed, shift=1e-3):
    """Random weighted graph Laplacian + shift*I (SPD)."""
    rng = np.random.default_rng(seed)
    m = n * deg // 2
    r = rng.integers(0, n, m)
    c = rng.integers(0, n, m)
    keep = r != c
    r, c = np.minimum(r[keep], c[keep]), np.maximum(r[keep], c[keep])
    key = np.unique(r.astype(np.int64) * n + c)
    r, c = key // n, key % n
    w = rng.lognormal(0.0, 1.0, r.shape[0])
    degw = np.bincount(r, w, n) + np.bincount(c, w, n)
    rows = np.concatenate([c, np.arange(n)])
    cols = np.concatenate([r, np.arange(n)])
    vals = np.concatenate([-w, degw + shift])
    return from_coo(n, rows, cols, vals, symmetric=True)


def star_graph(n, shift=1.0):
    """Hub row 0 joined to every other node: one row of n nonzeros."""
    rows = np.concatenate([np.arange(1, n), np.arange(n)])
    cols = np.concatenate([np.zeros(n - 1, dtype=np.int64), np.arange(n)])
    deg = np.ones(n)
    deg[0] = n - 1
    vals = np.concatenate([-np.ones(n - 1), deg + shift])
    return from_coo(n, rows, cols, vals, symmetric=True)


PROBLEMS = {
    "grid2d": lambda: grid_laplacian([9, 7], 1e-8),
    "grid3d": lambda: grid_laplacian([5, 4, 3], 1e-8),
    "graph": lambda: weighted_graph(300, 8, 7),
    "grid3d_16": lambda: grid_laplacian([16, 12, 10], 1e-8),
    "graph_big": lambda: weighted_graph(5000, 12, 11),
}
CPU_PROBLEMS = ["grid2d", "grid3d", "graph"]


def ring_chords(n, per, seed):
    """Connected graph: a ring plus about `per` random chords per vertex
    (edges (i < j), no duplicates)."""
    rng = np.random.default_rng(seed)
    s = np.arange(n)
    d = (s + 1) % n
    a = rng.integers(0, n, n * per)
    b = rng.integers(0, n, n * per)
    keep = a != b
    src = np.concatenate([s, a[keep]])
    dst = np.concatenate([d, b[keep]])
    lo, hi = np.minimum(src, dst), np.maximum(src, dst)
    key = np.unique(lo.astype(np.int64) * n + hi)
    return n, (key // n).astype(np.int32), (key % n).astype(np.int32)


def nine_point(nx, ny, shift=1e-8):
    """2D 9-point graph Laplacian: L+D rows of up to 5 entries (the 8-wide
    packed rows of the triangular solves)."""
    from lap_amd.problems import from_coo
    idx = np.arange(nx * ny).reshape(ny, nx)
    rows, cols = [], []
    for dy, dx in ((0, 1), (1, -1), (1, 0), (1, 1)):
        a = idx[max(0, -dy):ny - max(0, dy), max(0, -dx):nx - max(0, dx)]
        b = idx[max(0, dy):ny - max(0, -dy) if dy <= 0 else ny, max(0, dx):nx - max(0, -dx)]
        a, b = a.ravel(), b.ravel()
        rows.append(np.maximum(a, b))
        cols.append(np.minimum(a, b))
    r, c = np.concatenate(rows), np.concatenate(cols)
    deg = np.bincount(r, minlength=nx * ny) + np.bincount(c, minlength=nx * ny)
    n = nx * ny
    return from_coo(n, np.concatenate([r, np.arange(n)]), np.concatenate([c, np.arange(n)]),
                    np.concatenate([-np.ones(r.shape[0]), deg + shift]), symmetric=True)
