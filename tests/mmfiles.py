"""Matrix Market test files for the input-path tests.

Each writer produces a file whose parse the reference's read_mm_file /
read_adjacency_file defines (io_utils.c:15-130): a banner, optional comment
lines, the size line, then exactly nnz entry lines.  The cases cover the
formats and the quirks that matter for parity: symmetric and general,
pattern and real, duplicate (row, column) pairs, diagonal entries, unsorted
entry order, CRLF line ends, tabs, exponent / sign / hexadecimal value
syntax, long rows (several thousand entries), non-integer adjacency weights
(the int degree accumulator truncates) and an adjacency file's missing
diagonal.
"""
from __future__ import annotations

import os

import numpy as np

from lap_amd.problems import grid_laplacian, lower_coo


def _write(path, banner, n, m, rows, cols, vals=None, comments=(), crlf=False, fmt="%.17g",
           sep=" "):
    nl = "\r\n" if crlf else "\n"
    lines = [banner] + ["%" + c for c in comments] + [f"{n} {m} {len(rows)}"]
    if vals is None:
        lines += [f"{r + 1}{sep}{c + 1}" for r, c in zip(rows, cols)]
    else:
        lines += [f"{r + 1}{sep}{c + 1}{sep}" + (v if isinstance(v, str) else fmt % v)
                  for r, c, v in zip(rows, cols, vals)]
    with open(path, "w", newline="") as f:
        f.write(nl.join(lines) + nl)
    return path


def grid_symmetric(d, dims=(6, 5, 4), shift=1e-3, shuffle=True):
    """Lower triangle of a shifted grid Laplacian, entries shuffled."""
    A = grid_laplacian(list(dims), shift)
    r, c, v = lower_coo(A)
    if shuffle:
        p = np.random.default_rng(1).permutation(r.shape[0])
        r, c, v = r[p], c[p], v[p]
    return _write(os.path.join(d, "grid_sym.mtx"),
                  "%%MatrixMarket matrix coordinate real symmetric", A.n, A.n, r, c, v,
                  comments=["generated grid", " second comment"])


def general_with_duplicates(d, n=300, seed=3):
    """General real matrix, random order, duplicate pairs, varied value
    syntax, a few rows with thousands of entries, CRLF line ends, tabs."""
    rng = np.random.default_rng(seed)
    k = 6 * n
    rows = rng.integers(0, n, k)
    cols = rng.integers(0, n, k)
    dup = rng.integers(0, k, 200)
    rows = np.concatenate([rows, rows[dup], np.full(5000, 7), np.full(4500, 11)])
    cols = np.concatenate([cols, cols[dup], rng.integers(0, n, 5000), rng.integers(0, n, 4500)])
    vals = list(rng.standard_normal(rows.shape[0]))
    special = ["1e-3", "-0.5", "3", "+2.5E+01", "0x1.8p1", "-7.25e-300", "1.0000000000000002"]
    for i, s in enumerate(special):
        vals[37 * i + 1] = s
    return _write(os.path.join(d, "general_dup.mtx"),
                  "%%MatrixMarket matrix coordinate real general", n, n, rows, cols, vals,
                  crlf=True, sep="\t")


def value_formats(d, n=100, k=30000, seed=8):
    """Values in every syntax strtod accepts that a file is likely to hold:
    %.1g .. %.17g, %e, integers, signs, leading zeros, '5.', '.5', large and
    tiny exponents, subnormals, -0, and digit strings longer than 19."""
    rng = np.random.default_rng(seed)
    rows, cols = rng.integers(0, n, k), rng.integers(0, n, k)
    mags = 10.0 ** rng.uniform(-30, 30, k)
    x = rng.choice([-1.0, 1.0], k) * mags * rng.random(k)
    vals = []
    fixed = ["0", "-0", "+0.0", "5.", ".5", "-.25", "007", "1e5", "1E-5", "1e+22", "1e-22",
             "9007199254740993", "9007199254740992", "12345678901234567890", "1e23",
             "4.9e-324", "2.2250738585072014e-308", "1.7976931348623157e308",
             "0.000000000000000000000000123", "123456789012345678901234567890e-20",
             "3.14159265358979323846264338327950288", "-1", "+1", "1.5e-5", "6.00000001"]
    for i in range(k):
        if i < len(fixed):
            vals.append(fixed[i])
            continue
        f = i % 6
        if f == 0:
            vals.append("%.*g" % (int(rng.integers(1, 18)), x[i]))
        elif f == 1:
            vals.append("%.*e" % (int(rng.integers(0, 17)), x[i]))
        elif f == 2:
            vals.append(str(int(rng.integers(-10**6, 10**6))))
        elif f == 3:
            vals.append("%.*f" % (int(rng.integers(0, 12)), x[i] / mags[i] * 1000))
        elif f == 4:
            vals.append("%.17g" % x[i])
        else:
            vals.append(repr(float(x[i])))
    return _write(os.path.join(d, "value_formats.mtx"),
                  "%%MatrixMarket matrix coordinate real general", n, n, rows, cols, vals)


def long_rows_symmetric(d, n=20000, seed=4):
    """Symmetric file whose mirrored entries make two rows of ~9000 and
    ~20000 entries (the multi-run merge path), plus a random sparse part."""
    rng = np.random.default_rng(seed)
    rows = list(range(n))  # diagonal
    cols = list(range(n))
    for hub, count in ((3, 9000), (5, n - 6)):
        others = rng.choice(np.arange(hub + 1, n), size=count, replace=False)
        rows += list(others)
        cols += [hub] * count
    extra_r = rng.integers(1, n, 3 * n)
    extra_c = (extra_r * rng.random(3 * n)).astype(np.int64)
    rows += list(extra_r)
    cols += list(extra_c)
    vals = rng.standard_normal(len(rows))
    return _write(os.path.join(d, "long_rows.mtx"),
                  "%%MatrixMarket matrix coordinate real symmetric", n, n, np.array(rows),
                  np.array(cols), vals)


def adjacency_pattern(d, n=500, seed=5):
    """Pattern adjacency of a random graph (row > column, no diagonal)."""
    rng = np.random.default_rng(seed)
    edges = set()
    while len(edges) < 4 * n:
        i, j = (int(v) for v in rng.integers(0, n, 2))
        if i != j:
            edges.add((max(i, j), min(i, j)))
    e = np.array(sorted(edges, key=lambda t: rng.random()))
    return _write(os.path.join(d, "adj_pattern.mtx"),
                  "%%MatrixMarket matrix coordinate pattern symmetric", n, n, e[:, 0], e[:, 1])


def adjacency_weighted(d, n=400, seed=6):
    """Real weights in (0.1, 2.6): the int degree accumulator truncates at
    every step, so degrees differ from the rounded weight sums."""
    rng = np.random.default_rng(seed)
    edges = set()
    while len(edges) < 3 * n:
        i, j = (int(v) for v in rng.integers(0, n, 2))
        if i != j:
            edges.add((max(i, j), min(i, j)))
    e = np.array(sorted(edges))
    w = 0.1 + 2.5 * rng.random(e.shape[0])
    return _write(os.path.join(d, "adj_weighted.mtx"),
                  "%%MatrixMarket matrix coordinate real symmetric", n, n, e[:, 0], e[:, 1], w)


def rhs_file(d, n, seed=7):
    b = np.random.default_rng(seed).standard_normal(n)
    path = os.path.join(d, "rhs.mtx")
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix array real general\n% comment\n")
        f.write(f"{n} 1\n")
        f.write("".join(f"{v:.17g}\n" for v in b))
    return path, b


ALL_MATRICES = [grid_symmetric, general_with_duplicates, value_formats, long_rows_symmetric]
ALL_ADJACENCY = [adjacency_pattern, adjacency_weighted]


def edge_cases(d):
    """Small files at the edges: 1x1, no entries, empty rows, diagonal only,
    pattern general, and values 0, -0, nan, inf, -inf (kept bit for bit)."""
    out = {}
    out["one_by_one"] = _write(os.path.join(d, "e_one.mtx"),
                               "%%MatrixMarket matrix coordinate real symmetric", 1, 1,
                               [0], [0], [4.5])
    out["no_entries"] = _write(os.path.join(d, "e_none.mtx"),
                               "%%MatrixMarket matrix coordinate real general", 5, 5, [], [], [])
    out["empty_rows"] = _write(os.path.join(d, "e_rows.mtx"),
                               "%%MatrixMarket matrix coordinate real general", 9, 9,
                               [0, 8, 8, 4], [8, 0, 8, 4], [1.0, 2.0, 3.0, 4.0])
    out["diagonal_only"] = _write(os.path.join(d, "e_diag.mtx"),
                                  "%%MatrixMarket matrix coordinate real symmetric", 6, 6,
                                  list(range(6)), list(range(6)), [float(i + 1) for i in range(6)])
    out["pattern_general"] = _write(os.path.join(d, "e_pat.mtx"),
                                    "%%MatrixMarket matrix coordinate pattern general", 4, 4,
                                    [0, 1, 3, 3, 2], [1, 0, 3, 0, 2])
    out["special_values"] = _write(os.path.join(d, "e_special.mtx"),
                                   "%%MatrixMarket matrix coordinate real symmetric", 4, 4,
                                   [0, 1, 2, 3, 3, 2], [0, 0, 1, 3, 1, 2],
                                   ["0", "-0", "nan", "inf", "-inf", "-0.0"])
    return out
