"""Input path on the host (no GPU): the library's Matrix Market parser, its
right-hand-side reader and writer, the binary CSR format, and its
definitions of the reference's io_utils.h functions.

Bars (all bitwise):
  * lap_mm_load's COO == the reference's read_mm_file / read_adjacency_file
    (compiled from io_utils.c into oracle/_ref) == the oracle's orc_mm_read;
  * the library's coo_to_csr / split / create_L_and_split (the io_utils
    drop-in) == the reference's, every array and every mmatrix field;
  * the oracle's orc_coo_to_csr / orc_split / orc_adjacency_laplacian ==
    the reference's (pinning the oracle the device tests compare against),
    the augmented split == the numpy restatement of cg_driver2.c:95-144, and
    L + D + U == A;
  * lap_rhs_load == read_rhs; writer and binary CSR round trips are exact.
"""
import ctypes as C
import os

import numpy as np
import pytest

import mmfiles as MF
from lap_amd import io as IO
from lap_amd import lib as L
from lap_amd.problems import augment_lower, augment_upper, from_coo

ORACLE_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")

pytestmark = pytest.mark.skipif(not os.path.exists(L.LIB_PATH), reason="liblapgfx.so not built")


@pytest.fixture(scope="module")
def files(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("mm"))
    mats = {f.__name__: f(d) for f in MF.ALL_MATRICES}
    adj = {f.__name__: f(d) for f in MF.ALL_ADJACENCY}
    return d, mats, adj


@pytest.fixture(scope="module")
def ref_io(ref):
    import oracle as O
    return O.IoUtils(ref.lib)


@pytest.fixture(scope="module")
def our_io():
    import oracle as O
    return O.IoUtils(C.CDLL(L.LIB_PATH))


def csr_equal(X, Y):
    return (X.n == Y.n and np.array_equal(X.ia, Y.ia) and np.array_equal(X.ja, Y.ja)
            and np.array_equal(X.a.view(np.int64), Y.a.view(np.int64)))


def bits(v):
    return np.ascontiguousarray(v, np.float64).view(np.int64)


@pytest.mark.parametrize("adjacency", [False, True])
def test_parser_matches_reference_and_oracle(files, ref_io, orc, adjacency):
    _, mats, adj = files
    for name, path in (adj if adjacency else mats).items():
        info, r, c, v = IO.read_coo(path, adjacency)
        n, m, nnz, sym, rr, rc, rv = ref_io.read(path, adjacency)
        assert (info.n, info.m, info.nnz, int(info.symmetric)) == (n, m, nnz, sym), name
        assert np.array_equal(r, rr) and np.array_equal(c, rc), name
        assert np.array_equal(bits(v), bits(rv)), name
        on, om, osym, orr, orc_c, orv = orc.mm_read(path, adjacency)
        assert (on, om, osym) == (n, m, bool(sym)) and np.array_equal(orr, rr), name
        assert np.array_equal(orc_c, rc) and np.array_equal(bits(orv), bits(rv)), name
        mirrored = int(np.sum(r != c)) if info.symmetric else 0
        assert info.nnz_csr == info.nnz + mirrored


@pytest.mark.parametrize("weighted", [None, 0, 1])
def test_io_utils_dropin_matches_reference(files, ref_io, our_io, weighted):
    """read -> coo_to_csr -> split / create_L_and_split through the
    reference's io_utils.c and through the library's definitions."""
    _, mats, adj = files
    cases = list(mats.values()) if weighted is None else list(adj.values())
    for path in cases:
        adjacency = weighted is not None
        R = ref_io.load(path, adjacency, weighted)
        G = our_io.load(path, adjacency, weighted)
        assert R["fields"] == G["fields"], path
        for key in ("A", "L", "U"):
            assert csr_equal(R[key], G[key]), (path, key)
        for key in ("D_vals", "D_ia", "D_ja"):
            assert np.array_equal(bits(R[key]) if key == "D_vals" else R[key],
                                  bits(G[key]) if key == "D_vals" else G[key]), (path, key)


def test_oracle_conversions_pinned_to_reference(files, ref_io, orc):
    _, mats, adj = files
    for path in mats.values():
        R = ref_io.load(path)
        n, m, sym, r, c, v = orc.mm_read(path)
        A = orc.coo_to_csr(n, r, c, v, sym)
        assert csr_equal(A, R["A"]), path
        Ls, d, Us = orc.split(A)
        assert csr_equal(Ls, R["L"]) and csr_equal(Us, R["U"]), path
        assert np.array_equal(bits(d), bits(R["D_vals"])), path
        # augmented triangles: numpy restatement of cg_driver2.c:95-144
        La, da, Ua = orc.split(A, augmented=True)
        if np.all(np.diff(A.ia) > 0) and path.endswith("grid_sym.mtx"):
            assert csr_equal(La, augment_lower(A)) and csr_equal(Ua, augment_upper(A))
        assert np.array_equal(bits(da), bits(d))
        # A = L + D + U
        dense = Ls.to_dense() + Us.to_dense() + np.diag(d)
        if A.n <= 2000 and path.endswith(("grid_sym.mtx",)):
            assert np.array_equal(dense, A.to_dense())
    for w in (0, 1):
        for path in adj.values():
            R = ref_io.load(path, True, w)
            n, m, sym, r, c, v = orc.mm_read(path, True)
            A = orc.coo_to_csr(n, r, c, v, sym)
            a, deg = orc.adjacency_laplacian(A, bool(w))
            assert np.array_equal(bits(a), bits(R["A"].a)), (path, w)
            if w == 0 and "pattern" in path:
                # unweighted Laplacian of a simple graph: zero row sums
                assert np.all(np.add.reduceat(a, A.ia[:-1]) == 0.0)


def test_oracle_coo_to_csr_is_stable_sort_of_fill(orc):
    """Closed form: equal (row, column) keys keep the fill order, mirrors
    right after the entries before them (a stable sort of coo_to_csr's fill)."""
    rows = np.array([2, 0, 2, 1, 2, 2], np.int32)
    cols = np.array([1, 0, 1, 0, 0, 1], np.int32)
    vals = np.arange(1.0, 7.0)
    A = orc.coo_to_csr(3, rows, cols, vals, True)
    # fill: row 0 <- (0,2) (1,4) (2,5); row 1 <- (2,1) (2,3) (0,4) (2,6);
    # row 2 <- (1,1) (1,3) (0,5) (1,6); each then sorted by column, stably
    assert A.ia.tolist() == [0, 3, 7, 11]
    assert A.ja[:3].tolist() == [0, 1, 2] and A.a[:3].tolist() == [2.0, 4.0, 5.0]
    assert A.ja[3:7].tolist() == [0, 2, 2, 2] and A.a[3:7].tolist() == [4.0, 1.0, 3.0, 6.0]
    assert A.ja[7:].tolist() == [0, 1, 1, 1] and A.a[7:].tolist() == [5.0, 1.0, 3.0, 6.0]
    # a mirror and an entry with the same key: fill order decides
    # (row 0 <- mirror of entry 0, then entry 1; row 1 <- entry 0, then mirror of 1)
    A2 = orc.coo_to_csr(2, np.array([1, 0], np.int32), np.array([0, 1], np.int32),
                        np.array([7.0, 8.0]), True)
    assert A2.a.tolist() == [7.0, 8.0, 7.0, 8.0]
    B2 = from_coo(2, [1, 0], [0, 1], [7.0, 8.0], True)
    assert B2.a.tolist() == [7.0, 8.0, 7.0, 8.0]
    B = from_coo(3, rows, cols, vals, True)  # numpy lexsort restatement (stable)
    assert np.array_equal(A.ia, B.ia) and np.array_equal(A.ja, B.ja)
    assert np.array_equal(A.a, B.a)


def test_rhs_reader_matches_reference(files, ref_io, our_io):
    d, _, _ = files
    path, b = MF.rhs_file(d, 777)
    got = IO.load_rhs(path, 777, device=False)
    assert np.array_equal(bits(got), bits(b))
    assert np.array_equal(bits(our_io.read_rhs(path, 777)), bits(ref_io.read_rhs(path, 777)))
    assert np.array_equal(bits(ref_io.read_rhs(path, 777)), bits(b))
    with pytest.raises(L.LapError):
        IO.load_rhs(path, 778, device=False)


def test_writer_round_trip(files, ref_io, tmp_path):
    from lap_amd.problems import grid_laplacian
    A = grid_laplacian([7, 6, 5], 1e-3)
    A.a *= 1.0 / 3.0  # values that need all 17 digits (symmetric still)
    path = str(tmp_path / "w.mtx")
    IO.write_mm(path, A)
    R = ref_io.load(path)
    assert csr_equal(R["A"], A)


def test_binary_csr_round_trip(tmp_path):
    from lap_amd.problems import grid_laplacian
    A = grid_laplacian([9, 8, 7], 1e-3)
    path = str(tmp_path / "a.lapcsr")
    IO.save_csr(path, A)
    assert IO.csr_info(path) == (A.n, A.nnz)
    assert csr_equal(IO.load_csr(path, device=False), A)
    assert os.path.getsize(path) == 32 + 4 * (A.n + 1) + 12 * A.nnz
    with open(path, "r+b") as f:  # truncated file
        f.truncate(os.path.getsize(path) - 8)
    with pytest.raises(L.LapError):
        IO.csr_info(path)


@pytest.mark.parametrize("body,what", [
    ("3 3 2\n1 1 1.0\n", "declares"),                 # fewer entries than declared
    ("3 3 1\n1 1 1.0\n2 2 1.0\n", "declares"),        # more entries than declared
    ("3 3 1\n4 1 1.0\n", "outside"),                  # row out of range
    ("3 3 1\n1 x 1.0\n", "malformed"),                # not a number
    ("3 3 1\n0 1 1.0\n", "malformed"),                # 0-based index
    ("3 x 1\n1 1 1.0\n", "size line"),
])
def test_parser_errors(tmp_path, body, what):
    path = tmp_path / "bad.mtx"
    path.write_text("%%MatrixMarket matrix coordinate real general\n" + body)
    with pytest.raises(L.LapError, match=what):
        IO.read_coo(str(path))
    with pytest.raises(L.LapError, match="cannot read"):
        IO.read_coo(str(tmp_path / "missing.mtx"))


def test_comments_and_blank_lines_between_entries(tmp_path, orc):
    """Comment and blank lines after the size line are skipped (the
    reference would store its previous entry again for each)."""
    path = tmp_path / "c.mtx"
    path.write_text("%%MatrixMarket matrix coordinate real general\n% c\n\n2 2 2\n"
                    "1 1 2.5\n% mid\n\n2 2 -1\n")
    info, r, c, v = IO.read_coo(str(path))
    assert r.tolist() == [0, 1] and c.tolist() == [0, 1] and v.tolist() == [2.5, -1.0]


def test_edge_case_files_match_reference(tmp_path, ref_io, our_io, orc):
    """1x1, no entries, empty rows, diagonal only, pattern general and the
    values 0, -0, nan, +-inf: parser and COO -> CSR bitwise the reference's
    (the oracle's too); split where the reference's own split is defined."""
    for name, path in MF.edge_cases(str(tmp_path)).items():
        info, r, c, v = IO.read_coo(path)
        n, m, nnz, sym, rr, rc, rv = ref_io.read(path)
        assert (info.n, info.nnz) == (n, nnz), name
        assert np.array_equal(r, rr) and np.array_equal(c, rc), name
        assert np.array_equal(bits(v), bits(rv)), name
        # the reference's split needs every diagonal entry (it sizes L and U
        # as nnz - n); the others stop after coo_to_csr
        full = name in ("one_by_one", "diagonal_only")
        R, G = ref_io.load(path, do_split=full), our_io.load(path, do_split=full)
        assert R["fields"] == G["fields"], name
        for key in (("A", "L", "U") if full else ("A",)):
            assert csr_equal(R[key], G[key]), (name, key)
        if full:
            assert np.array_equal(bits(R["D_vals"]), bits(G["D_vals"])), name
        on, om, osym, orr, occ, ov = orc.mm_read(path)
        A = orc.coo_to_csr(on, orr, occ, ov, osym)
        assert csr_equal(A, R["A"]), name
