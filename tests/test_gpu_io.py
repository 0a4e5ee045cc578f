"""Input path on the device (lap_coo_to_csr, lap_csr_split,
lap_adjacency_laplacian, lap_rhs_load / lap_csr_load to device).

Bars (bitwise): the device CSR, triangles, diagonal and Laplacian values
equal the oracle's orc_coo_to_csr / orc_split / orc_adjacency_laplacian,
which tests/test_io.py pins bitwise to the reference's io_utils.c.  Cases:
shuffled symmetric grid, general matrix with duplicates and rows of 5000
entries (the LDS bitonic class), symmetric file with rows of 9000 and 20000
entries (the multi-run merge class), pattern and real-weighted adjacency
files.  A matrix loaded from a file solves to the same bits as the matrix
it was written from.
"""
import numpy as np
import pytest
import torch

import mmfiles as MF
from lap_amd import io as IO
from lap_amd import lib as L
from lap_amd import solver as S
from lap_amd.problems import CSR, grid_laplacian, rhs_normal

pytestmark = pytest.mark.gpu


def host(A):
    return CSR(A.n, A.ia.cpu().numpy(), A.ja.cpu().numpy(), A.a.cpu().numpy())


def same(X, Y):
    return (X.n == Y.n and np.array_equal(X.ia, Y.ia) and np.array_equal(X.ja, Y.ja)
            and np.array_equal(X.a.view(np.int64), Y.a.view(np.int64)))


@pytest.fixture(scope="module")
def files(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("mm"))
    return ({f.__name__: f(d) for f in MF.ALL_MATRICES},
            {f.__name__: f(d) for f in MF.ALL_ADJACENCY}, d)


@pytest.mark.parametrize("name", [f.__name__ for f in MF.ALL_MATRICES])
def test_device_csr_and_split_match_oracle(orc, files, name):
    path = files[0][name]
    print(f"[case] coo_to_csr + split {name}", flush=True)
    A, info = IO.load_mm(path)
    n, m, sym, r, c, v = orc.mm_read(path)
    Ao = orc.coo_to_csr(n, r, c, v, sym)
    assert info.nnz_csr == Ao.nnz
    assert same(host(A), Ao)
    rows = np.diff(Ao.ia)
    print(f"[diag] {name}: n {n}, nnz {Ao.nnz}, longest row {rows.max()}", flush=True)
    for aug in (False, True):
        Ld, dd, Ud = IO.split(A, augmented=aug)
        Lo, do, Uo = orc.split(Ao, augmented=aug)
        assert same(host(Ld), Lo) and same(host(Ud), Uo), aug
        assert np.array_equal(dd.cpu().numpy().view(np.int64), do.view(np.int64))


@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("name", [f.__name__ for f in MF.ALL_ADJACENCY])
def test_device_adjacency_laplacian_matches_oracle(orc, files, name, weighted):
    path = files[1][name]
    print(f"[case] adjacency {name} weighted={weighted}", flush=True)
    A, info = IO.load_mm(path, adjacency=True)
    n, m, sym, r, c, v = orc.mm_read(path, True)
    Ao = orc.coo_to_csr(n, r, c, v, sym)
    assert same(host(A), Ao)
    deg = IO.adjacency_laplacian(A, weighted)
    ao, dego = orc.adjacency_laplacian(Ao, weighted)
    assert np.array_equal(deg.cpu().numpy(), dego)
    assert np.array_equal(A.a.cpu().numpy().view(np.int64), ao.view(np.int64))


def test_loaded_matrix_solves_like_its_source(tmp_path):
    A = grid_laplacian([24, 20, 16], 1e-3)
    path = str(tmp_path / "g.mtx")
    IO.write_mm(path, A)
    D, info, t = IO.load_mm(path, times=True)
    assert same(host(D), A)
    b = rhs_normal(A.n, 3)
    rpath = str(tmp_path / "b.mtx")
    with open(rpath, "w") as f:
        f.write(f"%%MatrixMarket matrix array real general\n{A.n} 1\n")
        f.write("".join(f"{x:.17g}\n" for x in b))
    bd = IO.load_rhs(rpath, A.n)
    assert np.array_equal(bd.cpu().numpy(), b)
    P1 = S.make_preconditioner(D, A, "line_jacobi")
    D2 = S.DeviceCSR.from_csr(A)
    P2 = S.make_preconditioner(D2, A, "line_jacobi")
    r1 = S.pcg(D, bd, P1, tol=1e-10, maxit=5000)
    r2 = S.pcg(D2, S.to_dev(b), P2, tol=1e-10, maxit=5000)
    assert r1.iterations == r2.iterations
    assert torch.equal(r1.x, r2.x)
    print(f"[diag] load 24x20x16: parse {t.parse_s*1e3:.1f} ms, upload {t.upload_s*1e3:.1f} ms,"
          f" convert {t.convert_s*1e3:.1f} ms", flush=True)


def test_binary_csr_device_round_trip(tmp_path):
    A = grid_laplacian([40, 30, 20], 1e-3)
    D = S.DeviceCSR.from_csr(A)
    path = str(tmp_path / "a.lapcsr")
    IO.save_csr(path, D)  # from device memory
    E = IO.load_csr(path)  # to device memory through pinned buffers
    assert same(host(E), A)
    assert same(IO.load_csr(path, device=False), A)


def test_device_entry_points_reject_bad_input():
    rows = torch.tensor([0, 5], dtype=torch.int32, device="cuda")
    cols = torch.tensor([0, 1], dtype=torch.int32, device="cuda")
    vals = torch.ones(2, dtype=torch.float64, device="cuda")
    with pytest.raises(L.LapError, match="outside"):
        IO.coo_to_csr(3, rows, cols, vals, False)
    rows = torch.tensor([0, 2], dtype=torch.int32, device="cuda")
    A = IO.coo_to_csr(3, rows, cols, vals, True)
    assert A.ia.cpu().tolist() == [0, 1, 2, 3] and A.ja.cpu().tolist() == [0, 2, 1]
    ia = torch.empty(4, dtype=torch.int32, device="cuda")
    ja = torch.empty(1, dtype=torch.int32, device="cuda")
    a = torch.empty(1, dtype=torch.float64, device="cuda")
    import ctypes as C
    cnt = C.c_int(0)
    st = L.load().lap_coo_to_csr(3, 3, 2, S.ptr(rows), S.ptr(cols), S.ptr(vals), 1, 1,
                                 S.ptr(ia), S.ptr(ja), S.ptr(a), C.byref(cnt))
    assert st == 1 and cnt.value == 3  # capacity 1 < 3 entries: refused, size reported
    with pytest.raises(L.LapError, match="capacity"):
        L.check()
    with pytest.raises(L.LapError):  # host pointers are refused
        L.load().lap_coo_to_csr(3, 3, 2, rows.cpu().numpy().ctypes.data, None, None, 0, 0,
                                None, None, None, C.byref(cnt))
        L.check()


def test_edge_case_files_on_the_device(orc, tmp_path):
    """The same edge files through lap_coo_to_csr and lap_csr_split: bitwise
    the oracle (NaN and infinities carried as bits)."""
    for name, path in MF.edge_cases(str(tmp_path)).items():
        print(f"[case] edge file {name}", flush=True)
        A, info = IO.load_mm(path)
        n, m, sym, r, c, v = orc.mm_read(path)
        Ao = orc.coo_to_csr(n, r, c, v, sym)
        assert same(host(A), Ao), name
        for aug in (False, True):
            Ld, dd, Ud = IO.split(A, augmented=aug)
            Lo, do, Uo = orc.split(Ao, augmented=aug)
            assert same(host(Ld), Lo) and same(host(Ud), Uo), (name, aug)
            assert np.array_equal(dd.cpu().numpy().view(np.int64), do.view(np.int64)), name
