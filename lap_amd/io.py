"""Input path: Matrix Market and binary CSR files to device matrices.

    A, info = load_mm("matrix.mtx")             # parse (host, parallel) + CSR (device)
    L, d, U = split(A, augmented=True)           # GS_std / IC0 triangles, on the device
    b = load_rhs("rhs.mtx", A.n)                 # device vector
    deg = adjacency_laplacian(G)                 # adjacency -> Laplacian, in place

The reference reads with read_mm_file / read_adjacency_file, converts with
coo_to_csr and splits on the host (io_utils.c:15-491, cg_driver2.c:95-196,
cg_driver.c:42-47) before copying every array to the device.  Here the text
is parsed by lap_mm_load on all host cores, the COO arrays are copied to the
device once, and CSR, the triangles and the Laplacian are built there
(lap_coo_to_csr, lap_csr_split, lap_adjacency_laplacian), with the
reference's arrays as the result.  save_csr / load_csr keep a binary CSR that
loads at file-read speed.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass

import numpy as np
import torch

from . import lib as L
from .problems import CSR
from .solver import DEV, DeviceCSR, _sync, ptr


@dataclass
class MMInfo:
    n: int
    m: int
    nnz: int        # COO entries (adjacency mode: with the n diagonal slots)
    nnz_csr: int    # CSR entries after mirroring
    symmetric: bool
    pattern: bool


def _call(name, *args):
    st = getattr(L.load(), name)(*args)
    if st != 0:
        L.check()
        raise L.LapError(f"{name} returned {st}")


class _Parsed:
    """Owns a lap_mm handle (COO in file order, host memory)."""

    def __init__(self, path: str, adjacency: bool):
        self.h = C.c_void_p()
        info = L.MmInfo()
        _call("lap_mm_load", str(path).encode(), int(adjacency), C.byref(self.h), C.byref(info))
        self.info = MMInfo(info.n, info.m, info.nnz, info.nnz_csr, bool(info.symmetric),
                           bool(info.pattern))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        L.load().lap_mm_free(self.h)
        self.h = None
        return False


def read_coo(path: str, adjacency: bool = False):
    """(info, rows, cols, vals): the file's COO, 0-based, in file order, as
    read_mm_file / read_adjacency_file produce it (host numpy arrays)."""
    with _Parsed(path, adjacency) as P:
        k = P.info.nnz
        rows, cols = np.empty(k, np.int32), np.empty(k, np.int32)
        vals = np.empty(k)
        _call("lap_mm_coo", P.h, rows.ctypes.data, cols.ctypes.data, vals.ctypes.data, 0)
        return P.info, rows, cols, vals


def coo_to_csr(n: int, rows: torch.Tensor, cols: torch.Tensor, vals: torch.Tensor,
               symmetric: bool, m: int | None = None) -> DeviceCSR:
    """COO (device int32/int32/float64) -> CSR on the device, rows sorted by
    column, equal to the reference's coo_to_csr."""
    m = n if m is None else m
    nnz = int(rows.shape[0])
    ia = torch.empty(n + 1, dtype=torch.int32, device=DEV)
    cnt = C.c_int(0)
    _sync()
    _call("lap_coo_to_csr", n, m, nnz, ptr(rows), ptr(cols), ptr(vals), int(symmetric), 0,
          ptr(ia), None, None, C.byref(cnt))
    ja = torch.empty(max(cnt.value, 1), dtype=torch.int32, device=DEV)
    a = torch.empty(max(cnt.value, 1), dtype=torch.float64, device=DEV)
    _call("lap_coo_to_csr", n, m, nnz, ptr(rows), ptr(cols), ptr(vals), int(symmetric),
          cnt.value, ptr(ia), ptr(ja), ptr(a), C.byref(cnt))
    return DeviceCSR(n, ia, ja[: cnt.value], a[: cnt.value])


@dataclass
class LoadTimes:
    parse_s: float
    upload_s: float
    convert_s: float

    @property
    def total_s(self) -> float:
        return self.parse_s + self.upload_s + self.convert_s


def load_mm(path: str, adjacency: bool = False, times: bool = False):
    """Matrix Market file -> (DeviceCSR, MMInfo[, LoadTimes])."""
    t0 = time.perf_counter()
    with _Parsed(path, adjacency) as P:
        t1 = time.perf_counter()
        k = P.info.nnz
        rows = torch.empty(max(k, 1), dtype=torch.int32, device=DEV)
        cols = torch.empty(max(k, 1), dtype=torch.int32, device=DEV)
        vals = torch.empty(max(k, 1), dtype=torch.float64, device=DEV)
        _sync()
        _call("lap_mm_coo", P.h, ptr(rows), ptr(cols), ptr(vals), 1)
        info = P.info
    t2 = time.perf_counter()
    A = coo_to_csr(info.n, rows[:k], cols[:k], vals[:k], info.symmetric, info.m)
    del rows, cols, vals
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    if times:
        return A, info, LoadTimes(t1 - t0, t2 - t1, t3 - t2)
    return A, info


def split(A: DeviceCSR, augmented: bool = False):
    """(L, d, U) on the device: strict triangles (split, io_utils.c:268-363)
    or, augmented, L+D with the diagonal last and D+U with it first
    (cg_driver2.c:95-196)."""
    i32 = dict(dtype=torch.int32, device=DEV)
    lia, uia = torch.empty(A.n + 1, **i32), torch.empty(A.n + 1, **i32)
    d = torch.empty(A.n, dtype=torch.float64, device=DEV)
    nl, nu = C.c_int(0), C.c_int(0)
    _sync()
    _call("lap_csr_split", A.n, ptr(A.ia), ptr(A.ja), ptr(A.a), int(augmented), 0, ptr(lia),
          None, None, 0, ptr(uia), None, None, None, C.byref(nl), C.byref(nu))
    lja, uja = torch.empty(max(nl.value, 1), **i32), torch.empty(max(nu.value, 1), **i32)
    la = torch.empty(max(nl.value, 1), dtype=torch.float64, device=DEV)
    ua = torch.empty(max(nu.value, 1), dtype=torch.float64, device=DEV)
    _call("lap_csr_split", A.n, ptr(A.ia), ptr(A.ja), ptr(A.a), int(augmented),
          max(nl.value, 1), ptr(lia), ptr(lja), ptr(la), max(nu.value, 1), ptr(uia), ptr(uja),
          ptr(ua), ptr(d), C.byref(nl), C.byref(nu))
    return (DeviceCSR(A.n, lia, lja[: nl.value], la[: nl.value]), d,
            DeviceCSR(A.n, uia, uja[: nu.value], ua[: nu.value]))


def adjacency_laplacian(A: DeviceCSR, weighted: bool = False) -> torch.Tensor:
    """In place: A (an adjacency matrix read in adjacency mode, so every row
    holds its diagonal slot) becomes the graph Laplacian exactly as
    create_L_and_split makes it (io_utils.c:383-425).  Returns the int
    degrees."""
    deg = torch.empty(A.n, dtype=torch.int32, device=DEV)
    _sync()
    _call("lap_adjacency_laplacian", A.n, ptr(A.ia), ptr(A.ja), ptr(A.a), int(weighted),
          ptr(deg))
    return deg


def load_rhs(path: str, n: int, device: bool = True):
    """First column of a Matrix Market array file (read_rhs,
    io_utils.c:239-264), as a device tensor or a numpy array."""
    if device:
        b = torch.empty(max(n, 1), dtype=torch.float64, device=DEV)
        _sync()
        _call("lap_rhs_load", str(path).encode(), n, ptr(b), 1)
        return b[:n]
    b = np.empty(n)
    _call("lap_rhs_load", str(path).encode(), n, b.ctypes.data, 0)
    return b


def write_mm(path: str, A: CSR) -> None:
    """Lower triangle of a host CSR as coordinate real symmetric (%.17g)."""
    ia = np.ascontiguousarray(A.ia, np.int32)
    ja = np.ascontiguousarray(A.ja, np.int32)
    a = np.ascontiguousarray(A.a, np.float64)
    _call("lap_mm_write", str(path).encode(), A.n, ia.ctypes.data, ja.ctypes.data, a.ctypes.data)


def save_csr(path: str, A) -> None:
    """Binary CSR (lap_csr_save) from a DeviceCSR or a host CSR."""
    if isinstance(A, DeviceCSR):
        _sync()
        _call("lap_csr_save", str(path).encode(), A.n, ptr(A.ia), ptr(A.ja), ptr(A.a))
        return
    ia = np.ascontiguousarray(A.ia, np.int32)
    ja = np.ascontiguousarray(A.ja, np.int32)
    a = np.ascontiguousarray(A.a, np.float64)
    _call("lap_csr_save", str(path).encode(), A.n, ia.ctypes.data, ja.ctypes.data, a.ctypes.data)


def csr_info(path: str):
    n, nnz = C.c_int(0), C.c_int(0)
    _call("lap_csr_info", str(path).encode(), C.byref(n), C.byref(nnz))
    return n.value, nnz.value


def load_csr(path: str, device: bool = True):
    """Binary CSR -> DeviceCSR (streamed through pinned buffers) or CSR."""
    n, nnz = csr_info(path)
    if device:
        ia = torch.empty(n + 1, dtype=torch.int32, device=DEV)
        ja = torch.empty(max(nnz, 1), dtype=torch.int32, device=DEV)
        a = torch.empty(max(nnz, 1), dtype=torch.float64, device=DEV)
        _sync()
        _call("lap_csr_load", str(path).encode(), ptr(ia), ptr(ja), ptr(a), 1)
        return DeviceCSR(n, ia, ja[:nnz], a[:nnz])
    ia, ja, a = np.empty(n + 1, np.int32), np.empty(nnz, np.int32), np.empty(nnz)
    _call("lap_csr_load", str(path).encode(), ia.ctypes.data, ja.ctypes.data, a.ctypes.data, 0)
    return CSR(n, ia, ja, a)


class MMatrix(C.Structure):
    """lap_mmatrix (include/lap_io.h): the layout of the reference's mmatrix
    (io_utils.h:27-40)."""
    _fields_ = [("coo_rows", C.POINTER(C.c_int)), ("coo_cols", C.POINTER(C.c_int)),
                ("coo_vals", C.POINTER(C.c_double)), ("csr_ia", C.POINTER(C.c_int)),
                ("csr_ja", C.POINTER(C.c_int)), ("csr_vals", C.POINTER(C.c_double)),
                ("n", C.c_int), ("m", C.c_int), ("nnz", C.c_int), ("nnz_unpacked", C.c_int),
                ("symmetric", C.c_int)]


def host_load(path: str, adjacency: bool = False, weighted: int | None = None):
    """The reference's host set-up through the library's io_utils.h
    definitions: read_mm_file (or read_adjacency_file), coo_to_csr, then
    split (or create_L_and_split(weighted)).  Returns (A, L, d, U) as host
    CSR / numpy arrays."""
    lib = L.load()
    P = C.POINTER(MMatrix)
    for nm in ("read_mm_file", "read_adjacency_file"):
        getattr(lib, nm).argtypes = [C.c_char_p, P]
        getattr(lib, nm).restype = None
    lib.coo_to_csr.argtypes = [P]
    lib.coo_to_csr.restype = None
    lib.split.argtypes = [P, P, P, P]
    lib.split.restype = None
    lib.create_L_and_split.argtypes = [P, P, P, P, C.c_int]
    lib.create_L_and_split.restype = None
    libc = C.CDLL(None)
    libc.free.argtypes = [C.c_void_p]
    A, Lm, U, D = MMatrix(), MMatrix(), MMatrix(), MMatrix()
    (lib.read_adjacency_file if adjacency else lib.read_mm_file)(str(path).encode(), C.byref(A))
    if A.n == 0:
        L.check()
        raise L.LapError(f"cannot read {path}")
    lib.coo_to_csr(C.byref(A))
    if weighted is None:
        lib.split(C.byref(A), C.byref(Lm), C.byref(U), C.byref(D))
    else:
        lib.create_L_and_split(C.byref(A), C.byref(Lm), C.byref(U), C.byref(D), int(weighted))
    n = A.n

    def take(M, with_vals=True):
        ia = np.ctypeslib.as_array(M.csr_ia, shape=(n + 1,)).copy()
        k = int(ia[-1])
        ja = np.ctypeslib.as_array(M.csr_ja, shape=(k,)).copy() if k else np.zeros(0, np.int32)
        a = np.ctypeslib.as_array(M.csr_vals, shape=(k,)).copy() if k else np.zeros(0)
        return CSR(n, ia, ja, a)

    out = (take(A), take(Lm), np.ctypeslib.as_array(D.csr_vals, shape=(n,)).copy(), take(U))
    for M in (A, Lm, U, D):
        for p in (M.csr_ia, M.csr_ja, M.csr_vals):
            if p:
                libc.free(C.cast(p, C.c_void_p))
    return out
