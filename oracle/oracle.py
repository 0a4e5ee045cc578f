"""ctypes bindings of the oracle -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.

Loads ``oracle/liblaporacle.so`` (the C restatement of the reference's serial
path) and, when it has been built, ``oracle/_ref/liblapref_cpu.so`` /
``liblapref_omp.so`` (the reference's own CPU path compiled from its sources
by ``oracle/build_ref.sh``).  Only tests/, ``__graft_entry__.smoke()`` and
bench.py's cpu_baseline leg import this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liblaporacle.so")
REF_CPU = os.path.join(HERE, "_ref", "liblapref_cpu.so")
REF_OMP = os.path.join(HERE, "_ref", "liblapref_omp.so")

PI = C.POINTER(C.c_int)
PD = C.POINTER(C.c_double)

PREC_KINDS = {"none": 0, "line_jacobi": 1, "it_jacobi": 2, "GS_std": 3,
              "GS_it": 4, "GS_it2": 5, "ichol": 6}


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _load(path):
    if not os.path.exists(path):
        build()
    return C.CDLL(path)


def ip(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(PI)


def dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(PD)


class OrcPrec(C.Structure):
    _fields_ = [("kind", C.c_int), ("n", C.c_int), ("m", C.c_int), ("k", C.c_int),
                ("lia", PI), ("lja", PI), ("la", PD),
                ("uia", PI), ("uja", PI), ("ua", PD),
                ("d", PD), ("d_r", PD),
                ("aux1", PD), ("aux2", PD), ("aux3", PD)]


class RefPData(C.Structure):
    """Layout of the reference's pdata (common.h:29-50)."""
    _fields_ = [("lia", PI), ("lja", PI), ("la", PD), ("lnnz", C.c_int),
                ("uia", PI), ("uja", PI), ("ua", PD), ("unnz", C.c_int),
                ("ichol_vals", PD), ("d", PD), ("d_r", PD), ("n", C.c_int),
                ("aux_vec1", PD), ("aux_vec2", PD), ("aux_vec3", PD),
                ("prec_op", C.c_char_p), ("m", C.c_int), ("k", C.c_int)]


class Oracle:
    def __init__(self, path: str = LIB):
        L = self.lib = _load(path)
        L.orc_dot.restype = C.c_double
        L.orc_dot.argtypes = [C.c_int, PD, PD]
        L.orc_scal.argtypes = [C.c_int, C.c_double, PD]
        L.orc_axpy.argtypes = [C.c_int, C.c_double, PD, PD]
        L.orc_csr_matvec.argtypes = [C.c_int, PI, PI, PD, PD, PD, C.c_double, C.c_double]
        L.orc_omp_csr_matvec.argtypes = [C.c_int, PI, PI, PD, PD, PD, C.c_double, C.c_double, C.c_int]
        for f in (L.orc_lower_solve, L.orc_upper_solve):
            f.argtypes = [C.c_int, PI, PI, PD, PD, PD, PD]
        for f in (L.orc_vec_vec,):
            f.argtypes = [C.c_int, PD, PD, PD]
        for f in (L.orc_reciprocal, L.orc_sqrt):
            f.argtypes = [C.c_int, PD, PD]
        L.orc_ichol_factor.argtypes = [C.c_int, PI, PI, PD, PI, PI, PD]
        L.orc_ichol_apply.argtypes = [C.c_int, PI, PI, PD, PI, PI, PD, PD, PD, PD]
        L.orc_prec_apply.argtypes = [PI, PI, PD, C.POINTER(OrcPrec), PD, PD]
        L.orc_cg.argtypes = [C.c_int, PI, PI, PD, PD, PD, C.c_double, C.POINTER(OrcPrec),
                             C.c_int, PI, PI, PD, C.c_int]
        L.orc_omp_cg.argtypes = [C.c_int, PI, PI, PD, PD, PD, C.c_double, C.POINTER(OrcPrec),
                                 C.c_int, PI, PI, PD, C.c_int, C.c_int]
        L.orc_graph_values.argtypes = [C.c_int, C.c_int, PI, PI, PI, PI, PD, C.c_double, PD]
        L.orc_edge_grad.argtypes = [C.c_int, PI, PI, PD, PD, PD]
        L.orc_mm_read.restype = C.c_long
        L.orc_mm_read.argtypes = [C.c_char_p, C.c_int, PI, PI, PI, PI, PI, PD]
        L.orc_coo_to_csr.restype = C.c_long
        L.orc_coo_to_csr.argtypes = [C.c_int, C.c_long, PI, PI, PD, C.c_int, PI, PI, PD]
        L.orc_split.restype = C.c_long
        L.orc_split.argtypes = [C.c_int, PI, PI, PD, C.c_int, PI, PI, PD, PI, PI, PD, PD,
                                C.POINTER(C.c_long)]
        L.orc_adjacency_laplacian.argtypes = [C.c_int, PI, PI, PD, C.c_int, PI]

    # -- kernels --------------------------------------------------------
    def dot(self, v, w):
        return self.lib.orc_dot(v.shape[0], dp(v), dp(w))

    def matvec(self, A, x, y=None, alpha=1.0, beta=0.0):
        y = np.zeros(A.n) if y is None else y.copy()
        self.lib.orc_csr_matvec(A.n, ip(A.ia), ip(A.ja), dp(A.a), dp(x), dp(y), alpha, beta)
        return y

    def graph_values(self, n, ia, ja, epos, dpos, w, shift):
        """Values of L(w) on a graph pattern (orc_graph_values)."""
        a = np.zeros(ja.shape[0])
        self.lib.orc_graph_values(n, w.shape[0], ip(ia), ip(ja), ip(epos), ip(dpos), dp(w),
                                  shift, dp(a))
        return a

    def edge_grad(self, src, dst, lam, x):
        g = np.zeros(src.shape[0])
        self.lib.orc_edge_grad(src.shape[0], ip(src), ip(dst), dp(lam), dp(x), dp(g))
        return g

    # -- input path -------------------------------------------------------
    def mm_read(self, path: str, adjacency: bool = False):
        """(n, m, symmetric, rows, cols, vals): 0-based COO in file order."""
        n, m, sym = C.c_int(0), C.c_int(0), C.c_int(0)
        k = self.lib.orc_mm_read(path.encode(), int(adjacency), C.byref(n), C.byref(m),
                                 C.byref(sym), None, None, None)
        if k < 0:
            raise ValueError(f"oracle cannot parse {path}")
        rows, cols = np.zeros(k, np.int32), np.zeros(k, np.int32)
        vals = np.zeros(k)
        self.lib.orc_mm_read(path.encode(), int(adjacency), C.byref(n), C.byref(m),
                             C.byref(sym), ip(rows), ip(cols), dp(vals))
        return n.value, m.value, bool(sym.value), rows, cols, vals

    def coo_to_csr(self, n, rows, cols, vals, symmetric):
        from lap_amd.problems import CSR
        rows = np.ascontiguousarray(rows, np.int32)
        cols = np.ascontiguousarray(cols, np.int32)
        vals = np.ascontiguousarray(vals, np.float64)
        ia = np.zeros(n + 1, np.int32)
        tot = self.lib.orc_coo_to_csr(n, rows.shape[0], ip(rows), ip(cols), dp(vals),
                                      int(symmetric), ip(ia), None, None)
        ja, a = np.zeros(tot, np.int32), np.zeros(tot)
        self.lib.orc_coo_to_csr(n, rows.shape[0], ip(rows), ip(cols), dp(vals), int(symmetric),
                                ip(ia), ip(ja), dp(a))
        return CSR(n, ia, ja, a)

    def split(self, A, augmented: bool = False):
        """(L, d, U) of A = L + D + U; augmented adds the diagonal to both."""
        from lap_amd.problems import CSR
        lia, uia = np.zeros(A.n + 1, np.int32), np.zeros(A.n + 1, np.int32)
        nu = C.c_long(0)
        nl = self.lib.orc_split(A.n, ip(A.ia), ip(A.ja), dp(A.a), int(augmented), ip(lia), None,
                                None, ip(uia), None, None, None, C.byref(nu))
        lja, la = np.zeros(nl, np.int32), np.zeros(nl)
        uja, ua = np.zeros(nu.value, np.int32), np.zeros(nu.value)
        d = np.zeros(A.n)
        self.lib.orc_split(A.n, ip(A.ia), ip(A.ja), dp(A.a), int(augmented), ip(lia), ip(lja),
                           dp(la), ip(uia), ip(uja), dp(ua), dp(d), C.byref(nu))
        return CSR(A.n, lia, lja, la), d, CSR(A.n, uia, uja, ua)

    def adjacency_laplacian(self, A, weighted: bool = False):
        """(values of the Laplacian on A's pattern, int degrees)."""
        a = A.a.copy()
        deg = np.zeros(A.n, np.int32)
        self.lib.orc_adjacency_laplacian(A.n, ip(A.ia), ip(A.ja), dp(a), int(weighted), ip(deg))
        return a, deg

    def lower_solve(self, T, x, diag=None):
        y = np.zeros(T.n)
        self.lib.orc_lower_solve(T.n, ip(T.ia), ip(T.ja), dp(T.a),
                                 dp(diag) if diag is not None else None, dp(x), dp(y))
        return y

    def upper_solve(self, T, x, diag=None):
        y = np.zeros(T.n)
        self.lib.orc_upper_solve(T.n, ip(T.ia), ip(T.ja), dp(T.a),
                                 dp(diag) if diag is not None else None, dp(x), dp(y))
        return y

    def ichol_factor(self, A):
        """IC0 of A (full symmetric CSR): returns (Lfac, Ufac) as CSR with
        the diagonal last / first, as the reference CPU driver sets up."""
        from lap_amd.problems import augment_lower, augment_upper
        Lf, Uf = augment_lower(A), augment_upper(A)
        self.lib.orc_ichol_factor(A.n, ip(Uf.ia), ip(Uf.ja), dp(Uf.a),
                                  ip(Lf.ia), ip(Lf.ja), dp(Lf.a))
        return Lf, Uf

    # -- preconditioned CG -------------------------------------------------
    def make_prec(self, A, kind: str, m: int = 1, k: int = 1):
        """Operands exactly as cg_driver2.c sets them up on the CPU path."""
        from lap_amd.problems import split
        L, d, U = split(A)
        keep = {}
        if kind == "ichol":
            L, U = self.ichol_factor(A)
        d_r = np.zeros(A.n)
        self.lib.orc_reciprocal(A.n, dp(d), dp(d_r))
        aux = [np.zeros(A.n) for _ in range(3)]
        keep.update(L=L, U=U, d=d, d_r=d_r, aux=aux)
        p = OrcPrec(PREC_KINDS[kind], A.n, m, k, ip(L.ia), ip(L.ja), dp(L.a),
                    ip(U.ia), ip(U.ja), dp(U.a), dp(d), dp(d_r),
                    dp(aux[0]), dp(aux[1]), dp(aux[2]))
        return p, keep

    def prec_apply(self, A, prec, x):
        p, _keep = prec
        y = np.zeros(A.n)
        self.lib.orc_prec_apply(ip(A.ia), ip(A.ja), dp(A.a), C.byref(p), dp(x), dp(y))
        return y

    def cg(self, A, b, kind="line_jacobi", tol=1e-8, maxit=25000, m=1, k=1, x0=None,
           prec=None, quiet=True, omp=None):
        prec = prec or self.make_prec(A, kind, m, k)
        p, _keep = prec
        x = np.zeros(A.n) if x0 is None else x0.copy()
        hist = np.zeros(maxit + 2)
        it, flag = C.c_int(0), C.c_int(0)
        if omp is None:
            self.lib.orc_cg(A.n, ip(A.ia), ip(A.ja), dp(A.a), dp(x), dp(b), tol, C.byref(p),
                            maxit, C.byref(it), C.byref(flag), dp(hist), 1 if quiet else 0)
        else:
            spmv_threads, dot_threads = omp
            self.lib.orc_omp_cg(A.n, ip(A.ia), ip(A.ja), dp(A.a), dp(x), dp(b), tol,
                                C.byref(p), maxit, C.byref(it), C.byref(flag), dp(hist),
                                spmv_threads, dot_threads)
        return x, it.value, flag.value, hist[: it.value + 1].copy()


def have_ref() -> bool:
    return os.path.exists(REF_CPU)


class Reference:
    """The reference's own CPU path (compiled from its sources into _ref/)."""

    def __init__(self, path: str = REF_CPU):
        L = self.lib = C.CDLL(path)
        self.omp = path.endswith("omp.so")
        pre = "openmp_" if self.omp else "simple_"
        self.pre = pre
        getattr(L, pre + "dot").restype = C.c_double
        getattr(L, pre + "dot").argtypes = [C.c_int, PD, PD]
        getattr(L, pre + "csr_matvec").argtypes = [C.c_int, C.c_int, PI, PI, PD, PD, PD, PD, PD]
        for nm in ("lower_triangular_solve", "upper_triangular_solve"):
            getattr(L, pre + nm).argtypes = [C.c_int, C.c_int, PI, PI, PD, PD, PD, PD]
        L.initialize_ichol.argtypes = [C.c_int, C.c_int, PI, PI, PD, PI, PI, PD]
        L.prec_function.argtypes = [PI, PI, PD, C.c_int, C.POINTER(RefPData), PD, PD]
        L.cg.argtypes = [C.c_int, C.c_double, PI, PI, PD, PD, PD, C.c_double,
                         C.POINTER(RefPData), C.c_int, PI, PI, PD]

    def fn(self, name):
        return getattr(self.lib, self.pre + name)

    def make_pdata(self, A, kind: str, m: int = 1, k: int = 1):
        """pdata exactly as cg_driver2.c builds it on the CPU path (:146-196,
        :354-387): strict L/U, or the IC0 factors for ichol."""
        from lap_amd.problems import augment_lower, augment_upper, split
        L, d, U = split(A)
        if kind == "ichol":
            L, U = augment_lower(A), augment_upper(A)
            self.lib.initialize_ichol(A.n, U.nnz, ip(U.ia), ip(U.ja), dp(U.a),
                                      ip(L.ia), ip(L.ja), dp(L.a))
        d_r = np.where(d != 0.0, 1.0 / np.where(d != 0.0, d, 1.0), 0.0)
        aux = [np.zeros(A.n) for _ in range(3)]
        name = kind.encode()
        pd = RefPData(ip(L.ia), ip(L.ja), dp(L.a), L.nnz, ip(U.ia), ip(U.ja), dp(U.a), U.nnz,
                      None, dp(d), dp(d_r), A.n, dp(aux[0]), dp(aux[1]), dp(aux[2]),
                      name, m, k)
        return pd, dict(L=L, U=U, d=d, d_r=d_r, aux=aux, name=name)

    def prec_apply(self, A, pdk, x):
        pd, _keep = pdk
        y = np.zeros(A.n)
        self.lib.prec_function(ip(A.ia), ip(A.ja), dp(A.a), A.nnz, C.byref(pd), dp(x), dp(y))
        return y

    def cg(self, A, b, kind="line_jacobi", tol=1e-8, maxit=25000, m=1, k=1, pdk=None):
        """Calls the reference's cg(); its per-iteration log goes to fd 1."""
        pdk = pdk or self.make_pdata(A, kind, m, k)
        pd, _keep = pdk
        x = np.zeros(A.n)
        bb = b.copy()
        hist = np.zeros(maxit + 2)
        it, flag = C.c_int(0), C.c_int(0)
        with _silence_stdout():
            self.lib.cg(A.n, float(A.nnz), ip(A.ia), ip(A.ja), dp(A.a), dp(x), dp(bb), tol,
                        C.byref(pd), maxit, C.byref(it), C.byref(flag), dp(hist))
        return x, it.value, flag.value, hist[: it.value + 1].copy()


class MMatrix(C.Structure):
    """Layout of the reference's mmatrix (io_utils.h:27-40)."""
    _fields_ = [("coo_rows", PI), ("coo_cols", PI), ("coo_vals", PD),
                ("csr_ia", PI), ("csr_ja", PI), ("csr_vals", PD),
                ("n", C.c_int), ("m", C.c_int), ("nnz", C.c_int), ("nnz_unpacked", C.c_int),
                ("symmetric", C.c_int)]


class IoUtils:
    """Calls the io_utils.h functions (read_mm_file, read_adjacency_file,
    coo_to_csr, read_rhs, split, create_L_and_split) of a shared library that
    defines them -- the reference's compiled CPU path or a library that
    replaces io_utils.c -- and copies every result into numpy arrays."""

    def __init__(self, lib: C.CDLL):
        self.lib = lib
        P = C.POINTER(MMatrix)
        for nm in ("read_mm_file", "read_adjacency_file"):
            getattr(lib, nm).argtypes = [C.c_char_p, P]
            getattr(lib, nm).restype = None
        lib.coo_to_csr.argtypes = [P]
        lib.coo_to_csr.restype = None
        lib.read_rhs.argtypes = [C.c_char_p, PD]
        lib.read_rhs.restype = None
        lib.split.argtypes = [P, P, P, P]
        lib.split.restype = None
        lib.create_L_and_split.argtypes = [P, P, P, P, C.c_int]
        lib.create_L_and_split.restype = None
        self.libc = C.CDLL(None)
        self.libc.free.argtypes = [C.c_void_p]

    @staticmethod
    def _arr(p, count, dtype):
        if count <= 0 or not p:
            return np.zeros(0, dtype)
        return np.ctypeslib.as_array(p, shape=(count,)).astype(dtype, copy=True)

    def _free(self, *ptrs):
        for p in ptrs:
            if p:
                self.libc.free(C.cast(p, C.c_void_p))

    def read(self, path: str, adjacency: bool = False):
        """COO after read_mm_file / read_adjacency_file:
        (n, m, nnz, symmetric, rows, cols, vals)."""
        A = MMatrix()
        fn = self.lib.read_adjacency_file if adjacency else self.lib.read_mm_file
        with _silence_stdout():
            fn(path.encode(), C.byref(A))
        out = (A.n, A.m, A.nnz, A.symmetric, self._arr(A.coo_rows, A.nnz, np.int32),
               self._arr(A.coo_cols, A.nnz, np.int32), self._arr(A.coo_vals, A.nnz, np.float64))
        self._free(A.coo_rows, A.coo_cols, A.coo_vals)
        return out

    def _csr(self, M, n, nnz):
        from lap_amd.problems import CSR
        ia = self._arr(M.csr_ia, n + 1, np.int32)
        return CSR(n, ia, self._arr(M.csr_ja, int(ia[-1]), np.int32),
                   self._arr(M.csr_vals, int(ia[-1]), np.float64))

    def load(self, path: str, adjacency: bool = False, weighted: int | None = None,
             do_split: bool = True):
        """read -> coo_to_csr -> split (or create_L_and_split when weighted is
        not None).  Returns dict(A, L, U, D, fields) with every array copied.
        do_split=False stops after coo_to_csr: the reference's split sizes L
        and U as nnz - n (io_utils.c:281-287), so a file without every
        diagonal entry makes it write out of bounds."""
        A, L, U, D = MMatrix(), MMatrix(), MMatrix(), MMatrix()
        fn = self.lib.read_adjacency_file if adjacency else self.lib.read_mm_file
        with _silence_stdout():
            fn(path.encode(), C.byref(A))
            self.lib.coo_to_csr(C.byref(A))
            if not do_split:
                pass
            elif weighted is None:
                self.lib.split(C.byref(A), C.byref(L), C.byref(U), C.byref(D))
            else:
                self.lib.create_L_and_split(C.byref(A), C.byref(L), C.byref(U), C.byref(D),
                                            int(weighted))
        n = A.n
        if not do_split:
            res = dict(A=self._csr(A, n, A.nnz_unpacked),
                       fields=dict(n=A.n, m=A.m, nnz=A.nnz, nnz_unpacked=A.nnz_unpacked,
                                   symmetric=A.symmetric))
            self._free(A.csr_ia, A.csr_ja, A.csr_vals)
            return res
        res = dict(A=self._csr(A, n, A.nnz_unpacked), L=self._csr(L, n, L.nnz),
                   U=self._csr(U, n, U.nnz),
                   D_vals=self._arr(D.csr_vals, n, np.float64),
                   D_ia=self._arr(D.csr_ia, n + 1, np.int32),
                   D_ja=self._arr(D.csr_ja, n, np.int32),
                   fields=dict(n=A.n, m=A.m, nnz=A.nnz, nnz_unpacked=A.nnz_unpacked,
                               symmetric=A.symmetric, L_nnz=L.nnz, U_nnz=U.nnz, D_nnz=D.nnz))
        for M in (A, L, U, D):
            self._free(M.csr_ia, M.csr_ja, M.csr_vals)
        return res

    def read_rhs(self, path: str, n: int):
        b = np.zeros(n)
        with _silence_stdout():
            self.lib.read_rhs(path.encode(), dp(b))
        return b


class _silence_stdout:
    """Redirect fd 1 to /dev/null around a C call that prints per iteration."""

    def __enter__(self):
        import sys
        sys.stdout.flush()
        C.CDLL(None).fflush(None)
        self.saved = os.dup(1)
        self.null = os.open(os.devnull, os.O_WRONLY)
        os.dup2(self.null, 1)
        return self

    def __exit__(self, *exc):
        C.CDLL(None).fflush(None)
        os.dup2(self.saved, 1)
        os.close(self.saved)
        os.close(self.null)
        return False
