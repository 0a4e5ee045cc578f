"""Host-side mirror of the reference's solver setup, over the C ABI.

Device memory comes from torch (int32 / float64 CUDA tensors); every compute
call goes through liblapgfx.so.  The setup follows the reference's GPU
driver (cg_driver2.c:95-351): GS_std gets L+D / D+U with the pivots inside
the matrices, ichol gets a device copy of A's values factored by
initialize_ichol, d_r = 1/diag(A) comes from hip_vector_reciprocal.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np
import torch

from . import lib as L
from .problems import CSR, diagonal

DEV = "cuda"


def ptr(t):
    """Device pointer of a contiguous CUDA tensor (None passes NULL).  An
    empty view of a non-empty buffer passes the buffer's address (torch
    reports 0 for it), so zero-length arrays still name device memory."""
    if t is None:
        return None
    if not t.is_cuda or not t.is_contiguous():
        raise L.LapError("expected a contiguous CUDA tensor")
    if t.numel() == 0:
        base = t.untyped_storage().data_ptr()
        return C.c_void_p(base + t.storage_offset() * t.element_size() if base else 0)
    return C.c_void_p(t.data_ptr())


def _sync():
    torch.cuda.synchronize()


@dataclass
class DeviceCSR:
    n: int
    ia: torch.Tensor
    ja: torch.Tensor
    a: torch.Tensor

    @property
    def nnz(self) -> int:
        return int(self.ja.shape[0])

    @classmethod
    def from_csr(cls, A: CSR, device: str = DEV) -> "DeviceCSR":
        A.check()
        D = cls(A.n, torch.from_numpy(A.ia).to(device), torch.from_numpy(A.ja).to(device),
                torch.from_numpy(np.ascontiguousarray(A.a)).to(device))
        D.forget()
        return D

    def forget(self) -> None:
        """The torch allocator may hand these arrays addresses that held
        another matrix: drop any plan cached for those addresses."""
        lib = L.load()
        torch.cuda.synchronize()
        for t in (self.ia, self.ja, self.a):
            lib.lap_forget(C.c_void_p(t.data_ptr()))


def to_dev(v: np.ndarray) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(v, dtype=np.float64)).to(DEV)


# ----------------------------------------------------------------- kernels

def spmv(A: DeviceCSR, x: torch.Tensor, y: torch.Tensor | None = None, alpha: float = 1.0,
         beta: float = 0.0) -> torch.Tensor:
    lib = L.load()
    if y is None:
        y = torch.zeros(A.n, dtype=torch.float64, device=x.device)
    al, be = C.c_double(alpha), C.c_double(beta)
    _sync()
    lib.hip_csr_matvec(A.n, A.nnz, ptr(A.ia), ptr(A.ja), ptr(A.a), ptr(x), ptr(y),
                       C.byref(al), C.byref(be), b"A")
    L.check()
    return y


def dot(x: torch.Tensor, y: torch.Tensor) -> float:
    _sync()
    v = L.load().hip_dot(x.shape[0], ptr(x), ptr(y))
    L.check()
    return v


def tri_solve(T: DeviceCSR, x: torch.Tensor, lower: bool, diag: torch.Tensor | None = None,
              out: torch.Tensor | None = None) -> torch.Tensor:
    lib = L.load()
    out = torch.empty_like(x) if out is None else out
    fn = lib.hip_lower_triangular_solve if lower else lib.hip_upper_triangular_solve
    _sync()
    fn(T.n, T.nnz, ptr(T.ia), ptr(T.ja), ptr(T.a), ptr(diag), ptr(x), ptr(out))
    L.check()
    return out


def ichol_factor(A: DeviceCSR) -> torch.Tensor:
    """IC0 in A's full pattern (lower triangle holds L), as initialize_ichol."""
    vals = A.a.clone()
    _sync()
    L.load().initialize_ichol(A.n, A.nnz, ptr(A.ia), ptr(A.ja), ptr(vals))
    L.check()
    return vals


def ichol_refactor(A: DeviceCSR, vals: torch.Tensor) -> None:
    """Factors A's values held in `vals` in place again (initialize_ichol on
    the same buffer), e.g. after the matrix values changed."""
    _sync()
    L.load().initialize_ichol(A.n, A.nnz, ptr(A.ia), ptr(A.ja), ptr(vals))
    L.check()


# -------------------------------------------------------- preconditioners

@dataclass
class Preconditioner:
    """Operands of one preconditioner, kept alive with the pdata pointing at
    them."""
    kind: str
    pdata: L.PData
    keep: dict = field(default_factory=dict)


def make_preconditioner(A: DeviceCSR, host: CSR, kind: str, m: int = 1, k: int = 1,
                        ) -> Preconditioner:
    """Mirrors cg_driver2.c's GPU-path setup (:95-144, :219-252, :328-338)."""
    from .io import split as device_split  # (io imports this module)
    lib = L.load()
    n = A.n
    # the strict triangles and the diagonal, or for GS_std L+D and D+U, split
    # on the device (lap_csr_split: the same arrays as a host split of `host`,
    # which describes the same matrix)
    Ld, d_dev, Ud = device_split(A, augmented=(kind == "GS_std"))
    d_r = torch.empty(n, dtype=torch.float64, device=DEV)
    _sync()
    lib.hip_vector_reciprocal(n, ptr(d_dev), ptr(d_r))
    aux = [torch.zeros(n, dtype=torch.float64, device=DEV) for _ in range(3)]
    keep = dict(L=Ld, U=Ud, d=d_dev, d_r=d_r, aux=aux, name=kind.encode())
    ichol_vals = None
    if kind == "GS_std":
        lib.initialize_and_analyze_L_and_U_solve(n, Ld.nnz, ptr(Ld.ia), ptr(Ld.ja), ptr(Ld.a),
                                                 Ud.nnz, ptr(Ud.ia), ptr(Ud.ja), ptr(Ud.a))
    if kind == "ichol":
        ichol_vals = A.a.clone()
        lib.lap_forget(ptr(ichol_vals))
        lib.initialize_ichol(n, A.nnz, ptr(A.ia), ptr(A.ja), ptr(ichol_vals))
        keep["ichol_vals"] = ichol_vals
    L.check()
    pd = L.PData(ptr(Ld.ia), ptr(Ld.ja), ptr(Ld.a), Ld.nnz, ptr(Ud.ia), ptr(Ud.ja), ptr(Ud.a),
                 Ud.nnz, ptr(ichol_vals), ptr(d_dev), ptr(d_r), n, ptr(aux[0]), ptr(aux[1]),
                 ptr(aux[2]), keep["name"], m, k)
    return Preconditioner(kind, pd, keep)


def prec_apply(A: DeviceCSR, P: Preconditioner, r: torch.Tensor) -> torch.Tensor:
    w = torch.zeros_like(r)
    _sync()
    L.load().prec_function(ptr(A.ia), ptr(A.ja), ptr(A.a), A.nnz, C.byref(P.pdata), ptr(r),
                           ptr(w))
    L.check()
    return w


# ------------------------------------------------------------------- PCG

@dataclass
class Result:
    x: torch.Tensor
    iterations: int
    flag: int
    history: np.ndarray
    res0: float
    true_res: float
    gpu_ms: float


def pcg(A: DeviceCSR, b: torch.Tensor, P: Preconditioner, tol: float = 1e-10,
        maxit: int = 25000, x0: torch.Tensor | None = None) -> Result:
    """lap_pcg: the device-resident solve with the reference's recurrence."""
    lib = L.load()
    x = torch.zeros(A.n, dtype=torch.float64, device=DEV) if x0 is None else x0.clone()
    hist = np.zeros(maxit + 2)
    info = L.SolveInfo()
    _sync()
    st = lib.lap_pcg(A.n, A.nnz, ptr(A.ia), ptr(A.ja), ptr(A.a), ptr(x), ptr(b), tol,
                     C.byref(P.pdata), maxit, hist.ctypes.data_as(C.POINTER(C.c_double)),
                     C.byref(info))
    if st != 0:
        L.check()
        raise L.LapError(f"lap_pcg returned {st}")
    return Result(x, info.iterations, info.flag, hist[: info.iterations + 1].copy(), info.res0,
                  info.true_res, info.gpu_ms)


def cg_reference_abi(A: DeviceCSR, b: torch.Tensor, P: Preconditioner, tol: float,
                     maxit: int):
    """Calls the exported cg() exactly as the reference's drivers do
    (cg_driver2.c:392-406); it prints the reference's log lines."""
    lib = L.load()
    x = torch.zeros(A.n, dtype=torch.float64, device=DEV)
    hist = np.zeros(maxit + 2)
    it, flag = C.c_int(0), C.c_int(0)
    _sync()
    lib.cg(A.n, float(A.nnz), ptr(A.ia), ptr(A.ja), ptr(A.a), ptr(x), ptr(b), tol,
           C.byref(P.pdata), maxit, C.byref(it), C.byref(flag),
           hist.ctypes.data_as(C.POINTER(C.c_double)))
    L.check()
    return x, it.value, flag.value, hist[: it.value + 1].copy()


def set_option(name: str, value: int) -> int:
    return L.load().lap_set_option(name.encode(), int(value))


def spmv_timed(A: DeviceCSR, x: torch.Tensor, y: torch.Tensor, reps: int,
               fused: bool = False) -> float:
    """Average device ms per SpMV launch (HIP events on the library stream);
    fused=True times the PCG's q = A p kernel with the p.q epilogue."""
    _sync()
    fn = L.load().lap_pcg_spmv_timed if fused else L.load().lap_spmv_timed
    ms = fn(A.n, A.nnz, ptr(A.ia), ptr(A.ja), ptr(A.a), ptr(x), ptr(y), reps)
    L.check()
    return ms


@dataclass
class MixedResult(Result):
    refinements: int = 0


def pcg_mixed(A: DeviceCSR, b: torch.Tensor, P: Preconditioner, tol: float = 1e-10,
              maxit: int = 25000, delta: float = 0.01,
              x0: torch.Tensor | None = None) -> MixedResult:
    """lap_pcg_mixed: FP32 inner PCG in FP64 reliable-update refinement."""
    lib = L.load()
    x = torch.zeros(A.n, dtype=torch.float64, device=DEV) if x0 is None else x0.clone()
    hist = np.zeros(maxit + 2)
    info = L.SolveInfo()
    _sync()
    st = lib.lap_pcg_mixed(A.n, A.nnz, ptr(A.ia), ptr(A.ja), ptr(A.a), ptr(x), ptr(b), tol,
                           C.byref(P.pdata), maxit, delta,
                           hist.ctypes.data_as(C.POINTER(C.c_double)), C.byref(info))
    if st != 0:
        L.check()
        raise L.LapError(f"lap_pcg_mixed returned {st}")
    return MixedResult(x, info.iterations, info.flag, hist[: info.iterations + 1].copy(),
                       info.res0, info.true_res, info.gpu_ms, info.refinements)


# --------------------------------------------------------------- batching

@dataclass
class Batch:
    """Many systems as one block-diagonal device CSR (see lap_pcg_batched)."""
    nsys: int
    row_off: torch.Tensor
    A: DeviceCSR
    d_r: torch.Tensor
    b: torch.Tensor
    sizes: np.ndarray

    @classmethod
    def from_systems(cls, systems, rhs) -> "Batch":
        sizes = np.array([S_.n for S_ in systems], dtype=np.int64)
        row_off = np.zeros(len(systems) + 1, dtype=np.int64)
        np.cumsum(sizes, out=row_off[1:])
        nnz = np.array([S_.nnz for S_ in systems], dtype=np.int64)
        nz_off = np.zeros(len(systems) + 1, dtype=np.int64)
        np.cumsum(nnz, out=nz_off[1:])
        if row_off[-1] >= 2**31 or nz_off[-1] >= 2**31:
            raise L.LapError("batch exceeds int32 CSR")
        ia = np.concatenate([S_.ia[:-1].astype(np.int64) + o for S_, o in zip(systems, nz_off)]
                            + [nz_off[-1:]])
        ja = np.concatenate([S_.ja.astype(np.int64) + o for S_, o in zip(systems, row_off)])
        a = np.concatenate([S_.a for S_ in systems])
        big = CSR(int(row_off[-1]), ia.astype(np.int32), ja.astype(np.int32), a)
        d = diagonal(big)
        d_r = np.where(d != 0.0, 1.0 / np.where(d != 0.0, d, 1.0), 0.0)
        B = cls(len(systems), torch.from_numpy(row_off.astype(np.int32)).to(DEV),
                DeviceCSR.from_csr(big), to_dev(d_r), to_dev(np.concatenate(rhs)), sizes)
        L.load().lap_forget(ptr(B.row_off))
        return B

    def split(self, v: torch.Tensor):
        off = self.row_off.cpu().numpy()
        return [v[off[i]:off[i + 1]] for i in range(self.nsys)]


def batched_tune(B: Batch, trial_iters: int = 20):
    bs, tpr = C.c_int(0), C.c_int(0)
    _sync()
    L.load().lap_pcg_batched_tune(B.nsys, ptr(B.row_off), B.A.n, B.A.nnz, ptr(B.A.ia),
                                  ptr(B.A.ja), ptr(B.A.a), ptr(B.d_r), ptr(B.b), trial_iters,
                                  C.byref(bs), C.byref(tpr))
    L.check()
    return bs.value, tpr.value


def pcg_batched(B: Batch, tol: float = 1e-10, maxit: int = 25000, block: int = 0,
                tpr: int = 0, x0: torch.Tensor | None = None):
    x = torch.zeros(B.A.n, dtype=torch.float64, device=DEV) if x0 is None else x0.clone()
    it = np.zeros(B.nsys, dtype=np.int32)
    fl = np.zeros(B.nsys, dtype=np.int32)
    r0 = np.zeros(B.nsys)
    rs = np.zeros(B.nsys)
    ms = C.c_double(0.0)
    _sync()
    L.load().lap_pcg_batched(B.nsys, ptr(B.row_off), B.A.n, B.A.nnz, ptr(B.A.ia), ptr(B.A.ja),
                             ptr(B.A.a), ptr(B.d_r), ptr(x), ptr(B.b), tol, maxit, block, tpr,
                             it.ctypes.data_as(C.POINTER(C.c_int)),
                             fl.ctypes.data_as(C.POINTER(C.c_int)),
                             r0.ctypes.data_as(C.POINTER(C.c_double)),
                             rs.ctypes.data_as(C.POINTER(C.c_double)), C.byref(ms))
    L.check()
    return x, it, fl, r0, rs, ms.value


def spmm(A: DeviceCSR, X: torch.Tensor, tile: int = 0) -> torch.Tensor:
    """Y = A X for an (n, k) float64 CUDA tensor X (lap_csr_matmat)."""
    X = X.contiguous()
    n, k = X.shape
    Y = torch.empty_like(X)
    _sync()
    L.load().lap_csr_matmat(A.n, A.nnz, ptr(A.ia), ptr(A.ja), ptr(A.a), k, ptr(X), ptr(Y), tile)
    L.check()
    return Y


@dataclass
class BlockResult:
    x: torch.Tensor          # (n, k)
    iterations: np.ndarray   # per column
    flags: np.ndarray
    res0: np.ndarray
    res: np.ndarray
    gpu_ms: float
    history: np.ndarray | None = None  # (iterations.max()+1, k)


def pcg_block(A: DeviceCSR, B: torch.Tensor, d_r: torch.Tensor, tol: float = 1e-10,
              maxit: int = 25000, tile: int = 0, x0: torch.Tensor | None = None,
              history: bool = False) -> BlockResult:
    """k independent line_jacobi PCG solves of A X = B sharing each SpMM
    (lap_pcg_block); B is (n, k) with k a power of two <= 64."""
    B = B.contiguous()
    n, k = B.shape
    X = torch.zeros_like(B) if x0 is None else x0.clone().contiguous()
    it = np.zeros(k, dtype=np.int32)
    fl = np.zeros(k, dtype=np.int32)
    r0 = np.zeros(k)
    rs = np.zeros(k)
    hist = np.zeros((maxit + 2, k)) if history else None
    ms = C.c_double(0.0)
    _sync()
    L.load().lap_pcg_block(A.n, A.nnz, ptr(A.ia), ptr(A.ja), ptr(A.a), ptr(d_r), k, ptr(X),
                           ptr(B), tol, maxit, tile,
                           hist.ctypes.data_as(C.POINTER(C.c_double)) if history else None,
                           it.ctypes.data_as(C.POINTER(C.c_int)),
                           fl.ctypes.data_as(C.POINTER(C.c_int)),
                           r0.ctypes.data_as(C.POINTER(C.c_double)),
                           rs.ctypes.data_as(C.POINTER(C.c_double)), C.byref(ms))
    L.check()
    if history:
        hist = hist[: int(it.max()) + 1]
    return BlockResult(X, it, fl, r0, rs, ms.value, hist)
