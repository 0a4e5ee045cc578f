"""Weighted graph-Laplacian solves on the device, differentiable in the edge
weights and the right-hand side.

    G = GraphLaplacian(n, src, dst, shift=1e-4)      # pattern, built once
    x = laplacian_solve(G, w, b)                     # torch.autograd.Function
    loss = 0.5 * ((x - x_star) ** 2).sum(); loss.backward()   # w.grad, b.grad

Forward: L(w) is assembled on the device from w (lap_graph_values) and
solved by the device PCG (lap_pcg) with line_jacobi or IC0.  Backward: one
adjoint PCG L(w) lambda = dl/dx with the forward's preconditioner and the
fused edge kernel dl/dw_e = -(lambda_s - lambda_d)(x_s - x_d); dl/db =
lambda (lap_pcg_adjoint).  No reference counterpart (the reference solves
one unweighted Laplacian and has no derivative).

Node order: by default the nodes are renumbered internally by reverse
Cuthill-McKee, so that rows sharing columns sit together and the SpMV's
gathers of x stay in cache (a random geometric graph in the caller's random
order otherwise gathers from all of x: 2.8x slower per iteration at 1M
nodes).  Vectors cross the interface in the caller's order; edges keep
theirs.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import lib as L
from .solver import DEV, DeviceCSR, Preconditioner, _sync, ptr


def _idx(v) -> torch.Tensor:
    if isinstance(v, torch.Tensor):
        return v.to(device=DEV, dtype=torch.int32).contiguous()
    return torch.from_numpy(np.ascontiguousarray(v, dtype=np.int32)).to(DEV)


def _host(v) -> np.ndarray:
    if isinstance(v, torch.Tensor):
        return v.detach().cpu().numpy().astype(np.int64)
    return np.asarray(v, dtype=np.int64)


def rcm_order(n: int, src: np.ndarray, dst: np.ndarray) -> np.ndarray:
    """perm[k] = the node placed at internal row k: reverse Cuthill-McKee of
    the undirected graph (scipy.sparse.csgraph; 0.3 s at 1M nodes, 5M
    edges)."""
    import scipy.sparse as sp
    from scipy.sparse.csgraph import reverse_cuthill_mckee
    M = sp.coo_matrix((np.ones(len(src), np.int8), (src, dst)), shape=(n, n)).tocsr()
    return reverse_cuthill_mckee((M + M.T).tocsr(), symmetric_mode=True).astype(np.int64)


class GraphLaplacian:
    """CSR pattern of L(w) = D(w) - W(w) + shift*I for fixed edges.

    order "rcm" (default) numbers the nodes internally by reverse
    Cuthill-McKee; "natural" keeps the caller's numbering.  The pattern
    (ia, ja, epos, dpos), assemble() and preconditioner() are in the internal
    order: internal row k is the caller's node perm[k] (perm is None for the
    natural order).  laplacian_solve takes b and returns x (and dl/db) in
    the caller's order; the edge order (w, dl/dw) is never changed.  src and
    dst are the caller's; isrc and idst the same edges in internal numbers."""

    def __init__(self, n: int, src, dst, shift: float = 0.0, order: str = "rcm"):
        if order not in ("rcm", "natural"):
            raise L.LapError(f"order must be 'rcm' or 'natural', not {order!r}")
        self.n = int(n)
        self.order = order
        self.src, self.dst = _idx(src), _idx(dst)
        self.m = int(self.src.shape[0])
        if self.dst.shape[0] != self.m:
            raise L.LapError("src and dst differ in length")
        self.shift = float(shift)
        self.perm = self.iperm = None
        self.isrc, self.idst = self.src, self.dst
        if order == "rcm" and self.m:
            hs, hd = _host(src), _host(dst)
            if hs.min() >= 0 and hd.min() >= 0 and max(hs.max(), hd.max()) < self.n:
                perm = rcm_order(self.n, hs, hd)
                iperm = np.empty(self.n, np.int64)
                iperm[perm] = np.arange(self.n)
                self.perm = torch.from_numpy(perm).to(DEV)
                self.iperm = torch.from_numpy(iperm).to(DEV)
                self.isrc, self.idst = _idx(iperm[hs]), _idx(iperm[hd])
            # (out-of-range nodes: left to lap_graph_pattern to reject)
        nnz = self.n + 2 * self.m
        i32 = dict(dtype=torch.int32, device=DEV)
        self.ia = torch.empty(self.n + 1, **i32)
        self.ja = torch.empty(nnz, **i32)
        self.epos = torch.empty(max(2 * self.m, 1), **i32)
        self.dpos = torch.empty(self.n, **i32)
        _sync()
        st = L.load().lap_graph_pattern(self.n, self.m, ptr(self.isrc), ptr(self.idst),
                                        ptr(self.ia), ptr(self.ja), ptr(self.epos),
                                        ptr(self.dpos))
        if st != 0:
            L.check()
            raise L.LapError(f"lap_graph_pattern returned {st}")
        self.last_forward = None
        self.last_adjoint = None

    @property
    def nnz(self) -> int:
        return self.n + 2 * self.m

    def to_internal(self, v: torch.Tensor) -> torch.Tensor:
        """A node vector in the caller's order -> internal order."""
        return v if self.perm is None else v[self.perm]

    def to_caller(self, v: torch.Tensor) -> torch.Tensor:
        """A node vector in internal order -> the caller's order."""
        return v if self.iperm is None else v[self.iperm]

    def assemble(self, w: torch.Tensor) -> tuple[DeviceCSR, torch.Tensor]:
        """(L(w) on the device, 1/diag(L(w)))."""
        w = w.detach().to(device=DEV, dtype=torch.float64).contiguous()
        if w.shape != (self.m,):
            raise L.LapError(f"w has shape {tuple(w.shape)}, expected ({self.m},)")
        a = torch.empty(self.nnz, dtype=torch.float64, device=DEV)
        d_r = torch.empty(self.n, dtype=torch.float64, device=DEV)
        _sync()
        st = L.load().lap_graph_values(self.n, self.m, ptr(self.ia), ptr(self.ja), ptr(self.epos),
                                       ptr(self.dpos), ptr(w), self.shift, ptr(a), ptr(d_r))
        if st != 0:
            L.check()
            raise L.LapError(f"lap_graph_values returned {st}")
        return DeviceCSR(self.n, self.ia, self.ja, a), d_r

    def preconditioner(self, A: DeviceCSR, d_r: torch.Tensor, kind: str = "line_jacobi",
                       ) -> Preconditioner:
        """line_jacobi (d_r) or ichol (IC0 of A, factored on the device)."""
        lib = L.load()
        keep = dict(d_r=d_r, name=kind.encode(), A=A)
        ichol_vals = None
        if kind == "ichol":
            ichol_vals = A.a.clone()
            _sync()
            lib.lap_forget(ptr(ichol_vals))
            lib.initialize_ichol(A.n, A.nnz, ptr(A.ia), ptr(A.ja), ptr(ichol_vals))
            L.check()
            keep["ichol_vals"] = ichol_vals
        elif kind not in ("line_jacobi", "none"):
            raise L.LapError(f"graph solves support line_jacobi, ichol and none, not {kind}")
        pd = L.PData(None, None, None, 0, None, None, None, 0, ptr(ichol_vals), None, ptr(d_r),
                     A.n, None, None, None, keep["name"], 1, 1)
        return Preconditioner(kind, pd, keep)


@dataclass
class SolveStats:
    iterations: int
    flag: int
    res0: float
    true_res: float
    gpu_ms: float


class _LaplacianSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, w, b, G: GraphLaplacian, prec: str, tol: float, maxit: int):
        from .solver import pcg
        A, d_r = G.assemble(w)
        P = G.preconditioner(A, d_r, prec)
        bb = G.to_internal(b.detach().to(device=DEV, dtype=torch.float64)).contiguous()
        res = pcg(A, bb, P, tol=tol, maxit=maxit)
        G.last_forward = SolveStats(res.iterations, res.flag, res.res0, res.true_res, res.gpu_ms)
        if res.flag != 0:
            raise L.LapError(f"forward solve stopped with flag {res.flag}")
        ctx.G, ctx.A, ctx.P, ctx.tol, ctx.maxit = G, A, P, tol, maxit
        ctx.save_for_backward(res.x)  # internal order
        return G.to_caller(res.x)

    @staticmethod
    def backward(ctx, gx):
        (x,) = ctx.saved_tensors
        G, A, P = ctx.G, ctx.A, ctx.P
        gx = G.to_internal(gx.detach().to(dtype=torch.float64)).contiguous()
        lam = torch.empty(G.n, dtype=torch.float64, device=DEV)
        gw = torch.empty(G.m, dtype=torch.float64, device=DEV) if ctx.needs_input_grad[0] else None
        info = L.SolveInfo()
        _sync()
        st = L.load().lap_pcg_adjoint(A.n, A.nnz, ptr(A.ia), ptr(A.ja), ptr(A.a),
                                      C.byref(P.pdata), ptr(gx), ptr(lam), ctx.tol, ctx.maxit,
                                      G.m, ptr(G.isrc), ptr(G.idst), ptr(x), ptr(gw), None,
                                      C.byref(info))
        G.last_adjoint = SolveStats(info.iterations, info.flag, info.res0, info.true_res,
                                    info.gpu_ms)
        if st != 0:
            L.check()
            raise L.LapError(f"lap_pcg_adjoint returned {st}")
        return gw, (G.to_caller(lam) if ctx.needs_input_grad[1] else None), None, None, None, None


def laplacian_solve(G: GraphLaplacian, w: torch.Tensor, b: torch.Tensor,
                    prec: str = "line_jacobi", tol: float = 1e-10,
                    maxit: int = 25000) -> torch.Tensor:
    """x = L(w)^{-1} b, differentiable in w (m,) and b (n,)."""
    return _LaplacianSolve.apply(w, b, G, prec, tol, maxit)


def edge_grad(G: GraphLaplacian, lam: torch.Tensor, x: torch.Tensor) -> torch.Tensor:
    """-(lam_s - lam_d)(x_s - x_d) per edge (lap_edge_grad)."""
    g = torch.empty(G.m, dtype=torch.float64, device=DEV)
    _sync()
    st = L.load().lap_edge_grad(G.n, G.m, ptr(G.src), ptr(G.dst), ptr(lam), ptr(x), ptr(g))
    if st != 0:
        L.check()
        raise L.LapError(f"lap_edge_grad returned {st}")
    return g
