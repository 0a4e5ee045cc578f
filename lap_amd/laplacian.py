"""Laplacian mode: the reference's graph-Laplacian driver (cg_driver.c) on
the device.

    P = LaplacianProblem.from_adjacency("graph.mtx", prec="GS_std")
    res = P.solve(tol=1e-10)

The driver reads an adjacency file (read_adjacency_file, io_utils.c:80-130),
turns it into the graph Laplacian -- unweighted L = D - A, or the
normalised I - D^-1/2 A D^-1/2 -- (create_L_and_split, io_utils.c:367-439),
takes b_i = (-1)^i, projects b off the Laplacian's null space (the constant
vector, or D^1/2 e when normalised; cg_driver.c:188-230) and solves the
singular system with PCG from x = 0 (cg_driver.c:383-397).  Here every step
runs on the device: the file is parsed by the library's host reader, the
CSR, the Laplacian values and the triangles are built by lap_coo_to_csr,
lap_adjacency_laplacian and lap_csr_split, and the projection uses the
library's BLAS-1 in the driver's operation order.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import io as IO
from . import lib as L
from .solver import DEV, DeviceCSR, Preconditioner, Result, _sync, pcg, ptr


def alternating_rhs(n: int) -> torch.Tensor:
    """b_i = (-1)^i (cg_driver.c:156-160: pow(-1.0, i % 2))."""
    b = torch.ones(n, dtype=torch.float64, device=DEV)
    b[1::2] = -1.0
    return b


def project_rhs(b: torch.Tensor, d: torch.Tensor | None = None, weighted: bool = False):
    """In place: b minus its component along the Laplacian's null vector,
    with the driver's operations (cg_driver.c:188-230).  Unweighted:
    e = ones/sqrt(n), b += (-(e.b)) e.  Weighted: w = sqrt(d) .* ones scaled
    to unit norm (1/sqrt(w.w)), b += (-(b.w)) w."""
    lib = L.load()
    n = int(b.shape[0])
    e = torch.ones(n, dtype=torch.float64, device=DEV)
    _sync()
    if not weighted:
        one_over_norme = 1.0 / math.sqrt(n)  # 1./(double) sqrt(A->n)
        lib.hip_scal(n, one_over_norme, ptr(e))
        be = lib.hip_dot(n, ptr(e), ptr(b))
        lib.hip_axpy(n, -be, ptr(e), ptr(b))
    else:
        if d is None:
            raise L.LapError("the weighted projection needs the diagonal d")
        aux = torch.empty_like(e)
        lib.hip_vector_sqrt(n, ptr(d), ptr(aux))
        lib.hip_vec_vec(n, ptr(aux), ptr(e), ptr(aux))
        de = lib.hip_dot(n, ptr(aux), ptr(aux))
        lib.hip_scal(n, 1.0 / math.sqrt(de), ptr(aux))
        bwe = lib.hip_dot(n, ptr(b), ptr(aux))
        lib.hip_axpy(n, bwe * -1.0, ptr(aux), ptr(b))
    L.check()
    return b


@dataclass
class LaplacianProblem:
    A: DeviceCSR           # the Laplacian (singular)
    deg: torch.Tensor      # int degrees (create_L_and_split's DD)
    d: torch.Tensor        # diagonal of A
    P: Preconditioner
    b: torch.Tensor        # projected right-hand side
    weighted: bool

    @classmethod
    def from_adjacency(cls, path: str, weighted: bool = False, prec: str = "line_jacobi",
                       m: int = 1, k: int = 1) -> "LaplacianProblem":
        """The driver's set-up, on the device.  GS_std gets the triangles
        with the diagonal added (cg_driver.c:89-136), the others the strict
        ones; d_r = 1/d, or ones when weighted (cg_driver.c:253-262)."""
        A, info = IO.load_mm(path, adjacency=True)
        deg = IO.adjacency_laplacian(A, weighted)
        Lt, d, Ut = IO.split(A, augmented=(prec == "GS_std"))
        lib = L.load()
        n = A.n
        d_r = torch.empty(n, dtype=torch.float64, device=DEV)
        _sync()
        if weighted:
            d_r.fill_(1.0)
        else:
            lib.hip_vector_reciprocal(n, ptr(d), ptr(d_r))
        aux = [torch.zeros(n, dtype=torch.float64, device=DEV) for _ in range(3)]
        keep = dict(L=Lt, U=Ut, d=d, d_r=d_r, aux=aux, name=prec.encode(), A=A)
        ichol_vals = None
        if prec == "GS_std":
            lib.initialize_and_analyze_L_and_U_solve(n, Lt.nnz, ptr(Lt.ia), ptr(Lt.ja),
                                                     ptr(Lt.a), Ut.nnz, ptr(Ut.ia), ptr(Ut.ja),
                                                     ptr(Ut.a))
        elif prec == "ichol":
            ichol_vals = A.a.clone()
            _sync()
            lib.lap_forget(ptr(ichol_vals))
            lib.initialize_ichol(n, A.nnz, ptr(A.ia), ptr(A.ja), ptr(ichol_vals))
            keep["ichol_vals"] = ichol_vals
        L.check()
        pd = L.PData(ptr(Lt.ia), ptr(Lt.ja), ptr(Lt.a), Lt.nnz, ptr(Ut.ia), ptr(Ut.ja),
                     ptr(Ut.a), Ut.nnz, ptr(ichol_vals), ptr(d), ptr(d_r), n, ptr(aux[0]),
                     ptr(aux[1]), ptr(aux[2]), keep["name"], m, k)
        b = project_rhs(alternating_rhs(n), d, weighted)
        return cls(A, deg, d, Preconditioner(prec, pd, keep), b, weighted)

    def solve(self, tol: float = 1e-10, maxit: int = 25000) -> Result:
        return pcg(self.A, self.b, self.P, tol=tol, maxit=maxit)
