"""Host-side helpers for the weighted-graph tests."""
import numpy as np

from lap_amd.problems import CSR, EdgeGraph


def pattern_maps(g: EdgeGraph):
    """CSR pattern of L(w) (rows sorted, diagonal present) with the maps of
    lap_graph_pattern: epos[2e], epos[2e+1] = positions of (s,d), (d,s);
    dpos[i] = position of (i,i)."""
    rows = np.concatenate([g.src, g.dst, np.arange(g.n)]).astype(np.int64)
    cols = np.concatenate([g.dst, g.src, np.arange(g.n)]).astype(np.int64)
    tag = np.concatenate([2 * np.arange(g.m), 2 * np.arange(g.m) + 1, -np.ones(g.n, np.int64)])
    order = np.lexsort((cols, rows))
    ia = np.zeros(g.n + 1, np.int64)
    np.cumsum(np.bincount(rows, minlength=g.n), out=ia[1:])
    ja = cols[order].astype(np.int32)
    pos = np.empty(order.shape[0], np.int64)
    pos[order] = np.arange(order.shape[0])
    epos = np.zeros(2 * g.m, np.int32)
    t = tag >= 0
    epos[tag[t]] = pos[t]
    dpos = pos[~t].astype(np.int32)
    return ia.astype(np.int32), ja, epos, dpos


def small_graph(n=300, seed=3):
    from lap_amd.problems import random_geometric_graph
    return random_geometric_graph(n, seed)


def loss_and_grad_dense(g: EdgeGraph, w, b, xstar, shift):
    """l = 0.5 ||x - x*||^2 with L(w) x = b, and its gradient by dense FP64
    solves: lambda = L^{-1}(x - x*), dl/dw = -(lam_s - lam_d)(x_s - x_d),
    dl/db = lambda."""
    from lap_amd.problems import graph_dense
    Ld = graph_dense(g, w, shift)
    x = np.linalg.solve(Ld, b)
    lam = np.linalg.solve(Ld, x - xstar)
    gw = -(lam[g.src] - lam[g.dst]) * (x[g.src] - x[g.dst])
    return 0.5 * np.sum((x - xstar) ** 2), gw, lam, x
