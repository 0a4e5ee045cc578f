"""Full-size property checks of the multi-RHS, many-system and gradient
workloads, at the sizes they are benchmarked at (the reference has no
counterpart; the bars are the solver contract's):

* 128^3 7-point Laplacian + 1e-8 I with 64 right-hand sides in one block CG:
  every column converges (flag 0) with an FP64 true residual ||b - A x||
  (computed on the host with scipy, independent of the library's kernels)
  within 1e-10 ||r0||, and two columns agree with single-RHS lap_pcg solves;
* 256 2D grids of side U{32..448} in one persistent batched launch: every
  system's true residual within its tolerance;
* 1M-node random geometric graph, L_w + 1e-4 I: dl/dw . v from the adjoint
  solve agrees with a Richardson-extrapolated central difference of the loss
  along one random direction v to 1e-6 relative, and dl/db . u likewise.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

from lap_amd import solver as S
from lap_amd.problems import diagonal, grid_laplacian, rhs_normal

pytestmark = pytest.mark.gpu


def host_csr(Ah):
    return sp.csr_matrix((Ah.a, Ah.ja, Ah.ia), shape=(Ah.n, Ah.n))


def test_block_cg_128cubed_64_rhs_true_residuals():
    print("[case] full-size block CG 128^3 x 64", flush=True)
    Ah = grid_laplacian([128] * 3, 1e-8)
    k = 64
    Bh = np.random.default_rng(12345).standard_normal((Ah.n, k))
    Bh -= Bh.mean(axis=0)
    A = S.DeviceCSR.from_csr(Ah)
    res = S.pcg_block(A, S.to_dev(Bh), S.to_dev(1.0 / diagonal(Ah)), tol=1e-10)
    assert np.all(res.flags == 0), res.flags
    X = res.x.cpu().numpy()
    R = Bh - host_csr(Ah) @ X
    true = np.linalg.norm(R, axis=0) / np.linalg.norm(Bh, axis=0)
    print(f"[diag] block 128^3 x 64: iterations {res.iterations.min()}..{res.iterations.max()}, "
          f"true rel residual max {true.max():.2e}, {res.gpu_ms:.0f} ms", flush=True)
    assert true.max() <= 1e-10
    P = S.make_preconditioner(A, Ah, "line_jacobi")
    for j in (0, k - 1):
        one = S.pcg(A, S.to_dev(Bh[:, j].copy()), P, tol=1e-10)
        xo = one.x.cpu().numpy()
        dev = np.linalg.norm(X[:, j] - xo) / np.linalg.norm(xo)
        print(f"[diag] column {j}: block {res.iterations[j]} vs single {one.iterations} "
              f"iterations, x rel dev {dev:.2e}", flush=True)
        assert abs(int(res.iterations[j]) - one.iterations) <= 1 and one.flag == 0
        assert dev <= 1e-8


def test_batched_256_grids_true_residuals():
    print("[case] full-size batched 256 grids", flush=True)
    rng = np.random.default_rng(12345)
    sides = rng.integers(32, 449, 256)
    systems = [grid_laplacian([int(s), int(s)], 1e-8) for s in sides]
    rhs = [rhs_normal(Ah.n, 1000 + i) for i, Ah in enumerate(systems)]
    Bt = S.Batch.from_systems(systems, rhs)
    S.batched_tune(Bt, 20)
    tol = 1e-10
    x, it, fl, _, _, ms = S.pcg_batched(Bt, tol=tol)
    assert np.all(fl == 0), np.flatnonzero(fl)
    xh = x.cpu().numpy()
    worst = 0.0
    off = 0
    for Ah, b in zip(systems, rhs):
        xi = xh[off:off + Ah.n]
        off += Ah.n
        worst = max(worst, np.linalg.norm(b - host_csr(Ah) @ xi) / np.linalg.norm(b))
    print(f"[diag] batched 256 grids ({Bt.A.n} rows): iterations {it.min()}..{it.max()}, "
          f"true rel residual max {worst:.2e}, {ms:.0f} ms", flush=True)
    assert worst <= tol


def test_adjoint_1m_rgg_directional_derivatives():
    print("[case] full-size adjoint 1M RGG", flush=True)
    from lap_amd import graph as GR
    from lap_amd.problems import random_geometric_graph
    n = 1_000_000
    g = random_geometric_graph(n, 12345)
    rng = np.random.default_rng(12345)
    G = GR.GraphLaplacian(g.n, g.src, g.dst, 1e-4)
    b = rng.standard_normal(n)
    wstar = g.w * np.exp(0.1 * rng.standard_normal(g.m))
    bd = torch.from_numpy(b).cuda()
    xs = GR.laplacian_solve(G, torch.from_numpy(wstar).cuda(), bd, tol=1e-12).detach()

    w = torch.from_numpy(g.w).cuda().requires_grad_(True)
    bb = bd.clone().requires_grad_(True)
    x = GR.laplacian_solve(G, w, bb, tol=1e-12)
    loss = 0.5 * ((x - xs) ** 2).sum()
    loss.backward()
    gw, gb = w.grad.cpu().numpy(), bb.grad.cpu().numpy()

    def f(wv, bv=b):
        xx = GR.laplacian_solve(G, torch.from_numpy(wv).cuda(), torch.from_numpy(bv).cuda(),
                                tol=1e-12)
        return float(0.5 * ((xx - xs) ** 2).sum())

    v = g.w * rng.standard_normal(g.m)
    h = 1e-3
    D = lambda hh: (f(g.w + hh * v) - f(g.w - hh * v)) / (2 * hh)
    fd = (4 * D(h / 2) - D(h)) / 3  # Richardson-extrapolated central difference
    ew = abs(fd - gw @ v) / abs(gw @ v)
    u = rng.standard_normal(n)
    fdb = (f(g.w, b + 1e-4 * u) - f(g.w, b - 1e-4 * u)) / 2e-4
    eb = abs(fdb - gb @ u) / abs(gb @ u)
    print(f"[diag] adjoint 1M RGG (m={g.m}): forward {G.last_forward.iterations} iterations; "
          f"dl/dw.v {gw @ v:.6e} vs FD {fd:.6e} (rel {ew:.2e}); dl/db.u rel {eb:.2e}",
          flush=True)
    assert ew <= 1e-6 and eb <= 1e-6
