"""Differentiable weighted-graph Laplacian solve (lap_graph_pattern,
lap_graph_values, lap_edge_grad, lap_pcg_adjoint, lap_amd.graph).

Bars:
  * pattern maps address the right entries; assembled values and 1/diag are
    bit-identical to the oracle's (orc_graph_values);
  * the edge kernel is bit-identical to orc_edge_grad;
  * on a 2000-node random geometric graph (L_w + 1e-4 I, b ~ N(0,1), x* the
    solution for weights w*exp(0.1 xi), l = 0.5||x - x*||^2) the autograd
    gradients dl/dw and dl/db agree with dense FP64 solves to 1e-6 relative,
    and with central finite differences of l (evaluated with this solver)
    along random directions and for the five largest edge gradients to 1e-6
    relative, for line_jacobi and IC0 preconditioners;
  * torch.autograd.gradcheck passes on a 40-node graph;
  * the default node order (reverse Cuthill-McKee inside the solver) is the
    natural pattern renumbered, with a smaller bandwidth, and solves to the
    same x; the gradient tests run in that order.
"""
import numpy as np
import pytest
import torch

from graphs import loss_and_grad_dense, pattern_maps, small_graph
from lap_amd import graph as GR
from lap_amd import lib as L
from lap_amd.problems import graph_dense, random_geometric_graph
from lap_amd.solver import DEV

pytestmark = pytest.mark.gpu

SHIFT = 1e-4


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def case():
    g = random_geometric_graph(2000, 5)
    rng = np.random.default_rng(11)
    b = rng.standard_normal(g.n)
    wstar = g.w * np.exp(0.1 * rng.standard_normal(g.m))
    xstar = np.linalg.solve(graph_dense(g, wstar, SHIFT), b)
    return g, b, xstar


def test_pattern_and_values_bitwise(orc):
    g = small_graph(500, 2)
    G = GR.GraphLaplacian(g.n, g.src, g.dst, SHIFT, order="natural")
    ia, ja, epos, dpos = pattern_maps(g)
    assert np.array_equal(G.ia.cpu().numpy(), ia) and np.array_equal(G.ja.cpu().numpy(), ja)
    assert np.array_equal(G.epos.cpu().numpy(), epos)
    assert np.array_equal(G.dpos.cpu().numpy(), dpos)
    A, d_r = G.assemble(torch.from_numpy(g.w).to(DEV))
    a = orc.graph_values(g.n, ia, ja, epos, dpos, g.w, SHIFT)
    assert np.array_equal(A.a.cpu().numpy(), a)
    assert np.array_equal(d_r.cpu().numpy(), 1.0 / a[dpos])


def test_rcm_pattern_is_the_permuted_natural_one(orc):
    """order "rcm": internal row k is node perm[k]; the pattern and the edge
    values equal the natural ones with rows and columns renumbered (the
    degrees to rounding: summed in another order), and a solve returns the
    caller's order (same x as the natural order to 1e-10)."""
    g = random_geometric_graph(3000, 4)
    Gn = GR.GraphLaplacian(g.n, g.src, g.dst, SHIFT, order="natural")
    Gr = GR.GraphLaplacian(g.n, g.src, g.dst, SHIFT)
    perm = Gr.perm.cpu().numpy()
    assert np.array_equal(np.sort(perm), np.arange(g.n))
    w = torch.from_numpy(g.w).to(DEV)
    An, _ = Gn.assemble(w)
    Ar, _ = Gr.assemble(w)
    import scipy.sparse as sp
    Mn = sp.csr_matrix((An.a.cpu().numpy(), An.ja.cpu().numpy(), An.ia.cpu().numpy()))
    Mr = sp.csr_matrix((Ar.a.cpu().numpy(), Ar.ja.cpu().numpy(), Ar.ia.cpu().numpy()))
    P = Mn[perm][:, perm].tocsr()
    P.sort_indices()
    assert np.array_equal(P.indptr, Mr.indptr) and np.array_equal(P.indices, Mr.indices)
    off = P.indices != np.repeat(np.arange(g.n), np.diff(P.indptr))
    assert np.array_equal(P.data[off], Mr.data[off])  # the edge weights, moved
    # the diagonals are the degrees summed in the rows' (renumbered) column
    # order: equal to rounding
    np.testing.assert_allclose(P.data[~off], Mr.data[~off], rtol=1e-14, atol=0)
    bw = Mr.indices - np.repeat(np.arange(g.n), np.diff(Mr.indptr))
    bn = Mn.indices - np.repeat(np.arange(g.n), np.diff(Mn.indptr))
    print(f"[diag] rcm bandwidth {np.abs(bw).max()} vs natural {np.abs(bn).max()}", flush=True)
    assert np.abs(bw).max() < np.abs(bn).max()
    b = torch.from_numpy(np.random.default_rng(2).standard_normal(g.n)).to(DEV)
    xn = GR.laplacian_solve(Gn, w, b, tol=1e-12).cpu().numpy()
    xr = GR.laplacian_solve(Gr, w, b, tol=1e-12).cpu().numpy()
    assert rel(xr, xn) <= 1e-10


def test_edge_kernel_bitwise(orc):
    g = small_graph(500, 2)
    G = GR.GraphLaplacian(g.n, g.src, g.dst, SHIFT)
    rng = np.random.default_rng(3)
    lam, x = rng.standard_normal(g.n), rng.standard_normal(g.n)
    got = GR.edge_grad(G, torch.from_numpy(lam).to(DEV), torch.from_numpy(x).to(DEV))
    assert np.array_equal(got.cpu().numpy(), orc.edge_grad(g.src, g.dst, lam, x))


def test_pattern_rejects_bad_edges():
    with pytest.raises(L.LapError):
        GR.GraphLaplacian(4, [0, 1], [1, 1])  # self loop
    with pytest.raises(L.LapError):
        GR.GraphLaplacian(4, [0, 1], [1, 0])  # duplicate edge
    with pytest.raises(L.LapError):
        GR.GraphLaplacian(4, [0], [4])        # out of range


def _grads(g, b, xstar, prec, tol=1e-12):
    G = GR.GraphLaplacian(g.n, g.src, g.dst, SHIFT)
    w = torch.from_numpy(g.w.copy()).to(DEV).requires_grad_(True)
    bt = torch.from_numpy(b.copy()).to(DEV).requires_grad_(True)
    x = GR.laplacian_solve(G, w, bt, prec=prec, tol=tol)
    loss = 0.5 * ((x - torch.from_numpy(xstar).to(DEV)) ** 2).sum()
    loss.backward()
    return G, float(loss), w.grad.cpu().numpy(), bt.grad.cpu().numpy()


@pytest.mark.parametrize("prec", ["line_jacobi", "ichol"])
def test_gradient_matches_dense_solve(case, prec):
    g, b, xstar = case
    print(f"[case] adjoint dense prec={prec}", flush=True)
    G, loss, gw, gb = _grads(g, b, xstar, prec)
    loss_d, gw_d, lam_d, _ = loss_and_grad_dense(g, g.w, b, xstar, SHIFT)
    ew, eb = rel(gw, gw_d), rel(gb, lam_d)
    print(f"[diag] adjoint {prec}: forward it {G.last_forward.iterations}, adjoint it "
          f"{G.last_adjoint.iterations}, loss rel {abs(loss - loss_d) / loss_d:.2e}, "
          f"dl/dw rel {ew:.2e}, dl/db rel {eb:.2e}", flush=True)
    assert ew <= 1e-6 and eb <= 1e-6
    assert abs(loss - loss_d) <= 1e-8 * loss_d


def test_gradient_matches_central_differences(case):
    g, b, xstar = case
    print("[case] adjoint finite differences", flush=True)
    G, _, gw, gb = _grads(g, b, xstar, "line_jacobi")
    xs = torch.from_numpy(xstar).to(DEV)

    def f(w, bb=b):
        x = GR.laplacian_solve(G, torch.from_numpy(w).to(DEV), torch.from_numpy(bb).to(DEV),
                               tol=1e-12)
        return float(0.5 * ((x - xs) ** 2).sum())

    rng = np.random.default_rng(5)
    v = g.w * rng.standard_normal(g.m)
    h = 1e-3
    D = lambda hh: (f(g.w + hh * v) - f(g.w - hh * v)) / (2 * hh)
    fd = (4 * D(h / 2) - D(h)) / 3  # Richardson-extrapolated central difference
    worst = abs(fd - gw @ v) / abs(gw @ v)
    u = rng.standard_normal(g.n)
    fdb = (f(g.w, b + 1e-4 * u) - f(g.w, b - 1e-4 * u)) / 2e-4
    worst_b = abs(fdb - gb @ u) / abs(gb @ u)
    for e in np.argsort(-np.abs(gw))[:5]:
        he = 1e-5 * g.w[e]
        ep = np.zeros(g.m)
        ep[e] = he
        fde = (f(g.w + ep) - f(g.w - ep)) / (2 * he)
        worst = max(worst, abs(fde - gw[e]) / abs(gw[e]))
    print(f"[diag] adjoint vs central FD: dl/dw worst rel {worst:.2e}, dl/db rel {worst_b:.2e}",
          flush=True)
    assert worst <= 1e-6 and worst_b <= 1e-6


def test_gradcheck_small_graph():
    g = random_geometric_graph(40, 8, radius=0.35)
    G = GR.GraphLaplacian(g.n, g.src, g.dst, 1e-2)
    w = torch.from_numpy(g.w.copy()).to(DEV).requires_grad_(True)
    b = torch.from_numpy(np.random.default_rng(1).standard_normal(g.n)).to(DEV)
    b.requires_grad_(True)
    fn = lambda ww, bb: GR.laplacian_solve(G, ww, bb, tol=1e-13, maxit=2000)
    assert torch.autograd.gradcheck(fn, (w, b), eps=1e-6, atol=1e-6, rtol=1e-4)


def test_repeated_forwards_keep_their_own_operands(case):
    """Two forwards before one backward: each backward uses its own L(w)."""
    g, b, xstar = case
    G = GR.GraphLaplacian(g.n, g.src, g.dst, SHIFT)
    w1 = torch.from_numpy(g.w.copy()).to(DEV).requires_grad_(True)
    w2 = torch.from_numpy(g.w * 1.5).to(DEV).requires_grad_(True)
    bt = torch.from_numpy(b).to(DEV)
    xs = torch.from_numpy(xstar).to(DEV)
    x1 = GR.laplacian_solve(G, w1, bt, tol=1e-12)
    x2 = GR.laplacian_solve(G, w2, bt, tol=1e-12)
    (0.5 * ((x1 - xs) ** 2).sum()).backward()
    (0.5 * ((x2 - xs) ** 2).sum()).backward()
    _, g1, _, _ = loss_and_grad_dense(g, g.w, b, xstar, SHIFT)
    _, g2, _, _ = loss_and_grad_dense(g, g.w * 1.5, b, xstar, SHIFT)
    assert rel(w1.grad.cpu().numpy(), g1) <= 1e-6
    assert rel(w2.grad.cpu().numpy(), g2) <= 1e-6
