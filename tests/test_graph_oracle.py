"""CPU checks of the weighted-graph oracle functions (no reference
counterpart) against the mathematics: the assembled L(w) equals the dense
sum_e w_e (u_s - u_d)(u_s - u_d)^T + shift*I and has row sums shift; the
edge-gradient formula with dense FP64 solves equals central finite
differences of l(w, b) = 0.5 ||L(w)^{-1} b - x*||^2."""
import numpy as np
import pytest

from graphs import loss_and_grad_dense, pattern_maps, small_graph
from lap_amd.problems import graph_dense

SHIFT = 1e-3


@pytest.fixture(scope="module")
def g():
    return small_graph()


def test_pattern_maps_address_edges(g):
    ia, ja, epos, dpos = pattern_maps(g)
    rows = np.repeat(np.arange(g.n), np.diff(ia))
    assert np.all(np.diff(ja)[np.diff(rows) == 0] > 0)  # sorted, no duplicates
    assert np.array_equal(rows[epos[0::2]], g.src) and np.array_equal(ja[epos[0::2]], g.dst)
    assert np.array_equal(rows[epos[1::2]], g.dst) and np.array_equal(ja[epos[1::2]], g.src)
    assert np.array_equal(rows[dpos], np.arange(g.n)) and np.array_equal(ja[dpos], np.arange(g.n))


def test_graph_values_match_dense_definition(orc, g):
    ia, ja, epos, dpos = pattern_maps(g)
    a = orc.graph_values(g.n, ia, ja, epos, dpos, g.w, SHIFT)
    dense = np.zeros((g.n, g.n))
    rows = np.repeat(np.arange(g.n), np.diff(ia))
    dense[rows, ja] = a
    ref = graph_dense(g, g.w, SHIFT)
    np.testing.assert_allclose(dense, ref, rtol=1e-14, atol=1e-15)
    assert np.array_equal(dense, dense.T)
    np.testing.assert_allclose(dense.sum(axis=1), SHIFT, atol=1e-13)


def test_edge_gradient_matches_central_differences(orc, g):
    rng = np.random.default_rng(7)
    b = rng.standard_normal(g.n)
    xstar = np.linalg.solve(graph_dense(g, g.w * np.exp(0.1 * rng.standard_normal(g.m)), SHIFT), b)
    loss, _, lam, x = loss_and_grad_dense(g, g.w, b, xstar, SHIFT)
    gw = orc.edge_grad(g.src, g.dst, lam, x)

    def f(w, bb=b):
        return 0.5 * np.sum((np.linalg.solve(graph_dense(g, w, SHIFT), bb) - xstar) ** 2)

    v = g.w * rng.standard_normal(g.m)  # relative perturbation of the weights
    h = 1e-3
    D = lambda hh: (f(g.w + hh * v) - f(g.w - hh * v)) / (2 * hh)
    fd = (4 * D(h / 2) - D(h)) / 3  # Richardson: O(h^4)
    assert abs(fd - gw @ v) <= 1e-7 * abs(gw @ v)
    u = rng.standard_normal(g.n)
    Db = lambda hh: (f(g.w, b + hh * u) - f(g.w, b - hh * u)) / (2 * hh)
    assert abs(Db(1e-4) - lam @ u) <= 1e-7 * abs(lam @ u)  # quadratic in b: exact up to rounding
    # a planted sign error in the formula is caught
    assert abs(fd + gw @ v) > 1e-3 * abs(gw @ v)
