"""Node renumbering of the graph solves (lap_amd.graph.rcm_order), host only:
the order is a permutation of all nodes, isolated nodes included, and on a
random geometric graph whose nodes come in random order it brings the
bandwidth of the adjacency pattern far below the random numbering's (the
SpMV gathers then stay within a narrow window of x)."""
import numpy as np

from lap_amd.graph import rcm_order
from lap_amd.problems import random_geometric_graph


def bandwidth(src, dst, where):
    return int(np.abs(where[src] - where[dst]).max())


def test_rcm_is_a_permutation_with_isolated_nodes():
    n = 50
    src = np.array([0, 1, 2, 10, 11], dtype=np.int64)
    dst = np.array([1, 2, 3, 11, 12], dtype=np.int64)
    perm = rcm_order(n, src, dst)
    assert perm.dtype == np.int64 and perm.shape == (n,)
    assert np.array_equal(np.sort(perm), np.arange(n))


def test_rcm_narrows_a_randomly_numbered_geometric_graph():
    g = random_geometric_graph(20000, 3)
    rng = np.random.default_rng(4)
    shuffle = rng.permutation(g.n)  # node i becomes shuffle[i]
    src, dst = shuffle[g.src], shuffle[g.dst]
    natural = bandwidth(src, dst, np.arange(g.n))
    perm = rcm_order(g.n, src, dst)
    where = np.empty(g.n, np.int64)
    where[perm] = np.arange(g.n)  # internal row of each node
    renumbered = bandwidth(src, dst, where)
    assert natural > g.n // 2
    assert renumbered * 20 < natural, (renumbered, natural)
