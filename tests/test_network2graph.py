"""Treewidth contraction ordering (qaintensor.network2graph), host side, on
the CPU.  Mirrors test/test_treewidth.jl: the known answers (subset,
lacking edges, rem_vertex_fill!, treewidth of complete and local-circuit
graphs, the tree-decomposition examples), with chordality of the
triangulation checked independently by maximum cardinality search
(Tarjan & Yannakakis 1984) and the orders checked as permutations whose
contraction gives ncon's result (oracle).  The reference's planner is Julia
(LightGraphs) and is not run here; no fixture pins its exact orders."""
import warnings

import numpy as np
import pytest

from oracle import qtn_oracle as O
from qaintensor import network2graph as NG
from qaintensor import workloads as Wk
from qaintensor.network import (GeneralTensorNetwork, Summation, Tensor,
                                contract_rep)


def square_tn():
    """test/test_treewidth.jl:41-51: four 2x2 tensors in a ring."""
    A = np.random.default_rng(0).random((2, 2)) + 0j
    tensors = [Tensor(A.copy()) for _ in range(4)]
    contractions = [Summation(p) for p in ([(1, 2), (2, 1)], [(2, 2), (3, 1)],
                                           [(3, 2), (4, 1)], [(4, 2), (1, 1)])]
    return GeneralTensorNetwork(tensors, contractions, [])


def is_chordal(G):
    """Maximum cardinality search; the reverse visit order is a perfect
    elimination ordering iff G is chordal."""
    n = G.nv()
    weight = {v: 0 for v in range(1, n + 1)}
    order = []
    left = set(range(1, n + 1))
    while left:
        v = max(sorted(left), key=lambda u: weight[u])
        order.append(v)
        left.remove(v)
        for u in G.neighbors(v):
            if u in left:
                weight[u] += 1
    peo = order[::-1]
    pos = {v: i for i, v in enumerate(peo)}
    for v in peo:
        later = [u for u in G.neighbors(v) if pos[u] > pos[v]]
        if later:
            p = min(later, key=lambda u: pos[u])
            for u in later:
                if u != p and not G.has_edge(p, u):
                    return False
    return True


def test_subset_and_graph_generators():
    assert NG.subset([1, 2, 3], [1, 2, 3, 4])
    assert not NG.subset([1, 2, 3, 4], [1, 2, 3])
    assert NG.subset((1, 2), [1, 2, 3]) and NG.subset("ba", "abc")
    rng = np.random.default_rng(1)
    G = NG.random_graph(10, 20, rng)
    assert G.nv() == 10 and G.ne() == 20
    with pytest.raises(ValueError, match="Number of edges must be smaller"):
        NG.random_graph(10, 46, rng)
    G = NG.Graph(4)
    for e in ((1, 2), (2, 3), (3, 4)):
        G.add_edge(*e)
    assert G == NG.local_circuit_graph(4, 2)
    G.add_edge(1, 3)
    G.add_edge(2, 4)
    assert G == NG.local_circuit_graph(4, 3)
    G.add_edge(1, 4)
    assert G == NG.local_circuit_graph(4, 4) == NG.complete_graph(4)


def test_rem_vertex_moves_the_last_vertex():
    """LightGraphs rem_vertex!: vertex n takes the freed index."""
    G = NG.Graph(5)
    for e in ((1, 2), (2, 5), (5, 3), (4, 5)):
        G.add_edge(*e)
    G.rem_vertex(2)
    # old 5 is now 2: edges (2,3), (2,4); old 1 lost its only edge
    assert G.nv() == 4 and G.edges() == [(2, 3), (2, 4)]


def test_network_and_line_graph():
    TN0 = square_tn()
    G, _ = NG.network_graph(TN0)
    assert G.nv() == 4 and G.degree() == [2, 2, 2, 2]
    TN = TN0.copy()
    TN.contractions = [Summation([(1, 2), (2, 1), (2, 2), (3, 1)]),
                       Summation([(3, 2), (4, 1), (4, 2), (1, 1)])]
    with pytest.raises(ValueError, match="more than 2 tensors"):
        NG.network_graph(TN)
    LG, _ = NG.line_graph(TN0)
    assert LG.nv() == 4 and LG.degree() == [2, 2, 2, 2]
    rng = np.random.default_rng(2)
    TN = NG.random_tn(10, 20, rng)
    LG, nodeinfo = NG.line_graph(TN)
    expect = set()
    for k, c in enumerate(TN.contractions, start=1):
        i, j = c.idx[0][0], c.idx[1][0]
        expect.add((min(i, j), max(i, j), k))
    assert set(nodeinfo) == expect
    TN = TN0.copy()
    c = TN.contractions.pop()
    TN.openidx = list(c.idx)
    with pytest.warns(UserWarning, match="All open indices are disregarded"):
        NG.line_graph(TN)


def test_lacking_and_rem_vertex_fill():
    G = NG.complete_graph(5)
    assert NG.lacking_for_clique_neigh(G, 1) == (0, [])
    G.rem_edge(2, 3)
    assert NG.lacking_for_clique_neigh(G, 1) == (1, [(2, 3)])
    ordering, labels = [], [1, 2, 3, 4, 5]
    NG.rem_vertex_fill(G, 1, [(2, 3)], ordering, labels)
    assert G == NG.complete_graph(4)
    assert ordering == [1] and labels == [5, 2, 3, 4]


@pytest.mark.parametrize("seed", range(6))
def test_triangulation_is_chordal(seed):
    rng = np.random.default_rng(10 + seed)
    Nn = int(rng.integers(4, 13))
    Ne = int(rng.integers(Nn, Nn * (Nn - 1) // 2 + 1))
    G = NG.random_graph(Nn, Ne, rng)
    order = NG.min_fill_ordering(G)
    assert sorted(order) == list(range(1, Nn + 1))
    H = NG.triangulation(G, order)
    assert all(H.has_edge(u, v) for (u, v) in G.edges())
    assert is_chordal(H)
    # and the check itself rejects a chordless 4-cycle
    C4 = NG.Graph(4)
    for e in ((1, 2), (2, 3), (3, 4), (4, 1)):
        C4.add_edge(*e)
    assert not is_chordal(C4)


def test_treewidth_known_answers():
    for n in (10, 25, 50):
        tw, _, _ = NG.tree_decomposition(NG.complete_graph(n))
        assert tw == n - 1
    for n in range(2, 6):
        tw, _, _ = NG.tree_decomposition(NG.local_circuit_graph(10, n))
        assert tw == n - 1


def test_is_tree_decomposition_examples():
    G = NG.Graph(5)
    for e in ((1, 2), (1, 4), (2, 3), (4, 3), (5, 3), (4, 5)):
        G.add_edge(*e)
    tree = NG.Graph(3)
    tree.add_edge(1, 2)
    tree.add_edge(1, 3)
    assert NG.is_tree_decomposition(G, tree, [[2, 3, 4], [2, 4, 1], [3, 4, 5]])
    for bags, msg in (([[2, 3, 4], [2, 4], [3, 4, 5]],
                       "Union of bags is not equal to union of vertices"),
                      ([[3, 4], [2, 4, 1], [3, 4, 5]],
                       r"Edge \(2, 3\) not found in any bag"),
                      ([[2, 3], [2, 4, 1], [3, 4, 5]],
                       "Subgraph for vertex 4 not connected")):
        with pytest.warns(UserWarning, match=msg):
            assert not NG.is_tree_decomposition(G, tree, bags)


@pytest.mark.parametrize("seed", range(3))
def test_tree_decompositions_are_valid(seed):
    rng = np.random.default_rng(20 + seed)
    Nn = int(rng.integers(20, 51))
    Ne = int(rng.integers(3 * Nn, Nn * (Nn - 1) // 2 + 1))
    G = NG.random_graph(Nn, Ne, rng)
    _, tree, bags = NG.tree_decomposition(G)
    assert NG.is_tree_decomposition(G, tree, bags)
    for (Nn, Ne) in ((10, 10), (10, 20), (10, 50), (30, 60)):
        H, _ = NG.line_graph(NG.random_tn(Nn, Ne, rng))
        _, tree, bags = NG.tree_decomposition(H)
        assert NG.is_tree_decomposition(H, tree, bags)


def test_contraction_order_is_a_permutation():
    TN = square_tn()
    H, edges = NG.line_graph(TN)
    order = NG.contraction_order_graph(H, edges)
    assert sorted(e[2] for e in order) == list(range(1, len(edges) + 1))
    for (i, j, k) in edges:
        c = TN.contractions[k - 1]
        assert {c.idx[0][0], c.idx[1][0]} == {i, j}
    with pytest.raises(ValueError, match="Invalid list of edges"):
        NG.contraction_order_graph(H, edges[:-1])


def _ncon(net):
    return O.ncon([t.data for t in net.tensors], contract_rep(net))


def test_optimize_contraction_order_keeps_the_network():
    TN0 = square_tn()
    TN = TN0.copy()
    NG.optimize_contraction_order(TN)
    assert TN.tensors == TN0.tensors and TN.openidx == TN0.openidx
    assert sorted(map(repr, TN.contractions)) == \
        sorted(map(repr, TN0.contractions))
    assert np.allclose(_ncon(TN), _ncon(TN0))
    TN = TN0.copy()
    c = TN.contractions.pop()
    TN.openidx = list(c.idx)
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        NG.optimize_contraction_order(TN)
    msgs = [str(x.message) for x in w]
    assert any("treewidth algorithm is unlikely" in m for m in msgs)
    assert any("All open indices are disregarded" in m for m in msgs)
    rng = np.random.default_rng(5)
    for (Nn, Ne) in ((10, 10), (10, 20), (10, 50), (20, 60)):
        net0 = NG.random_tn(Nn, Ne, rng)
        net = net0.copy()
        NG.optimize_contraction_order(net)
        assert net.tensors == net0.tensors
        assert sorted(map(repr, net.contractions)) == \
            sorted(map(repr, net0.contractions))


def test_qft4_expectation_network_in_treewidth_order():
    """test/test_treewidth.jl:370-377: the N = 4 QFT expectation network
    contracts to the same scalar after optimize_contraction_order! (the
    decomposed-gate network, whose gate splits run on the GPU, is in the
    GPU network tests)."""
    net0, _ = Wk.qft_expectation_network(4, 7)
    net = net0.copy()
    NG.optimize_contraction_order(net)
    assert sorted(map(repr, net.contractions)) == \
        sorted(map(repr, net0.contractions))
    a, b = _ncon(net), _ncon(net0)
    assert abs(a - b) <= 1e-12 * abs(b)
