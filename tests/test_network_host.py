"""Host logic of contract(): the restated planner contract_order
(src/contract.jl:62-235, the breadth-first search of arXiv:1304.6112)
checked against brute force over every pairwise contraction tree under the
same cost model (a build costs the product of the sizes of all legs of its
two operands, plus their own costs; only operands sharing a leg are
joined), and the label bookkeeping it feeds the GPU executor."""
import numpy as np
import pytest

from qaintensor import workloads as Wk
from qaintensor.network import (GeneralTensorNetwork, Summation, Tensor,
                                contract_order, contract_rep, leg_costs)


def brute_force_cost(indexlist, costs):
    """Minimal total cost over all contraction trees (memoised on tensor
    subsets)."""
    n = len(indexlist)
    legs = [frozenset(abs(x) for x in l) for l in indexlist]
    best = {}

    def obj_legs(mask):
        acc = set()
        for i in range(n):
            if mask >> i & 1:
                acc ^= set(legs[i])
        return frozenset(acc)

    def solve(mask):
        if mask in best:
            return best[mask]
        if mask & (mask - 1) == 0:
            best[mask] = 0
            return 0
        res = float("inf")
        sub = (mask - 1) & mask
        while sub:
            other = mask ^ sub
            if sub < other:
                la, lb = obj_legs(sub), obj_legs(other)
                if la & lb:
                    ca, cb = solve(sub), solve(other)
                    c = 1
                    for leg in la | lb:
                        c *= costs[leg]
                    res = min(res, c + ca + cb)
            sub = (sub - 1) & mask
        best[mask] = res
        return res

    return solve((1 << n) - 1)


def random_network(rng, n, extra):
    """n tensors on a random spanning tree plus `extra` more contractions,
    two open legs; leg sizes 2..5."""
    legs = [[] for _ in range(n)]
    pairs = [(i, int(rng.integers(0, i))) for i in range(1, n)]
    for _ in range(extra):
        a, b = rng.choice(n, 2, replace=False)
        pairs.append((int(a), int(b)))
    sizes = []
    for (a, b) in pairs:
        d = int(rng.integers(2, 6))
        legs[a].append(len(sizes))
        legs[b].append(len(sizes))
        sizes.append(d)
    for t in rng.choice(n, 2, replace=False):
        legs[t].append(len(sizes))
        sizes.append(int(rng.integers(2, 4)))
    tensors, where = [], {}
    for t in range(n):
        shape = [sizes[l] for l in legs[t]]
        tensors.append(Tensor(np.zeros(shape, dtype=np.complex128)))
        for k, l in enumerate(legs[t]):
            where.setdefault(l, []).append((t + 1, k + 1))
    contr = [Summation(where[l]) for l in range(len(sizes))
             if len(where[l]) == 2]
    openidx = [where[l][0] for l in range(len(sizes)) if len(where[l]) == 1]
    return GeneralTensorNetwork(tensors, contr, openidx)


@pytest.mark.parametrize("seed", range(6))
def test_contract_order_is_optimal_for_its_cost_model(seed):
    rng = np.random.default_rng(seed)
    net = random_network(rng, 5 + seed % 2, 2)
    labels = contract_rep(net)
    costs = leg_costs(net)
    seq, cost = contract_order(net, costs, labels)
    assert cost == brute_force_cost(labels, costs)
    assert sorted(seq) == list(range(1, len(net.contractions) + 1))


def test_contract_order_qft4_expectation_network():
    """The reference's own planner case (test/test_treewidth.jl:370-377):
    the N = 4 QFT expectation network, every contraction once."""
    net, _ = Wk.qft_expectation_network(4, 3)
    seq, cost = contract_order(net, leg_costs(net), contract_rep(net))
    assert sorted(seq) == list(range(1, len(net.contractions) + 1))
    assert cost == 500


def test_contract_order_rejects_disconnected_network():
    from qaintensor.network import ErrorException
    a, b = Tensor(np.zeros((2, 3))), Tensor(np.zeros((3, 2)))
    c = Tensor(np.zeros((2,)))
    net = GeneralTensorNetwork([a, b, c], [Summation([(1, 2), (2, 1)])],
                               [(1, 1), (2, 2), (3, 1)])
    with pytest.raises(ErrorException, match="not connected"):
        contract_order(net, leg_costs(net), contract_rep(net))
