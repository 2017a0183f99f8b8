"""Treewidth contraction ordering of a tensor network (Qaintensor.jl
src/network2graph.jl): the network's line graph, a min-fill elimination
ordering, the tree decomposition it induces, and the contraction order that
eliminates the decomposition's leaves first (Markov & Shi, "Simulating
quantum computation by contracting tensor networks", SIAM J. Comput. 38
(2008), Theorem 4.6).  optimize_contraction_order(net) permutes
net.contractions so that contract(net) (ncon's default order) follows it.

Host-side planning, as in the reference.  Graph mirrors the LightGraphs
SimpleGraph operations the reference uses, with their vertex numbering
(1-based, sorted adjacency lists, and rem_vertex moving the last vertex
into the freed index), so the orderings and bags come out as the
reference's do.
"""
import bisect
import warnings

import numpy as np


class Graph:
    """Undirected simple graph on vertices 1..n (LightGraphs SimpleGraph)."""

    def __init__(self, n=0):
        self.adj = [[] for _ in range(n + 1)]  # adj[0] unused

    def nv(self):
        return len(self.adj) - 1

    def ne(self):
        loops = sum(1 for v in range(1, self.nv() + 1) if v in self.adj[v])
        return (sum(len(a) for a in self.adj[1:]) - loops) // 2 + loops

    def copy(self):
        g = Graph(0)
        g.adj = [list(a) for a in self.adj]
        return g

    def add_vertex(self):
        self.adj.append([])

    def has_edge(self, u, v):
        a = self.adj[u]
        i = bisect.bisect_left(a, v)
        return i < len(a) and a[i] == v

    def add_edge(self, u, v):
        if self.has_edge(u, v):
            return False
        bisect.insort(self.adj[u], v)
        if u != v:
            bisect.insort(self.adj[v], u)
        return True

    def rem_edge(self, u, v):
        if not self.has_edge(u, v):
            return False
        self.adj[u].remove(v)
        if u != v:
            self.adj[v].remove(u)
        return True

    def neighbors(self, v):
        return self.adj[v]

    def degree(self, v=None):
        if v is None:
            return [len(a) for a in self.adj[1:]]
        return len(self.adj[v])

    def rem_vertex(self, v):
        """LightGraphs rem_vertex!: drop v's edges, move the last vertex n
        into index v, drop index n."""
        n = self.nv()
        for s in list(self.adj[v]):
            self.rem_edge(s, v)
        neigs = list(self.adj[n])
        for s in neigs:
            self.rem_edge(s, n)
        self_loop_n = False
        if v != n:
            for s in neigs:
                if s != n:
                    self.add_edge(s, v)
                else:
                    self_loop_n = True
        if self_loop_n:
            self.add_edge(v, v)
        self.adj.pop()

    def edges(self):
        return [(u, v) for u in range(1, self.nv() + 1) for v in self.adj[u]
                if u <= v]

    def __eq__(self, other):
        return isinstance(other, Graph) and self.adj == other.adj


def complete_graph(n):
    g = Graph(n)
    for i in range(1, n + 1):
        for j in range(i + 1, n + 1):
            g.add_edge(i, j)
    return g


def _intersect(a, b):
    """Julia intersect: unique elements of a that are in b, in a's order."""
    bs, seen, out = set(b), set(), []
    for x in a:
        if x in bs and x not in seen:
            out.append(x)
            seen.add(x)
    return out


def _union(a, b):
    seen, out = set(), []
    for x in list(a) + list(b):
        if x not in seen:
            out.append(x)
            seen.add(x)
    return out


def _setdiff(a, b):
    bs, seen, out = set(b), set(), []
    for x in a:
        if x not in bs and x not in seen:
            out.append(x)
            seen.add(x)
    return out


def subset(a, b):
    """The reference's A ⊂ B: every element of A is in B."""
    bs = set(b)
    return all(x in bs for x in a)


def random_graph(Nn, Ne, rng):
    """src/network2graph.jl:18-30 (rng: a numpy Generator)."""
    if not Ne <= Nn * (Nn - 1) // 2:
        raise ValueError("Number of edges must be smaller or equal than "
                         "N(N-1)/2, with N the number of vertices")
    G = Graph(Nn)
    possible = [(i, j) for i in range(1, Nn + 1) for j in range(1, i)]
    for k in rng.choice(len(possible), Ne, replace=False):
        G.add_edge(*possible[int(k)])
    return G


def local_circuit_graph(N, k):
    """src/network2graph.jl:37-45: N qubits, gates on k-neighbours."""
    G = Graph(N)
    for i in range(1, N):
        for j in range(1, min(k - 1, N - i) + 1):
            G.add_edge(i, i + j)
    return G


def network_graph(net):
    """src/network2graph.jl:53-73: (graph over the tensors, {(i, j):
    indices of the contractions between tensors i <= j})."""
    G = Graph(len(net.tensors))
    edge_idx = {}
    for k, s in enumerate(net.contractions, start=1):
        if len(s.idx) != 2:
            raise ValueError("Contractions of more than 2 tensors not "
                             "supported")
        i, j = s.idx[0][0], s.idx[1][0]
        if i > j:
            i, j = j, i
        G.add_edge(i, j)
        edge_idx.setdefault((i, j), []).append(k)
    return G, edge_idx


def line_graph(net):
    """src/network2graph.jl:121-152: one vertex per contraction between two
    different tensors, nodeinfo[v-1] = (i, j, k); vertices adjacent when
    their contractions share a tensor.  Open legs are disregarded."""
    if len(net.openidx) != 0:
        warnings.warn("All open indices are disregarded")
    G, edge_idx = network_graph(net)
    LG = Graph(0)
    nodeinfo = []
    for i in range(1, G.nv() + 1):
        for j in G.neighbors(i):
            if j > i:
                for e in edge_idx[(i, j)]:
                    LG.add_vertex()
                    nodeinfo.append((i, j, e))
    by_tensor = {}
    for v, (i, j, _) in enumerate(nodeinfo, start=1):
        by_tensor.setdefault(i, []).append(v)
        by_tensor.setdefault(j, []).append(v)
    for vs in by_tensor.values():
        for a in range(len(vs)):
            for b in range(a + 1, len(vs)):
                LG.add_edge(vs[a], vs[b])
    return LG, nodeinfo


def lacking_for_clique_neigh(G, i):
    """src/network2graph.jl:187-196: the edges missing for i's neighbourhood
    to be a clique, (count, [(i1, i2), ...])."""
    neigh = G.neighbors(i)
    lacking = []
    for j, i2 in enumerate(neigh):
        for i1 in neigh[:j]:
            if not G.has_edge(i1, i2):
                lacking.append((i1, i2))
    return len(lacking), lacking


def rem_vertex_fill(G, i, lacking, ordering, vertex_label):
    """src/network2graph.jl:210-220: fill i's neighbourhood, eliminate i,
    record its original label."""
    for (a, b) in lacking:
        G.add_edge(a, b)
    ordering.append(vertex_label[i - 1])
    G.rem_vertex(i)
    v = vertex_label.pop()
    if i <= G.nv():
        vertex_label[i - 1] = v


def min_fill_ordering(G):
    """src/network2graph.jl:224-271: eliminate isolated and degree-one
    vertices first, otherwise a simplicial vertex (lowest degree first) or
    the vertex whose elimination adds the fewest edges (Bodlaender,
    "Discovering treewidth", 2005)."""
    H = G.copy()
    ordering = []
    vertex_label = list(range(1, H.nv() + 1))
    while H.nv() > 0:
        success = False
        for i in range(H.nv(), 0, -1):
            if i <= H.nv() and H.degree(i) == 0:
                rem_vertex_fill(H, i, [], ordering, vertex_label)
                success = True
        for i in range(H.nv(), 0, -1):
            if i <= H.nv() and H.degree(i) == 1:
                rem_vertex_fill(H, i, [], ordering, vertex_label)
                success = True
        if not success:
            degrees = H.degree()
            J = sorted(range(1, H.nv() + 1), key=lambda v: degrees[v - 1])
            found_clique = False
            v, best_n, best_lacking = 0, float("inf"), []
            for j in J:
                n_lacking, lacking = lacking_for_clique_neigh(H, j)
                if n_lacking == 0:
                    rem_vertex_fill(H, j, lacking, ordering, vertex_label)
                    found_clique = True
                    break
                elif n_lacking < best_n:
                    v, best_n, best_lacking = j, n_lacking, lacking
            if not found_clique:
                rem_vertex_fill(H, v, best_lacking, ordering, vertex_label)
    return ordering


def triangulation(G, ordering):
    """src/network2graph.jl:279-291: chordal completion along ordering."""
    H = G.copy()
    pos = {v: i for i, v in enumerate(ordering)}
    for i, v in enumerate(ordering):
        high = [u for u in H.neighbors(v) if pos[u] > i]
        for j, i1 in enumerate(high):
            for i2 in high[:j]:
                H.add_edge(i1, i2)
    return H


def tree_decomposition(G):
    """src/network2graph.jl:300-339: (width, tree, bags) from the min-fill
    ordering's triangulation."""
    ordering = min_fill_ordering(G)
    H = triangulation(G, ordering)
    n = H.nv()
    pos = {v: i for i, v in enumerate(ordering)}
    up_neighs = [[u for u in H.neighbors(ordering[i]) if pos[u] > i]
                 for i in range(n)]
    idx = next(i for i in range(n) if len(up_neighs[i]) == n - 1 - i)
    first_bag = _union(up_neighs[idx], [ordering[idx]])
    decomp = Graph(1)
    bags = [first_bag]
    tw = len(first_bag) - 1
    for i in range(idx - 1, -1, -1):
        neigh = up_neighs[i]
        old = 0
        for j, bag in enumerate(bags, start=1):
            if subset(neigh, bag):
                old = j
                break
        if old == 0:
            old = 1
        decomp.add_vertex()
        new_bag = neigh + [ordering[i]]
        bags.append(new_bag)
        tw = max(tw, len(new_bag) - 1)
        decomp.add_edge(old, decomp.nv())
    return tw, decomp, bags


def is_tree_decomposition(G, tree, bags):
    """src/network2graph.jl:346-381: union, edge and subtree properties
    (warnings name the failed one)."""
    allv = sorted(set(v for b in bags for v in b))
    if allv != list(range(1, G.nv() + 1)):
        warnings.warn("Union of bags is not equal to union of vertices")
        return False
    for (u, v) in G.edges():
        if not any(subset((u, v), b) for b in bags):
            warnings.warn("Edge (%d, %d) not found in any bag" % (u, v))
            return False
    holders = [[] for _ in range(G.nv() + 1)]
    for i, b in enumerate(bags, start=1):
        for v in b:
            holders[v].append(i)
    for v in range(1, G.nv() + 1):
        s = holders[v]
        if not s:
            continue
        inside, seen, stack = set(s), {s[0]}, [s[0]]
        while stack:
            x = stack.pop()
            for y in tree.neighbors(x):
                if y in inside and y not in seen:
                    seen.add(y)
                    stack.append(y)
        if seen != inside:
            warnings.warn("Subgraph for vertex %d not connected" % v)
            return False
    return True


def contraction_order_graph(H, edges):
    """contraction_order(H, edges) (src/network2graph.jl:391-424): leaf bags
    of the tree decomposition eliminated smallest first; each vertex is
    emitted when its last bag goes."""
    if len(edges) != H.nv():
        raise ValueError("Invalid list of edges for `H`; the length of "
                         "`edges` must equal the number of vertices of `H`")
    _, tree, bags = tree_decomposition(H)
    bags = [list(b) for b in bags]
    order = []
    degrees = tree.degree()
    while max(degrees) > 0:
        leaves = [i for i, d in enumerate(degrees, start=1) if d == 1]
        leaf = min(leaves, key=lambda i: len(bags[i - 1]))
        leaf_bag = bags[leaf - 1]
        neigh_bag = bags[tree.neighbors(leaf)[0] - 1]
        for i in _setdiff(leaf_bag, neigh_bag):
            order.append(edges[i - 1])
        tree.rem_vertex(leaf)
        moved = bags.pop()
        if leaf <= tree.nv():
            bags[leaf - 1] = moved
        degrees = tree.degree()
    assert len(bags) == 1
    for i in bags[0]:
        order.append(edges[i - 1])
    return order


def contraction_order(net):
    """contraction_order(net) (src/network2graph.jl:429-445): traces first,
    then the line graph's order, as (i, j, k) triples."""
    _, edge_idx = network_graph(net)
    H, edges = line_graph(net)
    con = contraction_order_graph(H, edges)
    auto = [(i, i, k) for i in range(1, len(net.tensors) + 1)
            for k in edge_idx.get((i, i), [])]
    return auto + con


def optimize_contraction_order(net):
    """optimize_contraction_order! (src/network2graph.jl:473-479): permute
    net.contractions into the treewidth order, in place."""
    if len(net.openidx) != 0:
        warnings.warn("For TensorNetworks with open indices the treewidth "
                      "algorithm is unlikely to optimize performance")
    order = contraction_order(net)
    net.contractions = [net.contractions[k - 1] for (_, _, k) in order]


def random_tn(Nn, Ne, rng):
    """The random network of test/test_treewidth.jl:20-33: Ne legs of size
    2 between random tensor pairs, no open legs."""
    from .network import GeneralTensorNetwork, Summation, Tensor
    nlegs = [0] * (Nn + 1)
    contractions = []
    for _ in range(Ne):
        n1 = int(rng.integers(1, Nn))
        n2 = int(rng.integers(n1 + 1, Nn + 1))
        nlegs[n1] += 1
        nlegs[n2] += 1
        contractions.append(Summation([(n1, nlegs[n1]), (n2, nlegs[n2])]))
    tensors = [Tensor(rng.random((2,) * n) if n > 0 else rng.random(1))
               for n in nlegs[1:]]
    return GeneralTensorNetwork(tensors, contractions, [])
