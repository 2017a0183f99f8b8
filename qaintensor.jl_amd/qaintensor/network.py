"""Tensor-network data model and contract() (Qaintensor.jl src/tensor.jl,
src/tensor_network.jl, src/contract.jl:39-60, 242-264).

Arrays keep the reference's (Julia) index order; a whole contraction runs
on the GPU in one qtn_contract_network call (every pairwise step one
gather-staged complex MFMA GEMM, intermediates kept on the device).  Leg
and tensor numbers are 1-based as in the reference.
"""
import numpy as np

from . import ops


class ErrorException(RuntimeError):
    """The reference's `error(msg)` (Julia ErrorException); str() is the
    reference's message."""


class Tensor:
    """src/tensor.jl:7-22: a wrapper of one complex array."""

    def __init__(self, data):
        self.data = np.asarray(data)

    def ndims(self):
        return self.data.ndim

    def size(self, d=None):
        if d is None:
            return tuple(self.data.shape)
        # Julia's size(A, d) is 1 beyond ndims
        return self.data.shape[d - 1] if d <= self.data.ndim else 1

    def reshape(self, *dims):
        if len(dims) == 1 and isinstance(dims[0], (tuple, list)):
            dims = tuple(dims[0])
        return Tensor(self.data.reshape(dims, order="F"))

    def transpose(self):
        return Tensor(self.data.T)

    def isapprox(self, other, rtol=np.sqrt(np.finfo(float).eps), atol=0.0):
        """Elementwise approximate equality (src/tensor.jl:20-22)."""
        return bool(np.all(np.isclose(self.data, other.data, rtol=rtol,
                                      atol=atol)))

    def __repr__(self):
        return "Tensor(shape=%s)" % (self.data.shape,)


class Summation:
    """src/tensor_network.jl:7-14: one contraction as (tensor, leg)
    pairs."""

    def __init__(self, idx):
        self.idx = [tuple(p) for p in idx]

    def __eq__(self, other):
        return isinstance(other, Summation) and self.idx == other.idx

    def __repr__(self):
        return "Summation(%s)" % (self.idx,)


class TensorNetwork:
    """Abstract network (src/tensor_network.jl:19)."""


class GeneralTensorNetwork(TensorNetwork):
    """src/tensor_network.jl:26-35: tensors, contractions, ordered open
    legs."""

    def __init__(self, tensors, contractions, openidx):
        self.tensors = list(tensors)
        self.contractions = list(contractions)
        self.openidx = [tuple(p) for p in openidx]

    def copy(self):
        return GeneralTensorNetwork(list(self.tensors),
                                    list(self.contractions),
                                    list(self.openidx))


def shift_summation(s, step):
    """src/helper.jl:6-8."""
    return Summation([(t + step, leg) for (t, leg) in s.idx[:2]])


def shift_pair(p, step):
    """src/helper.jl:15-17."""
    return (p[0] + step, p[1])


def is_power_two(i):
    """src/helper.jl:24-27."""
    return i != 0 and (i & (i - 1)) == 0


def contract_rep(net):
    """src/contract.jl:39-60: ncon label lists; contraction i -> label i,
    open leg i -> label -i - #contractions."""
    indexlist = [[0] * t.ndims() for t in net.tensors]
    for i, s in enumerate(net.contractions, start=1):
        for (t, leg) in s.idx:
            if not 1 <= leg <= net.tensors[t - 1].ndims():
                raise AssertionError("leg out of range")
            indexlist[t - 1][leg - 1] = i
    J = len(net.contractions)
    for i, (t, leg) in enumerate(net.openidx, start=1):
        if indexlist[t - 1][leg - 1] != 0:
            raise AssertionError("open leg also contracted")
        indexlist[t - 1][leg - 1] = -i - J
    for lst in indexlist:
        if any(l == 0 for l in lst):
            raise AssertionError("leg neither contracted nor open")
    return indexlist


def leg_costs(net):
    """The leg sizes of contract_rep(net, optimize) (src/contract.jl:8-32):
    {label: size} for contraction i -> label i and open leg i -> label
    i + #contractions."""
    costs = {}
    for i, c in enumerate(net.contractions, start=1):
        for (t, leg) in c.idx:
            costs[i] = net.tensors[t - 1].size(leg)
    J = len(net.contractions)
    for i, (t, leg) in enumerate(net.openidx, start=1):
        costs[i + J] = net.tensors[t - 1].size(leg)
    return costs


class _Obj:
    """ContractionObject (src/contract.jl:161-169): the legs and tensors of a
    partial contraction, the leg sequence that builds it and its cost."""
    __slots__ = ("legs", "tensors", "seq", "cost")

    def __init__(self, legs, tensors, seq, cost):
        self.legs, self.tensors, self.seq, self.cost = legs, tensors, seq, cost


def _build_cost(free, common, costs, mu_old, mu, isnew, c1, c2):
    """getBuildCost (src/contract.jl:62-86): cost of building a tensor from
    two (the product of the sizes of all their legs, plus the two costs);
    refused above the cap mu, or, for objects already built at an earlier
    cap, at or below mu_old."""
    new = 1
    for leg in free | common:
        new *= costs[leg]
    new += c1 + c2
    if new > mu:
        return new, False
    if not isnew and new <= mu_old:
        return float("inf"), False
    return new, True


def _check(S, a, b, costs, newflags, mus):
    """check_contraction! (src/contract.jl:88-158): every disjoint pair of an
    a-tensor and a b-tensor object that share a leg builds an (a+b)-tensor
    object if its cost is within the cap; the cheapest one per tensor set
    is kept."""
    mu_old, mu_cap, mu_next = mus
    Sab = S[a + b]
    for i, Ta in enumerate(S[a]):
        for j, Tb in enumerate(S[b]):
            if Ta.tensors & Tb.tensors:
                continue
            common = Ta.legs & Tb.legs
            if not common:
                continue
            free = Ta.legs ^ Tb.legs
            cost, ok = _build_cost(free, common, costs, mu_old, mu_cap,
                                   newflags[a][i] or newflags[b][j],
                                   Ta.cost, Tb.cost)
            if not ok:
                mu_next = min(mu_next, cost)
                continue
            tens = Ta.tensors | Tb.tensors
            seq = Ta.seq + Tb.seq + sorted(common)
            for k, o in enumerate(Sab):
                if o.tensors == tens:
                    if o.cost > cost:
                        Sab[k] = _Obj(free, tens, seq, cost)
                        newflags[a + b][k] = True
                    break
            else:
                Sab.append(_Obj(free, tens, seq, cost))
                newflags[a + b].append(True)
    return mu_old, mu_cap, mu_next


def contract_order(net, costs, indexlist):
    """contract_order (src/contract.jl:177-235): the breadth-first search of
    Pfeifer, Haegeman and Verstraete (arXiv:1304.6112) over objects built
    from 1, 2, ..., n tensors under a rising cost cap.  Returns (sequence,
    cost): the contracted labels in the order the cheapest full build
    contracts them.  (The reference loops forever on a network that is not
    connected; this raises instead.)"""
    n = len(net.tensors)
    S = [None, [_Obj(frozenset(abs(x) for x in lst), frozenset([i]), [], 0)
                for i, lst in enumerate(indexlist)]]
    newflags = [None, [True] * n]
    for _ in range(2, n + 1):
        S.append([])
        newflags.append([])
    mu_old, mu_cap, mu_next = 0, 1, float("inf")
    while not S[n]:
        for c in range(2, n + 1):
            for d in range(1, c // 2 + 1):
                mu_old, mu_cap, mu_next = _check(S, d, c - d, costs, newflags,
                                                 (mu_old, mu_cap, mu_next))
        if mu_next == float("inf") and not S[n]:
            raise ErrorException("contract_order: the network is not "
                                 "connected")
        mu_old, mu_cap, mu_next = mu_cap, mu_next, float("inf")
        for fl in newflags[1:]:
            fl[:] = [False] * len(fl)
    best = S[n][0]
    return list(best.seq), best.cost


def contract(net, optimize=False):
    """src/contract.jl:242-264.  With one tensor only the open legs are
    permuted.  Otherwise the whole network contracts on the GPU in one call
    (qtn_contract_network, intermediates on the device): by increasing label
    (ncon's default order), or with optimize=True in the order the
    reference's planner contract_order returns.  (The reference relabels the
    legs by their position in that sequence and then passes the old labels
    as ncon's order, src/contract.jl:250-257, so it does not contract in the
    order it planned; here the planned sequence is the order.)"""
    if len(net.tensors) == 1:
        perm = [leg for (_, leg) in net.openidx]
        return ops.permute(net.tensors[0].data, perm)
    labels = contract_rep(net)
    tensors = [t.data for t in net.tensors]
    if optimize:
        order, _ = contract_order(net, leg_costs(net), labels)
    else:
        order = sorted({x for l in labels for x in l if x > 0})
    return ops.contract_network(tensors, labels, order)
