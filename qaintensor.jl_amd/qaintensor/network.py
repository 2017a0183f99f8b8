"""Tensor-network data model and contract() (Qaintensor.jl src/tensor.jl,
src/tensor_network.jl, src/contract.jl:39-60, 242-264).

Arrays keep the reference's (Julia) index order; every pairwise step of a
contraction runs on the GPU through qtn_contract_pair (permute + complex
MFMA GEMM).  Leg and tensor numbers are 1-based as in the reference.
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


def _ncon(tensors, labels):
    """ncon of the labelled tensors on the GPU.  Schedule: positive labels in
    increasing order, each step contracting every label the two holders
    share; disconnected pieces are joined by outer products; a label held
    twice by one tensor (a trace) is contracted against an identity.  The
    result is ordered by negative label, -1 (or the largest) first."""
    items = [(np.asarray(t, dtype=np.complex128), list(l))
             for t, l in zip(tensors, labels)]
    for lab in sorted({x for _, l in items for x in l if x > 0}):
        holders = [i for i, (_, l) in enumerate(items) if lab in l]
        if not holders:
            continue  # already contracted together with an earlier label
        if len(holders) == 1:
            (A, la) = items[holders[0]]
            p, q = [k for k, x in enumerate(la) if x == lab]
            eye = np.eye(A.shape[p], dtype=np.complex128)
            C = ops.contract_pair(A, eye, [p + 1, q + 1], [1, 2])
            items[holders[0]] = (C, [x for x in la if x != lab])
            continue
        i, j = holders
        (A, la), (B, lb) = items[i], items[j]
        shared = [x for x in la if x > 0 and x in lb]
        legsA = [la.index(x) + 1 for x in shared]
        legsB = [lb.index(x) + 1 for x in shared]
        C = ops.contract_pair(A, B, legsA, legsB)
        lc = [x for x in la if x not in shared] + \
             [x for x in lb if x not in shared]
        items = [it for k, it in enumerate(items) if k not in (i, j)]
        items.append((C, lc))
    while len(items) > 1:
        (A, la), (B, lb) = items[0], items[1]
        C = ops.contract_pair(A, B, [], [])
        items = [(C, la + lb)] + items[2:]
    C, lc = items[0]
    order = sorted(range(len(lc)), key=lambda k: -lc[k])
    if order != list(range(len(lc))):
        C = ops.permute(C, [k + 1 for k in order])
    return C


def contract(net, optimize=False):
    """src/contract.jl:242-264.  With one tensor only the open legs are
    permuted.  optimize=True orders the pairwise steps greedily by the size
    of each step's result (the reference's opt_einsum-style planner,
    src/contract.jl:8-235, is host-side planning outside this library's
    scope; any order gives the same tensor)."""
    if len(net.tensors) == 1:
        perm = [leg for (_, leg) in net.openidx]
        return ops.permute(net.tensors[0].data, perm)
    labels = contract_rep(net)
    tensors = [t.data for t in net.tensors]
    if not optimize:
        return _ncon(tensors, labels)
    return _ncon_greedy(tensors, labels)


def _ncon_greedy(tensors, labels):
    items = [(np.asarray(t, dtype=np.complex128), list(l))
             for t, l in zip(tensors, labels)]
    dim = {}
    for t, l in items:
        for x, d in zip(l, t.shape):
            dim[x] = d
    while True:
        best = None
        for i in range(len(items)):
            for j in range(i + 1, len(items)):
                li, lj = items[i][1], items[j][1]
                shared = [x for x in li if x > 0 and x in lj]
                if not shared:
                    continue
                out = [x for x in li + lj if x not in shared]
                cost = float(np.prod([dim[x] for x in out])) if out else 1.0
                if best is None or cost < best[0]:
                    best = (cost, i, j, shared)
        if best is None:
            break
        _, i, j, shared = best
        (A, la), (B, lb) = items[i], items[j]
        C = ops.contract_pair(A, B, [la.index(x) + 1 for x in shared],
                              [lb.index(x) + 1 for x in shared])
        lc = [x for x in la if x not in shared] + \
             [x for x in lb if x not in shared]
        items = [it for k, it in enumerate(items) if k not in (i, j)]
        items.append((C, lc))
    while len(items) > 1:
        (A, la), (B, lb) = items[0], items[1]
        items = [(ops.contract_pair(A, B, [], []), la + lb)] + items[2:]
    C, lc = items[0]
    order = sorted(range(len(lc)), key=lambda k: -lc[k])
    if order != list(range(len(lc))):
        C = ops.permute(C, [k + 1 for k in order])
    return C
