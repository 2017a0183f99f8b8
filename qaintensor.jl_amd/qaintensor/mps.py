"""MPS containers, contract_svd and switch! (Qaintensor.jl src/mps.jl,
src/svd.jl, src/switch.jl).

The containers are host-side bookkeeping (tensors, contractions, open legs)
exactly as in the reference; every SVD, permute and contraction they need
runs on the GPU through libqtn (qtn_svd, qtn_contract_svd, qtn_permute).
"""
import numpy as np

from . import ops
from .network import (ErrorException, Summation, Tensor, TensorNetwork,
                      is_power_two)


class MPS(TensorNetwork):
    """src/mps.jl:7-29: a network in matrix-product form.  Construct from
    (tensors, contractions, openidx), or from a statevector psi of length
    2^N as MPS(psi) (src/mps.jl:55-89)."""

    def __init__(self, tensors, contractions=None, openidx=None):
        if contractions is None and openidx is None:
            built = _mps_from_state(tensors)
            tensors, contractions, openidx = built
        tensors = [t if isinstance(t, Tensor) else Tensor(t)
                   for t in tensors]
        contractions = [c if isinstance(c, Summation) else Summation(c)
                        for c in contractions]
        for c in contractions:
            if len(c.idx) != 2:
                raise AssertionError("length(id.idx) == 2")
            (t1, l1), (_, l2) = c.idx
            if l1 != tensors[t1 - 1].ndims():
                raise ErrorException(
                    "Tensor objects first leg must contract with last leg "
                    "of previous Tensor object")
            if l2 != 1:
                raise ErrorException(
                    "Tensor objects last leg must contract with first leg "
                    "of next Tensor object")
        for t in tensors:
            if t.ndims() not in (2, 3):
                raise ErrorException(
                    "Each Tensor object in MPS form can only have 2 or 3 "
                    "legs")
        self.tensors = tensors
        self.contractions = contractions
        self.openidx = [tuple(p) for p in openidx]

    def copy(self):
        """src/mps.jl:203."""
        return MPS(list(self.tensors), list(self.contractions),
                   list(self.openidx))


def check_mps(mps):
    """src/mps.jl:37-48: the structure invariants still hold."""
    for t in mps.tensors:
        if t.ndims() not in (2, 3):
            raise ErrorException(
                "Each Tensor object in MPS form can only have 2 or 3 legs")
    for c in mps.contractions:
        if len(c.idx) != 2:
            raise AssertionError("length(id.idx) == 2")
        (t1, l1), (_, l2) = c.idx
        if l1 != mps.tensors[t1 - 1].ndims():
            raise ErrorException(
                "Tensor objects first leg must contract with last leg of "
                "previous Tensor object")
        if l2 != 1:
            raise ErrorException(
                "Tensor objects last leg must contract with first leg of "
                "next Tensor object")


def _mps_from_state(psi):
    """src/mps.jl:55-89: sequential untruncated SVDs (qtn_svd with S
    absorbed into V^H), wire w <-> site w <-> bit w-1."""
    psi = np.asarray(psi)
    if psi.ndim != 1 or not is_power_two(psi.size):
        raise ErrorException("Input state must have length 2^N")
    M = psi.size.bit_length() - 1
    if M < 2:
        # the reference builds a contraction with tensor index 0 here
        raise IndexError("MPS(psi) needs at least two qubits")
    tensors, contractions, openidx = [], [], []
    m = psi.reshape((2, 2 ** (M - 1)), order="F")
    U, S, SVh = ops.svd(m, absorb=True)
    tensors.append(Tensor(U))
    openidx.append((1, 1))
    lbond = S.size
    m = SVh
    lastbit = 2
    for bit in range(2, M):
        m = m.reshape((lbond * 2, 2 ** (M - bit)), order="F")
        U, S, SVh = ops.svd(m, absorb=True)
        rbond = S.size
        m = SVh
        tensors.append(Tensor(U.reshape((lbond, 2, rbond), order="F")))
        contractions.append(Summation([(bit - 1, lastbit), (bit, 1)]))
        openidx.append((bit, 2))
        lbond = rbond
        lastbit = 3
    tensors.append(Tensor(m))
    contractions.append(Summation([(M - 1, lastbit), (M, 1)]))
    openidx.append((M, 2))
    return tensors, contractions, openidx


def OpenMPS(T, N=None):
    """src/mps.jl:99-121: open boundaries, every tensor 3 legs; OpenMPS(T, N)
    repeats one tensor N times."""
    if N is not None:
        T = [T] * N
    T = [t if isinstance(t, Tensor) else Tensor(t) for t in T]
    l = len(T)
    for t in T:
        if t.ndims() != 3:
            raise ErrorException("Tensors must have 3 legs")
    contractions = [Summation([(i, 3), (i + 1, 1)]) for i in range(1, l)]
    openidx = [(1, 1)] + [(i, 2) for i in range(1, l + 1)] + [(l, 3)]
    return MPS(T, contractions, openidx)


def ClosedMPS(*args):
    """src/mps.jl:130-154: closed boundaries, end tensors have 2 legs.
    ClosedMPS(T) or ClosedMPS(Tfirst, Tmiddle, Tend, N)."""
    if len(args) == 4:
        tf, tm, te, N = args
        T = [tf] + [tm] * (N - 2) + [te]
    else:
        (T,) = args
    T = [t if isinstance(t, Tensor) else Tensor(t) for t in T]
    l = len(T)
    if T[0].ndims() != 2:
        raise ErrorException("First tensor must have 2 legs")
    for t in T[1:l - 1]:
        if t.ndims() != 3:
            raise ErrorException(
                "Tensors must have 3 legs, except the first and last one")
    if T[l - 1].ndims() != 2:
        raise ErrorException("Last tensor must have 2 legs")
    contractions = [Summation([(1, 2), (2, 1)])] + \
        [Summation([(i, 3), (i + 1, 1)]) for i in range(2, l)]
    openidx = [(1, 1)] + [(i, 2) for i in range(2, l + 1)]
    return MPS(T, contractions, openidx)


def PeriodicMPS(T, N=None):
    """src/mps.jl:162-182: the last tensor's right leg joins the first
    tensor's left leg."""
    if N is not None:
        T = [T] * N
    T = [t if isinstance(t, Tensor) else Tensor(t) for t in T]
    l = len(T)
    for t in T:
        if t.ndims() != 3:
            raise AssertionError("Qaintensor.ndims(T[i]) == 3")
    contractions = [Summation([(i, 3), (i + 1, 1)]) for i in range(1, l)] + \
        [Summation([(l, 3), (1, 1)])]
    openidx = [(i, 2) for i in range(1, l + 1)]
    return MPS(T, contractions, openidx)


def contract_svd(T1, T2, indx, er=0.0):
    """src/svd.jl:7-38 on the GPU (qtn_contract_svd)."""
    if not er >= 0:
        raise ErrorException("Error must be positive")
    a = T1.data if isinstance(T1, Tensor) else np.asarray(T1)
    b = T2.data if isinstance(T2, Tensor) else np.asarray(T2)
    i1, i2 = indx
    # Julia's size(A, d) is 1 beyond ndims
    d1 = a.shape[i1 - 1] if i1 <= a.ndim else 1
    d2 = b.shape[i2 - 1] if i2 <= b.ndim else 1
    if d1 != d2:
        raise ErrorException("Dimensions of contraction legs do not match")
    return Tensor(ops.contract_svd(a, b, i1, i2, er))


def contract_svd_mps(tn, er=0.0):
    """src/mps.jl:190-201: contract the chain left to right with
    contract_svd at tail norm er per step."""
    if not er >= 0:
        raise ErrorException("Error must be positive")
    l = len(tn.tensors)
    for s in tn.contractions:
        if s == Summation([(l, 3), (1, 1)]) or s == Summation([(1, 1),
                                                                (l, 3)]):
            raise ErrorException("Function doesn't support periodic "
                                 "boundary conditions for now")
    t = tn.tensors[0]
    for j in range(1, l):
        t = contract_svd(t, tn.tensors[j], (t.ndims(), 1), er=er)
    return t.data


def switch_adjacent(mps, i):
    """switch!(mps, i) (src/switch.jl:18-56): exchange the physical legs of
    sites i and i+1 by contract_svd, a permute and an untruncated SVD, all
    on the GPU."""
    if i < 1 or i >= len(mps.tensors):
        # the reference indexes mps.tensors[i+1] and throws BoundsError
        raise IndexError("Attempt to access %d qubit in MPS of size %d"
                         % (i + 1, i))
    T1 = mps.tensors[i - 1]
    T2 = mps.tensors[i]
    d1, d2 = T1.size(), T2.size()
    T = contract_svd(T1, T2, (T1.ndims(), 1)).data
    if T1.ndims() == 2:
        T = ops.permute(T, [2, 1, 3]).reshape((2, 2 * d2[-1]), order="F")
    elif T2.ndims() == 2:
        T = ops.permute(T, [1, 3, 2]).reshape((2 * d1[0], 2), order="F")
    else:
        T = ops.permute(T, [1, 3, 2, 4]).reshape((2 * d1[0], 2 * d2[-1]),
                                                 order="F")
    U, S, SVh = ops.svd(T, absorb=True)
    k = S.size
    if T1.ndims() == 2:
        U = U.reshape((2, k), order="F")
        V = SVh.reshape((k, 2, d2[-1]), order="F")
    elif T2.ndims() == 2:
        U = U.reshape((d1[0], 2, k), order="F")
        V = SVh
    else:
        U = U.reshape((d1[0], 2, k), order="F")
        V = SVh.reshape((k, 2, d2[-1]), order="F")
    mps.tensors[i - 1] = Tensor(U)
    mps.tensors[i] = Tensor(V)


def switch(mps, i, j=None):
    """switch!(mps, i, j) (src/switch.jl:63-86): swap wires i and j by
    adjacent switches; switch!(mps, [(i, j), ...]) (src/switch.jl:6-11)
    swaps pair by pair; switch!(mps, i) swaps sites i and i+1."""
    if j is None and isinstance(i, (list, tuple)):
        check_mps(mps)
        for (a, b) in i:
            switch(mps, a, b)
        return mps
    if j is None:
        switch_adjacent(mps, i)
        return mps
    check_mps(mps)
    if not (i > 0 and j > 0):
        raise ErrorException("Wire indices `i` and `j` must be positive")
    n = len(mps.tensors)
    if not (i <= n and j <= n):
        raise ErrorException(
            "Indices to swap `i` and `j` must be less than or equal to the "
            "number of open wires in MPS")
    if i == j:
        return mps
    i = n - i + 1
    j = n - j + 1
    lo, hi = (i, j) if i < j else (j, i)
    for k in range(lo, hi):
        switch_adjacent(mps, k)
    for k in range(hi - 2, lo - 1, -1):
        switch_adjacent(mps, k)
    return mps


def permute_mps(mps, order):
    """Base.permute!(mps, order) (src/switch.jl:94-106)."""
    check_mps(mps)
    order = list(order)
    n = len(mps.tensors)
    if len(order) != n:
        raise ErrorException("Given permutation must be same length as "
                             "number of Tensors in MPS")
    if len(set(order)) != len(order):
        raise ErrorException("Permutation order cannot contain repeat "
                             "values")
    if not all(x > 0 for x in order):
        raise ErrorException("Permutation order can only contain positive "
                             "values")
    if max(order) != len(order):
        raise ErrorException("Wire numbers in permutation order cannot "
                             "exceed number of wires in MPS")
    seq = list(range(1, n + 1))
    for i in range(1, n + 1):
        loc = seq.index(order[i - 1]) + 1
        switch(mps, i, loc)
        seq[loc - 1], seq[i - 1] = seq[i - 1], seq[loc - 1]
    return mps
