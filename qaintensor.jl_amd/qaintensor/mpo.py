"""MPO construction and application, decompose! and tensor_circuit!
(Qaintensor.jl src/mpo.jl, src/decompose.jl, src/tensor_circuit.jl).

MPO(m) and decompose! split a gate matrix by sequential untruncated SVDs on
the GPU (qtn_svd, S absorbed into V^H); the permute runs on the GPU too
(qtn_permute).  extend_MPO, apply_MPO and tensor_circuit! only rewire the
network; contract(net) then evaluates it on the GPU.

Gate convention: a 2^M x 2^M matrix acts on wires iwire with gate-matrix
bit k-1 <-> wire iwire[k]; the state is contracted with the matrix's column
legs (src/mpo.jl:208-214).
"""
import numpy as np

from . import ops
from .network import (ErrorException, GeneralTensorNetwork, Summation,
                      Tensor, TensorNetwork, shift_pair, shift_summation)


class CircuitGate:
    """A gate matrix on wires iwire (the parts of Qaintessent.CircuitGate
    that Qaintensor reads: cg.iwire and the gate matrix)."""

    def __init__(self, iwire, matrix):
        if isinstance(iwire, int):
            iwire = (iwire,)
        self.iwire = tuple(int(w) for w in iwire)
        self.matrix = np.asarray(matrix, dtype=np.complex128)
        if self.matrix.shape != (2 ** len(self.iwire),) * 2:
            raise ValueError("gate matrix must be 2^M x 2^M for M wires")


def _split(m, M):
    """The SVD sweep shared by MPO(m) (src/mpo.jl:40-70) and decompose!
    (src/decompose.jl:16-48): tensors (2,2,b1), (b,2,2,b'), ..., (b,2,2)."""
    t = np.asarray(m, dtype=np.complex128).reshape((2,) * (2 * M),
                                                   order="F")
    dims = []
    for i in range(1, M + 1):
        dims += [i, i + M]
    t = ops.permute(t, dims)
    bond = 1
    U, S, SVh = ops.svd(t.reshape((4 * bond, -1), order="F"), absorb=True)
    bond = S.size
    mm = SVh
    out = [Tensor(U.reshape((2, 2, bond), order="F"))]
    for _ in range(2, M):
        U, S, SVh = ops.svd(mm.reshape((bond * 4, -1), order="F"),
                            absorb=True)
        nb = S.size
        mm = SVh
        out.append(Tensor(U.reshape((bond, 2, 2, nb), order="F")))
        bond = nb
    out.append(Tensor(mm.reshape((bond, 2, 2), order="F")))
    return out


class MPO(TensorNetwork):
    """src/mpo.jl:8-110.  MPO(m) builds the operator's tensor train from a
    2^M x 2^M matrix (or a CircuitGate's matrix); direct construction from
    (tensors, contractions, openidx) is refused as in the reference."""

    def __init__(self, m, contractions=None, openidx=None):
        if contractions is not None or openidx is not None:
            raise ErrorException(
                "Direct conversion to MPS form is not support, please "
                "construct MPO from matrix or CircuitGate objects")
        if isinstance(m, CircuitGate):
            m = m.matrix
        m = np.asarray(m)
        if m.ndim != 2 or m.shape[0] != m.shape[1]:
            raise AssertionError("size(m)[1] == size(m)[2]")
        M = int(np.log2(m.shape[0]))
        if 2 ** M != m.shape[0]:
            # Julia's Int(log2(n)) throws InexactError
            raise ValueError("matrix size must be a power of two")
        if M < 1:
            raise ErrorException("Need at least one qubit to act on.")
        self.contractions = []
        self.openidx = []
        if M > 1:
            self.tensors = _split(m, M)
            self.contractions.append(Summation([(1, 3), (2, 1)]))
            for i in range(2, M):
                self.contractions.append(Summation([(i, 4), (i + 1, 1)]))
            # output legs (rows) then input legs (columns), last wire first
            self.openidx = [(M - i + 1, 2) for i in range(1, M)] + [(1, 1)]
            self.openidx += [(M - i + 1, 3) for i in range(1, M)] + [(1, 2)]
        else:
            self.tensors = [Tensor(np.asarray(m, dtype=np.complex128)
                                   .reshape((2, 2), order="F"))]
            self.openidx = [(1, 2), (1, 1)]

    def isapprox(self, other, rtol=np.sqrt(np.finfo(float).eps)):
        """src/mpo.jl:254-256."""
        return (len(self.tensors) == len(other.tensors) and
                all(a.data.shape == b.data.shape and a.isapprox(b, rtol)
                    for a, b in zip(self.tensors, other.tensors)) and
                self.contractions == other.contractions and
                self.openidx == other.openidx)


def extend_MPO(mpo, iwire):
    """src/mpo.jl:122-173: insert identity pipes kron(I_bond, I_2) for the
    wires between iwire's ends (iwire sorted descending).  Accepts an MPO or
    a matrix."""
    iwire = tuple(iwire)
    if len(set(iwire)) != len(iwire):
        raise ErrorException("Repeated wires are not valid.")
    if not isinstance(mpo, MPO):
        if not all(w > 0 for w in iwire):
            raise ErrorException("Wires must be positive integers.")
        if list(iwire) != sorted(iwire, reverse=True):
            raise ErrorException("Wires not sorted")
        mpo = MPO(mpo)
    if list(iwire) != sorted(iwire, reverse=True):
        raise ErrorException("Wires not sorted")
    if not all(w > 0 for w in iwire):
        raise ErrorException("Wires must be positive integers.")
    M = len(iwire)
    iwire = tuple(reversed(iwire))
    N = iwire[-1] - iwire[0] + 1
    if len(mpo.tensors) != M:
        raise AssertionError("length(mpo.tensors) == M")
    if M == N:
        raise ErrorException("MPO is already decomposed in N tensors")
    qwire = list(range(iwire[0], iwire[-1] + 1))
    pipes = sorted(set(qwire) - set(iwire))
    qwire = qwire[::-1]
    d = 2
    for w in pipes:
        ind = qwire.index(w) + 1
        bond = mpo.tensors[ind - 2].size()[-1]
        eye = np.eye(bond * d, dtype=np.complex128)
        pipe = eye.reshape((bond, d, bond, d), order="F")
        pipe = ops.permute(pipe, [1, 2, 4, 3])
        mpo.tensors.insert(ind - 1, Tensor(pipe))
    for i in range(M, N):
        mpo.contractions.append(Summation([(i, 4), (i + 1, 1)]))
        mpo.openidx.insert(0, (i + 1, 2))
        mpo.openidx.insert(i + 1, (i + 1, 3))
    return mpo


def apply_MPO(psi, mpo, iwire=None):
    """src/mpo.jl:184-252: wire the MPO into the network: psi's open leg of
    wire iwire[k] contracts with the MPO's input leg of gate bit k-1 and is
    replaced by the matching output leg.  mpo may be an MPO (iwire sorted),
    a matrix (iwire any order: the matrix is permuted on the GPU) or a
    CircuitGate."""
    if isinstance(mpo, CircuitGate):
        return apply_MPO(psi, mpo.matrix, mpo.iwire)
    iwire = tuple(iwire)
    if not isinstance(mpo, MPO):
        m = np.asarray(mpo, dtype=np.complex128)
        if len(set(iwire)) != len(iwire):
            raise ErrorException("Repeated wires are not valid.")
        if not all(w > 0 for w in iwire):
            raise ErrorException("Wires must be positive integers.")
        M = len(iwire)
        srt = sorted(iwire)
        if srt != list(iwire):
            order = sorted(range(M), key=lambda k: iwire[k])
            perm = [k + 1 for k in order] + [k + 1 + M for k in order]
            t = m.reshape((2,) * (2 * M), order="F")
            m = ops.permute(t, perm).reshape((2 ** M, 2 ** M), order="F")
        return apply_MPO(psi, MPO(m), tuple(srt))
    if len(set(iwire)) != len(iwire):
        raise ErrorException("Repeated wires are not valid.")
    n = len(psi.openidx)
    if not all(0 < w <= n for w in iwire):
        raise ErrorException("Wires must be integers between 1 and n "
                             "(total number of qudits).")
    step = len(psi.tensors)
    N = len(iwire)
    iwire = tuple(reversed(iwire))
    out = GeneralTensorNetwork(
        list(psi.tensors) + list(mpo.tensors),
        list(psi.contractions) + [shift_summation(s, step)
                                  for s in mpo.contractions],
        list(psi.openidx))
    for i, w in enumerate(iwire, start=1):
        out.contractions.append(Summation(
            [psi.openidx[w - 1], shift_pair(mpo.openidx[i + N - 1], step)]))
    for i, q in enumerate(iwire, start=1):
        out.openidx[q - 1] = shift_pair(mpo.openidx[i - 1], step)
    return out


def decompose(cg):
    """decompose!(cg) (src/decompose.jl:6-52): (tensors, link legs, wires)
    of an M > 1 qubit gate split as in MPO(m)."""
    M = len(cg.iwire)
    if M <= 1:
        raise ErrorException("Only decompose Circuit Gates that apply to "
                             "multiple wires")
    t = _split(cg.matrix, M)
    c = [0, 3] + [4] * (M - 2)
    w = list(cg.iwire)
    return t, c, w


def tensor_circuit(psi, cg, is_decompose=False):
    """tensor_circuit!(psi, cg) (src/tensor_circuit.jl:14-79): append the
    gate (dense, or its decompose! chain) and rewire psi's open legs.  cg
    may be a CircuitGate, a list of them, or a list of moments (lists of
    gates).

    The dense path contracts the state with the gate's column legs, so it
    applies the matrix m; the reference's dense path (src/tensor_circuit.jl
    :44-51) contracts the row legs and so applies m^T, which agrees with
    this for every gate with m^T = +-m (the gates its tests use)."""
    if isinstance(cg, (list, tuple)):
        for g in cg:
            tensor_circuit(psi, g, is_decompose=is_decompose)
        return
    M = len(cg.iwire)
    if not max(cg.iwire) <= len(psi.openidx):
        raise AssertionError("Qaintessent.req_wires(cg) <= "
                             "length(psi.openidx)")
    d = 2
    if M > 1 and is_decompose:
        t, c, w = decompose(cg)
        for i, (ti, ci, wi) in enumerate(zip(t, c, w), start=1):
            psi.tensors.append(ti)
            L = len(psi.tensors)
            leg_in = 2 if i == 1 else 3
            psi.contractions.append(Summation([psi.openidx[wi - 1],
                                               (L, leg_in)]))
            if ci > 0:
                psi.contractions.append(Summation([(L - 1, ci), (L, 1)]))
            psi.openidx[wi - 1] = (L, 1) if i == 1 else (L, 2)
        return
    psi.tensors.append(Tensor(cg.matrix.reshape((d,) * (2 * M), order="F")))
    L = len(psi.tensors)
    for i, w in enumerate(cg.iwire, start=1):
        psi.contractions.append(Summation([psi.openidx[w - 1], (L, M + i)]))
        psi.openidx[w - 1] = (L, i)
