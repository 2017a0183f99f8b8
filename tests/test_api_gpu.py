"""The host mirror of the reference API (Tensor / GeneralTensorNetwork /
contract, contract_svd, MPS / OpenMPS / ClosedMPS, MPS(psi), switch!,
permute!, MPO(m), extend_MPO, apply_MPO, decompose!, tensor_circuit!)
driven through libqtn on the GPU, against the oracle and closed forms.

The cases follow the reference's own tests (test/test_svd.jl,
test/test_mps.jl, test/test_mpo.jl, test/test_tensor_circuit.jl) with
seeded inputs.  Tolerance: complex FP64 throughout, relative 1e-10 on every
compared tensor (the reference's tests use isapprox, rtol ~1.5e-8).
"""
import itertools

import numpy as np
import pytest

from oracle import qtn_oracle as O
from qaintensor import (MPO, MPS, CircuitGate, ClosedMPS, ErrorException,
                        GeneralTensorNetwork, OpenMPS, PeriodicMPS,
                        Summation, Tensor, apply_MPO, contract, contract_svd,
                        contract_svd_mps, decompose, extend_MPO,
                        permute_mps, switch, tensor_circuit)
from qaintensor import ops

pytestmark = pytest.mark.gpu
RTOL = 1e-10

X = np.array([[0, 1], [1, 0]], dtype=complex)
Y = np.array([[0, -1j], [1j, 0]], dtype=complex)
Z = np.diag([1, -1]).astype(complex)
H = np.array([[1, 1], [1, -1]], dtype=complex) / np.sqrt(2)
SWAP = np.eye(4, dtype=complex)[[0, 2, 1, 3]]


def close(a, b, rtol=RTOL):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, (a.shape, b.shape)
    return np.linalg.norm(a - b) <= rtol * max(np.linalg.norm(b), 1e-300)


def crand(rng, *shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))


def controlled(u, nctrl=1):
    """Matrix of u on bit 0 controlled by bits 1..nctrl (all set)."""
    d = 2 ** (nctrl + 1)
    m = np.eye(d, dtype=complex)
    m[d - 2:, d - 2:] = u
    # bit 0 is the fastest index: rows d-2, d-1 have every control bit set
    return m


def haar(rng, d):
    q, r = np.linalg.qr(crand(rng, d, d))
    return q * (np.diag(r) / np.abs(np.diag(r)))


def mps_network(rng):
    """The three-site network of test/test_tensor_circuit.jl:25-37."""
    t = [Tensor(crand(rng, 2, 6)), Tensor(crand(rng, 2, 6, 7)),
         Tensor(crand(rng, 2, 7))]
    return GeneralTensorNetwork(
        t, [Summation([(1, 2), (2, 2)]), Summation([(2, 3), (3, 2)])],
        [(1, 1), (2, 1), (3, 1)])


def net_oracle(net):
    return O.contract_network([t.data for t in net.tensors],
                              [s.idx for s in net.contractions],
                              net.openidx)


# -- contract(net) ---------------------------------------------------------
def test_contract_single_tensor_is_permute(gpu_ctx):
    """test/test_tensor_circuit.jl:13-23."""
    rng = np.random.default_rng(3)
    a = crand(rng, 2, 6)
    net = GeneralTensorNetwork([Tensor(a)], [], [(1, 1), (1, 2)])
    assert np.array_equal(contract(net), a)
    net = GeneralTensorNetwork([Tensor(a)], [], [(1, 2), (1, 1)])
    assert np.array_equal(contract(net), a.T)


@pytest.mark.parametrize("optimize", [False, True])
def test_contract_three_site_network(gpu_ctx, optimize):
    net = mps_network(np.random.default_rng(5))
    assert close(contract(net, optimize), net_oracle(net))


def test_contract_trace_and_outer_product(gpu_ctx):
    rng = np.random.default_rng(8)
    a = crand(rng, 3, 4, 3)
    b = crand(rng, 5)
    # leg 1 and leg 3 of a are traced; b is disconnected (outer product)
    net = GeneralTensorNetwork([Tensor(a), Tensor(b)],
                               [Summation([(1, 1), (1, 3)])],
                               [(2, 1), (1, 2)])
    ref = np.einsum("iji,k->kj", a, b)
    assert close(contract(net), ref)


def test_contract_random_networks_vs_einsum(gpu_ctx):
    """Random 4-tensor loop networks with two open legs."""
    rng = np.random.default_rng(11)
    for _ in range(5):
        d = rng.integers(1, 6, size=6)
        a = crand(rng, d[0], d[1], d[4])
        b = crand(rng, d[1], d[2])
        c = crand(rng, d[2], d[3], d[5])
        e = crand(rng, d[3], d[0])
        net = GeneralTensorNetwork(
            [Tensor(a), Tensor(b), Tensor(c), Tensor(e)],
            [Summation([(1, 2), (2, 1)]), Summation([(2, 2), (3, 1)]),
             Summation([(3, 2), (4, 1)]), Summation([(4, 2), (1, 1)])],
            [(3, 3), (1, 3)])
        ref = np.einsum("abx,bc,cdy,da->yx", a, b, c, e)
        assert close(contract(net), ref)
        assert close(contract(net, True), ref)


# -- contract_svd ----------------------------------------------------------
def test_contract_svd_exact_matches_product(gpu_ctx):
    """test/test_svd.jl:7-25."""
    rng = np.random.default_rng(13)
    t1, t2 = crand(rng, 4, 4), crand(rng, 4, 4)
    assert close(contract_svd(Tensor(t1), Tensor(t2), (2, 1)).data, t1 @ t2,
                 1e-12)
    t1 = Tensor(crand(rng, 2, 3, 4, 6))
    t2 = Tensor(crand(rng, 1, 5, 2, 4))
    tn1 = GeneralTensorNetwork([t1, t2], [Summation([(1, 1), (2, 3)])],
                               [(1, 2), (1, 3), (1, 4), (2, 1), (2, 2),
                                (2, 4)])
    tn2 = GeneralTensorNetwork([t1, t2], [Summation([(1, 3), (2, 4)])],
                               [(1, 1), (1, 2), (1, 4), (2, 1), (2, 2),
                                (2, 3)])
    assert close(contract_svd(t1, t2, (1, 3)).data, contract(tn1))
    assert close(contract_svd(t1, t2, (3, 4)).data, contract(tn2))
    assert close(contract_svd(t1, t2, (3, 4)).data,
                 O.contract_svd(t1.data, t2.data, (3, 4)))


def test_contract_svd_errors(gpu_ctx):
    """test/test_svd.jl:28-50."""
    rng = np.random.default_rng(17)
    t1 = Tensor(crand(rng, 2, 3, 4, 2))
    t2 = Tensor(crand(rng, 1, 5, 2, 4))
    with pytest.raises(ErrorException, match="^Error must be positive$"):
        contract_svd(t1, t2, (1, 3), er=-0.3)
    with pytest.raises(ErrorException,
                       match="^Dimensions of contraction legs do not match$"):
        contract_svd(t1, Tensor(crand(rng, 8, 8)), (1, 3))


def test_contract_svd_truncation_error_bound(gpu_ctx):
    """test/test_svd.jl:53-83: spectra exp(-(0:99)), er = 1e-10: the
    truncated product stays within 2 n t + t^2 of the exact one, and equals
    the oracle's truncated product."""
    rng = np.random.default_rng(19)
    s = np.exp(-np.arange(100.0))
    q = [np.linalg.qr(rng.random((100, 100)))[0] for _ in range(4)]
    T1 = (q[0] * s) @ q[1]
    T2 = (q[2] * s) @ q[3]
    er = 1e-10
    mps = ClosedMPS([Tensor(T1.astype(complex)), Tensor(T2.astype(complex))])
    approx = contract_svd_mps(mps, er=er)
    exact = contract(mps)
    k = O.truncation_rank(s, er)
    assert k == 24
    nt = np.linalg.norm(s[k:])
    bound = 2 * np.linalg.norm(s) * nt + nt ** 2
    assert np.linalg.norm(approx - exact) < bound
    ref = O.contract_svd(T1, T2, (2, 1), er=er)
    # the kept subspace is separated by a factor e from the dropped one
    assert np.linalg.norm(approx - ref) <= 1e-12


def test_contract_svd_mps_open_chain(gpu_ctx):
    """test/test_mps.jl:9-19 (and the periodic / negative-er errors)."""
    rng = np.random.default_rng(23)
    T = Tensor(crand(rng, 2, 2, 2))
    mps = OpenMPS(T, 3)
    g = GeneralTensorNetwork(mps.tensors, mps.contractions, mps.openidx)
    assert close(contract_svd_mps(mps, er=0.0), contract(g))
    with pytest.raises(ErrorException, match="periodic boundary"):
        contract_svd_mps(PeriodicMPS(T, 3), er=0.0)


# -- MPS(psi), switch!, permute! ------------------------------------------
def product_state(rng, n):
    b = [crand(rng, 2) for _ in range(n)]
    psi = b[0]
    for v in b[1:]:
        psi = np.kron(psi, v)
    return b, psi


def kron_all(vs):
    out = vs[0]
    for v in vs[1:]:
        out = np.kron(out, v)
    return out


def test_mps_from_product_state_round_trip(gpu_ctx):
    """test/test_mps.jl:34-49."""
    _, psi = product_state(np.random.default_rng(29), 5)
    mps = MPS(psi)
    assert close(contract(mps).reshape(-1, order="F"), psi)


def test_mps_from_random_state_matches_oracle(gpu_ctx):
    rng = np.random.default_rng(31)
    psi = crand(rng, 2 ** 8)
    mps = MPS(psi)
    ref = O.mps_from_state(psi)
    assert [t.size() for t in mps.tensors] == [t.shape for t in ref]
    assert close(contract(mps).reshape(-1, order="F"), psi)
    # a full-rank state: every site but the last is a left isometry
    for t in mps.tensors[:-1]:
        u = t.data.reshape((-1, t.size()[-1]), order="F")
        assert close(u.conj().T @ u, np.eye(u.shape[1]), 1e-12)


def test_svd_untruncated_factors(gpu_ctx):
    rng = np.random.default_rng(37)
    for (m, n) in [(8, 3), (3, 8), (64, 64), (130, 70)]:
        A = crand(rng, m, n)
        U, S, Vh = ops.svd(A)
        assert close(S, np.linalg.svd(A, compute_uv=False), 1e-12)
        assert close((U * S) @ Vh, A, 1e-12)
        assert close(U.conj().T @ U, np.eye(min(m, n)), 1e-12)
        _, S2, SVh = ops.svd(A, absorb=True)
        assert close(SVh, S[:, None] * Vh, 1e-12)


@pytest.mark.parametrize("order", [[1, 3, 2, 5, 6, 4], [4, 1, 6, 2, 3, 5]])
def test_permute_mps_product_state(gpu_ctx, order):
    """test/test_mps.jl:153-170."""
    b, psi = product_state(np.random.default_rng(41), 6)
    mps = MPS(psi)
    permute_mps(mps, order)
    ref = kron_all([b[o - 1] for o in order])
    assert close(contract(mps).reshape(-1, order="F"), ref)


def test_switch_pairs_product_state(gpu_ctx):
    """test/test_mps.jl:198-222."""
    b, psi = product_state(np.random.default_rng(43), 6)
    ref = kron_all([b[o - 1] for o in [1, 3, 2, 5, 6, 4]])
    mps = MPS(psi)
    switch(mps, [(2, 3), (4, 6), (4, 5)])
    assert close(contract(mps).reshape(-1, order="F"), ref)
    mps = MPS(psi)
    switch(mps, 2, 3)
    switch(mps, 4, 6)
    switch(mps, 4, 5)
    assert close(contract(mps).reshape(-1, order="F"), ref)


def test_switch_neighbouring_sites(gpu_ctx):
    """test/test_mps.jl:241-259."""
    b, psi = product_state(np.random.default_rng(47), 6)
    ref = kron_all([b[o - 1] for o in [1, 3, 4, 5, 6, 2]])
    mps = MPS(psi)
    for i in (4, 3, 2, 1):
        switch(mps, i)
    assert close(contract(mps).reshape(-1, order="F"), ref)


def test_switch_entangled_state_vs_dense(gpu_ctx):
    """A swap of wires i, j on a random (entangled) 7-qubit state equals the
    dense axis swap (the reference tests only product states)."""
    rng = np.random.default_rng(53)
    psi = crand(rng, 2 ** 7)
    psi /= np.linalg.norm(psi)
    mps = MPS(psi)
    switch(mps, 2, 6)
    t = psi.reshape((2,) * 7, order="F")
    ref = np.swapaxes(t, 1, 5).reshape(-1, order="F")
    assert close(contract(mps).reshape(-1, order="F"), ref, 1e-9)


# -- MPO, extend_MPO, apply_MPO --------------------------------------------
def mpo_tensor(mpo):
    """contract(MPO) as the reference does: output legs then input legs."""
    return contract(mpo)


def test_mpo_construction_and_contract(gpu_ctx):
    """test/test_mpo.jl:25-36 plus contract(MPO(U)) == U."""
    rng = np.random.default_rng(59)
    for M in (1, 2, 3):
        U = haar(rng, 2 ** M)
        mpo = MPO(U)
        assert len(mpo.tensors) == M
        assert mpo.isapprox(MPO(CircuitGate(tuple(range(1, M + 1)), U)))
        # output legs (rows) last wire first, then inputs: reversing both
        # halves gives U as a (2,..,2) tensor
        t = mpo_tensor(mpo)
        if M == 1:
            assert close(t, U.T)
            continue
        perm = list(range(M - 1, -1, -1)) + list(range(2 * M - 1, M - 1,
                                                       -1))
        t = np.transpose(t, perm)
        assert close(t.reshape((2 ** M, 2 ** M), order="F"), U)
        ref, spectra = O.mpo_from_matrix(U)
        assert [x.size() for x in mpo.tensors] == [x.shape for x in ref]


def test_mpo_operator_schmidt_closed_forms(gpu_ctx):
    """SWAP splits into four unit-weight terms, CNOT into two of weight
    sqrt(2)."""
    _, S, _ = ops.svd(ops.permute(SWAP.reshape((2,) * 4, order="F"),
                                  [1, 3, 2, 4]).reshape((4, 4), order="F"))
    assert close(S, np.ones(4), 1e-12)
    cnot = controlled(X)
    _, S, _ = ops.svd(ops.permute(cnot.reshape((2,) * 4, order="F"),
                                  [1, 3, 2, 4]).reshape((4, 4), order="F"))
    assert close(S[:2], np.full(2, np.sqrt(2)), 1e-12)
    assert np.all(np.abs(S[2:]) <= 1e-12)


def test_extend_mpo_self_consistency(gpu_ctx):
    """test/test_mpo.jl:38-59."""
    rng = np.random.default_rng(61)
    A, B = crand(rng, 2, 2), crand(rng, 2, 2)
    U1 = np.kron(np.kron(A, np.eye(2)), B)
    U2 = np.kron(A, B)
    ref = contract(MPO(U1))
    assert close(contract(extend_MPO(U2, (3, 1))), ref)
    assert close(contract(extend_MPO(MPO(U2), (3, 1))), ref)
    # against the dense operator of kron(QA, QB) on wires (3, 1)
    QA, QB = haar(rng, 2), haar(rng, 2)
    U = np.kron(QA, QB)
    dense = np.stack([O.apply_mpo_dense(e, U, (1, 3))
                      for e in np.eye(8, dtype=complex)], axis=1)
    t = contract(extend_MPO(U, (3, 1)))
    # reference layout: rows of wires 3,2,1 then columns of wires 3,2,1
    t = np.transpose(t, [2, 1, 0, 5, 4, 3]).reshape((8, 8), order="F")
    assert close(t, dense)


@pytest.mark.parametrize("targ,cntrl", [(1, 2), (4, 2), (5, 1), (2, 5)])
def test_apply_mpo_cnot(gpu_ctx, targ, cntrl):
    """test/test_mpo.jl:62-86."""
    rng = np.random.default_rng(67 + targ * 7 + cntrl)
    _, psi = product_state(rng, 5)
    mps = MPS(psi)
    w = (targ, cntrl)
    cnot = controlled(X)  # bit 0 target, bit 1 control
    ref = O.apply_mpo_dense(psi, cnot, w)
    # an MPO pairs gate bit k-1 with wire w[k] in any order; a matrix is
    # permuted to sorted wires first
    for net in (apply_MPO(mps, MPO(cnot), w), apply_MPO(mps, cnot, w)):
        assert close(contract(net).reshape(-1, order="F"), ref)


def test_apply_mpo_multi_qubit_all_wire_sets(gpu_ctx):
    """test/test_mpo.jl:88-121: a Haar 4-qubit gate on every sorted wire set
    of 8 qubits (a sample of them unsorted) of a random closed MPS."""
    rng = np.random.default_rng(71)
    N, M = 8, 4
    U = haar(rng, 2 ** M)
    tens = [crand(rng, 2, 2)] + [crand(rng, 2, 2, 2) for _ in range(N - 2)] \
        + [crand(rng, 2, 2)]
    mps = ClosedMPS([Tensor(t) for t in tens])
    psi = contract(mps).reshape(-1, order="F")
    combos = list(itertools.combinations(range(1, N + 1), M))
    for com in combos[::5]:
        net = apply_MPO(mps, MPO(U), com)
        ref = O.apply_mpo_dense(psi, U, com)
        assert close(contract(net).reshape(-1, order="F"), ref)
        w = tuple(int(x) for x in rng.permutation(np.arange(1, N + 1))[:M])
        net = apply_MPO(mps, U, w)
        ref = O.apply_mpo_dense(psi, U, w)
        assert close(contract(net).reshape(-1, order="F"), ref)
        assert close(contract(apply_MPO(mps, CircuitGate(w, U))),
                     contract(net))


# -- tensor_circuit! -------------------------------------------------------
def gates_qft(n):
    """QFT-like gate list (Hadamards and controlled phases)."""
    out = []
    for j in range(n, 0, -1):
        out.append(CircuitGate((j,), H))
        for k in range(j - 1, 0, -1):
            ph = np.diag([1, np.exp(1j * np.pi / 2 ** (j - k))])
            out.append(CircuitGate((j, k), controlled(ph)))
    return out


def dense_apply(psi, gates):
    for g in gates:
        psi = O.apply_mpo_dense(psi, g.matrix, g.iwire)
    return psi


def test_tensor_circuit_dense_gates(gpu_ctx):
    """test/test_tensor_circuit.jl:25-52 with the QFT, plus non-symmetric
    Haar gates (the dense path applies the matrix itself)."""
    rng = np.random.default_rng(73)
    net = mps_network(rng)
    psi = contract(net).reshape(-1, order="F")
    gates = gates_qft(3) + [CircuitGate((2, 3), haar(rng, 4)),
                            CircuitGate((1,), haar(rng, 2))]
    tensor_circuit(net, gates)
    ref = dense_apply(psi, gates)
    assert close(contract(net).reshape(-1, order="F"), ref)
    assert close(contract(net, True).reshape(-1, order="F"), ref)


def test_tensor_circuit_decompose_2q_and_3q(gpu_ctx):
    """test/test_tensor_circuit.jl:54-130 (controlled X/Y/Z, SWAP,
    Toffoli) plus a Haar 3-qubit gate, decomposed."""
    rng = np.random.default_rng(79)
    net = mps_network(rng)
    psi = contract(net).reshape(-1, order="F")
    gates = [CircuitGate((3, 1), controlled(X)),
             CircuitGate((3, 1), controlled(Y)),
             CircuitGate((1, 2), controlled(Y)),
             CircuitGate((2, 1), controlled(Z))]
    tensor_circuit(net, gates, is_decompose=True)
    assert close(contract(net).reshape(-1, order="F"),
                 dense_apply(psi, gates))

    t = [Tensor(crand(rng, 2, 6)), Tensor(crand(rng, 2, 6, 7)),
         Tensor(crand(rng, 2, 7, 5)), Tensor(crand(rng, 2, 5))]
    net = GeneralTensorNetwork(
        t, [Summation([(1, 2), (2, 2)]), Summation([(2, 3), (3, 2)]),
            Summation([(3, 3), (4, 2)])], [(1, 1), (2, 1), (3, 1), (4, 1)])
    psi = contract(net).reshape(-1, order="F")
    cswap = np.eye(8, dtype=complex)
    cswap[[5, 6]] = cswap[[6, 5]]  # SWAP of bits 0,1 controlled by bit 2
    gates = [CircuitGate((4, 1, 3), cswap), CircuitGate((2,), Y),
             CircuitGate((3, 1), controlled(X)), CircuitGate((2,), Y),
             CircuitGate((4, 1, 2), controlled(X, 2)),
             CircuitGate((3,), Z),
             CircuitGate((2, 4, 1), haar(rng, 8))]
    tensor_circuit(net, gates, is_decompose=True)
    assert close(contract(net).reshape(-1, order="F"),
                 dense_apply(psi, gates))
    tt, c, w = decompose(gates[-1])
    assert c == [0, 3, 4] and w == [2, 4, 1]
    b1 = tt[0].size()[2]
    assert tt[0].size() == (2, 2, b1)
    assert tt[1].size()[:3] == (b1, 2, 2)
    assert tt[2].size() == (tt[1].size()[3], 2, 2)
