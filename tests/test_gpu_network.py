"""contract(net) on the device (qtn_contract_network): whole networks with
their intermediates on the GPU, every pairwise step one gather-staged
complex MFMA GEMM.  Checked against the oracle's ncon (einsum, the
definition) and, for the QFT expectation networks of
test/test_treewidth.jl:352-368, against the dense bilinear form T0^T U T0
(statevector simulation of the circuit).  Tolerance 1e-10 relative (FP64)."""
import numpy as np
import pytest

from oracle import qtn_oracle as O
from qaintensor import circuits as C, ops
from qaintensor import network2graph as NG
from qaintensor import workloads as Wk
from qaintensor._lib import QTN_ERR_ARG, QtnError
from qaintensor.mps import ClosedMPS
from qaintensor.network import (GeneralTensorNetwork, Summation, Tensor,
                                contract, contract_rep)

pytestmark = pytest.mark.gpu


def crand(rng, *dims):
    return rng.standard_normal(dims) + 1j * rng.standard_normal(dims)


def close(a, b, tol=1e-10):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and \
        np.max(np.abs(a - b)) <= tol * max(1.0, np.max(np.abs(b)))


@pytest.mark.parametrize("order", ["increasing", "reversed", "shuffled"])
def test_network_executor_matches_ncon(gpu_ctx, order):
    """A ring of four tensors with two open legs, a trace inside one of
    them, and an unconnected vector (outer product), under three
    contraction orders."""
    rng = np.random.default_rng(3)
    a = crand(rng, 3, 4, 5, 2, 2)      # legs 1..3 ring/open, 4-5 a trace
    b = crand(rng, 4, 6)
    c = crand(rng, 6, 7, 8)
    e = crand(rng, 7, 3)
    v = crand(rng, 9)
    labels = [[4, 1, -2, 5, 5], [1, 2], [2, 3, -1], [3, 4], [-3]]
    pos = [1, 2, 3, 4, 5]
    ordr = {"increasing": pos, "reversed": pos[::-1],
            "shuffled": [3, 5, 1, 4, 2]}[order]
    got = ops.contract_network([a, b, c, e, v], labels, ordr)
    ref = O.ncon([a, b, c, e, v], labels)
    assert close(got, ref)


def test_network_executor_large_steps_and_leg_orders(gpu_ctx):
    """Steps big enough for 64 x 64 tiles, with the unit-stride leg among
    the contracted legs of one operand and the free legs of the other."""
    rng = np.random.default_rng(4)
    A = crand(rng, 16, 40, 12, 3)
    B = crand(rng, 12, 70, 16)
    Cm = crand(rng, 3, 70, 5)
    labels = [[1, -1, 2, 3], [2, 4, 1], [3, 4, -2]]
    got = ops.contract_network([A, B, Cm], labels, [1, 2, 3, 4])
    assert close(got, O.ncon([A, B, Cm], labels))
    got = ops.contract_network([A, B, Cm], labels, [4, 3, 2, 1])
    assert close(got, O.ncon([A, B, Cm], labels))


def test_contract_open_leg_order_and_single_tensor(gpu_ctx):
    rng = np.random.default_rng(5)
    a, b = crand(rng, 2, 3, 4), crand(rng, 4, 5)
    net = GeneralTensorNetwork([Tensor(a), Tensor(b)],
                               [Summation([(1, 3), (2, 1)])],
                               [(2, 2), (1, 1), (1, 2)])
    assert close(contract(net), np.einsum("ijk,kl->lij", a, b))
    single = GeneralTensorNetwork([Tensor(a)], [], [(1, 3), (1, 1), (1, 2)])
    assert close(contract(single), np.transpose(a, (2, 0, 1)))


def test_qft4_expectation_network_both_orders(gpu_ctx):
    """test/test_treewidth.jl:370-377: the N = 4 QFT expectation network
    contracts to the same scalar in ncon's default order and in the order
    contract_order plans (applied as planned), for the dense and the
    decomposed gates."""
    for is_decompose in (False, True):
        net, _ = Wk.qft_expectation_network(4, 7, is_decompose=is_decompose)
        ref = O.ncon([t.data for t in net.tensors], contract_rep(net))
        assert close(contract(net), ref)
        assert close(contract(net, True), ref)


@pytest.mark.parametrize("is_decompose", [False, True])
def test_qft_networks_in_treewidth_order(gpu_ctx, is_decompose):
    """optimize_contraction_order! (src/network2graph.jl:473-479) then
    contract(net): the QFT-4 expectation networks of
    test/test_treewidth.jl:370-377 against ncon in the original order, and
    QFT-10 against the dense bilinear form."""
    for n in (4, 10):
        net0, t0 = Wk.qft_expectation_network(n, 7, is_decompose=is_decompose)
        net = net0.copy()
        NG.optimize_contraction_order(net)
        got = complex(contract(net))
        if n == 4:
            ref = complex(O.ncon([t.data for t in net0.tensors],
                                 contract_rep(net0)))
        else:
            ref = dense_expectation(n, t0)
        assert abs(got - ref) <= 1e-10 * abs(ref)


def dense_expectation(n, t0):
    v = O.ncon(t0, contract_rep(ClosedMPS([Tensor(x) for x in t0])))
    v = np.asarray(v).reshape(-1, order="F")
    sv = O.StateVector(n)
    sv.psi = v.copy()
    for op in C.qft_gates(n):
        if op[0] == "1q":
            sv.apply_1q(op[1], op[2])
        else:
            sv.apply_2q_any(op[1], op[2], op[3])
    return np.sum(v * sv.psi)


@pytest.mark.parametrize("n", [8, 12])
def test_qft_expectation_network_against_statevector(gpu_ctx, n):
    """Larger QFT expectation networks (ncon's default order) against the
    dense bilinear form computed by statevector simulation."""
    net, t0 = Wk.qft_expectation_network(n, n)
    got = complex(contract(net))
    ref = dense_expectation(n, t0)
    assert abs(got - ref) <= 1e-10 * abs(ref)


def test_network_errors(gpu_ctx):
    a, b = np.zeros((2, 3), complex), np.zeros((4, 2), complex)
    with pytest.raises(QtnError) as e:
        ops.contract_network([a, b], [[1, -1], [1, -2]], [1])
    assert e.value.message == "Dimensions of contraction legs do not match"
    with pytest.raises(QtnError) as e:
        ops.contract_network([a, a.T], [[1, -1], [-2, 1]], [2])
    assert e.value.status == QTN_ERR_ARG
