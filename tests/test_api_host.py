"""Host-side logic of the API mirror that needs no GPU: structure checks,
error messages (the reference's exact strings) and network wiring."""
import numpy as np
import pytest

from qaintensor import (MPO, MPS, CircuitGate, ClosedMPS, ErrorException,
                        GeneralTensorNetwork, OpenMPS, PeriodicMPS,
                        Summation, Tensor, apply_MPO, check_mps, contract_rep,
                        contract_svd, contract_svd_mps, extend_MPO,
                        is_power_two, permute_mps, shift_pair,
                        shift_summation, switch, tensor_circuit)


def t3():
    return Tensor(np.ones((2, 2, 2), dtype=complex))


def t2():
    return Tensor(np.ones((2, 2), dtype=complex))


def test_tensor_helpers():
    """test/test_tensor_circuit.jl:7-19."""
    a = np.arange(12.0).reshape((2, 6), order="F") + 0j
    t = Tensor(a)
    assert t.reshape((3, 4)).isapprox(t.reshape(3, 4))
    assert np.array_equal(t.reshape((3, 4)).data,
                          a.reshape((3, 4), order="F"))
    assert t.reshape((3, 4)).transpose().isapprox(
        Tensor(a.reshape((3, 4), order="F").T))
    assert t.size() == (2, 6) and t.size(3) == 1 and t.ndims() == 2


def test_helpers():
    """test/test_helper.jl."""
    assert shift_summation(Summation([(1, 2), (3, 4)]), 2) == \
        Summation([(3, 2), (5, 4)])
    assert shift_pair((1, 2), 3) == (4, 2)
    assert [is_power_two(i) for i in (0, 1, 2, 3, 4, 6, 8, 1024)] == \
        [False, True, True, False, True, False, True, True]


def test_contract_rep_labels():
    """src/contract.jl:39-60: contraction i -> label i, open leg i -> label
    -i - #contractions."""
    net = GeneralTensorNetwork(
        [Tensor(np.ones((2, 6))), Tensor(np.ones((2, 6, 7))),
         Tensor(np.ones((2, 7)))],
        [Summation([(1, 2), (2, 2)]), Summation([(2, 3), (3, 2)])],
        [(1, 1), (2, 1), (3, 1)])
    assert contract_rep(net) == [[-3, 1], [-4, 1, 2], [-5, 2]]
    net.openidx[0] = (2, 2)
    with pytest.raises(AssertionError):
        contract_rep(net)


def test_mps_constructor_checks():
    """test/test_mps.jl:57-104 and src/mps.jl:15-27."""
    T = [t2(), t3(), t2()]
    with pytest.raises(ErrorException, match="^Tensor objects first leg"):
        MPS(T, [Summation([(1, 1), (2, 1)]), Summation([(2, 3), (3, 1)])],
            [(1, 1), (2, 2), (3, 2)])
    with pytest.raises(ErrorException, match="^Tensor objects last leg"):
        MPS(T, [Summation([(1, 2), (2, 2)]), Summation([(2, 3), (3, 1)])],
            [(1, 1), (2, 2), (3, 2)])
    with pytest.raises(ErrorException, match="can only have 2 or 3 legs"):
        MPS([t2(), Tensor(np.ones((2,) * 4)), t2()],
            [Summation([(1, 2), (2, 1)]), Summation([(2, 4), (3, 1)])],
            [(1, 1), (2, 2), (3, 2)])


def test_check_mps_errors():
    """test/test_mps.jl:21-32."""
    mps = OpenMPS(t3(), 3)
    check_mps(mps)
    mps.contractions[0] = Summation([(1, 2), (2, 1)])
    with pytest.raises(ErrorException, match="^Tensor objects first leg"):
        check_mps(mps)
    mps.contractions[0] = Summation([(1, 3), (2, 2)])
    with pytest.raises(ErrorException, match="^Tensor objects last leg"):
        check_mps(mps)
    mps.tensors[1] = Tensor(np.ones((2,) * 4))
    with pytest.raises(ErrorException, match="can only have 2 or 3 legs"):
        check_mps(mps)


def test_open_closed_periodic_layouts_and_errors():
    """src/mps.jl:99-182, test/test_mps.jl:106-150."""
    o = OpenMPS(t3(), 3)
    assert o.openidx == [(1, 1), (1, 2), (2, 2), (3, 2), (3, 3)]
    assert [s.idx for s in o.contractions] == [[(1, 3), (2, 1)],
                                               [(2, 3), (3, 1)]]
    c = ClosedMPS(t2(), t3(), t2(), 4)
    assert c.openidx == [(1, 1), (2, 2), (3, 2), (4, 2)]
    assert [s.idx for s in c.contractions] == [[(1, 2), (2, 1)],
                                               [(2, 3), (3, 1)],
                                               [(3, 3), (4, 1)]]
    p = PeriodicMPS(t3(), 3)
    assert p.contractions[-1] == Summation([(3, 3), (1, 1)])
    with pytest.raises(ErrorException, match="^Tensors must have 3 legs$"):
        OpenMPS([t2(), t3()])
    with pytest.raises(ErrorException, match="^Tensors must have 3 legs$"):
        OpenMPS(Tensor(np.ones((2,) * 4)), 3)
    with pytest.raises(ErrorException, match="^First tensor must have 2"):
        ClosedMPS([t3(), t3(), t2()])
    with pytest.raises(ErrorException, match="except the first and last"):
        ClosedMPS([t2(), t2(), t2()])
    with pytest.raises(ErrorException, match="^Last tensor must have 2"):
        ClosedMPS([t2(), t3(), t3()])
    # periodic and negative er are refused before any device work
    with pytest.raises(ErrorException, match="periodic boundary"):
        contract_svd_mps(p, er=0.0)
    with pytest.raises(ErrorException, match="^Error must be positive$"):
        contract_svd_mps(o, er=-0.5)
    with pytest.raises(ErrorException, match="^Error must be positive$"):
        contract_svd(t3(), t3(), (3, 1), er=-0.3)
    with pytest.raises(ErrorException, match="legs do not match"):
        contract_svd(t3(), Tensor(np.ones((8, 8))), (1, 1))


def test_state_length_error():
    """test/test_mps.jl:51-55."""
    with pytest.raises(ErrorException,
                       match="^Input state must have length 2\\^N$"):
        MPS(np.ones(6, dtype=complex))


def test_permute_and_switch_errors():
    """test/test_mps.jl:172-196, 224-239."""
    mps = ClosedMPS(t2(), t3(), t2(), 6)
    for order, msg in [([1, 2, 3], "must be same length"),
                       ([1, 2, 3, 4, 5, 6, 7], "must be same length"),
                       ([1] * 6, "cannot contain repeat values"),
                       ([1, -2, 3, 4, -5, 6], "only contain positive"),
                       ([1, 2, 8, 4, 5, 6], "cannot exceed number of wires")]:
        with pytest.raises(ErrorException, match=msg):
            permute_mps(mps, order)
    with pytest.raises(ErrorException, match="must be positive"):
        switch(mps, 1, -2)
    with pytest.raises(ErrorException, match="less than or equal to"):
        switch(mps, 7, 4)
    with pytest.raises(ErrorException, match="must be positive"):
        switch(mps, [(1, -2)])
    with pytest.raises(ErrorException, match="less than or equal to"):
        switch(mps, [(9, 10), (1, -2)])
    with pytest.raises(IndexError):
        switch(mps, 6)


def fake_mpo(M):
    """An MPO object with the layout of MPO(m) for M > 1, built without
    device work (the wiring checks come before any)."""
    mpo = MPO.__new__(MPO)
    mpo.tensors = [Tensor(np.ones((2, 2, 2)))] + \
        [Tensor(np.ones((2, 2, 2, 2))) for _ in range(M - 2)] + \
        [Tensor(np.ones((2, 2, 2)))]
    mpo.contractions = [Summation([(1, 3), (2, 1)])] + \
        [Summation([(i, 4), (i + 1, 1)]) for i in range(2, M)]
    mpo.openidx = [(M - i + 1, 2) for i in range(1, M)] + [(1, 1)] + \
        [(M - i + 1, 3) for i in range(1, M)] + [(1, 2)]
    return mpo


def test_apply_mpo_wire_errors():
    """test/test_mpo.jl:124-141."""
    psi = ClosedMPS(t2(), t3(), t2(), 5)
    with pytest.raises(ErrorException, match="^Repeated wires are not"):
        apply_MPO(psi, fake_mpo(3), (1, 3, 3))
    for w in [(1, -2, 4), (1, 3, 6)]:
        with pytest.raises(ErrorException, match="between 1 and n"):
            apply_MPO(psi, fake_mpo(3), w)
    with pytest.raises(ErrorException, match="^Wires not sorted$"):
        extend_MPO(fake_mpo(2), (1, 3))
    with pytest.raises(ErrorException, match="^Repeated wires"):
        extend_MPO(fake_mpo(2), (3, 3))
    with pytest.raises(ErrorException, match="already decomposed"):
        extend_MPO(fake_mpo(2), (2, 1))
    with pytest.raises(ErrorException, match="Direct conversion"):
        MPO([t2()], [], [])
    with pytest.raises(ErrorException, match="Need at least one qubit"):
        MPO(np.ones((1, 1)))


def test_apply_mpo_wiring():
    """src/mpo.jl:204-215: the state's open leg of wire iwire[k] meets the
    MPO input leg of gate bit k-1 and is replaced by its output leg."""
    psi = ClosedMPS(t2(), t3(), t2(), 4)
    net = apply_MPO(psi, fake_mpo(2), (2, 4))
    # MPO tensors 5 (bit 0 <-> wire 2) and 6 (bit 1 <-> wire 4)
    assert net.contractions[-2:] == [Summation([(4, 2), (6, 3)]),
                                     Summation([(2, 2), (5, 2)])]
    assert net.openidx == [(1, 1), (5, 1), (3, 2), (6, 2)]
    assert net.contractions[-3] == Summation([(5, 3), (6, 1)])
    # one-qubit MPO: the reference's openidx is (col, row)
    m = MPO(np.array([[1, 2], [3, 4]], dtype=complex))
    assert m.openidx == [(1, 2), (1, 1)]


def test_tensor_circuit_wiring():
    """src/tensor_circuit.jl:14-52: the gate tensor is appended and the
    wires re-pointed; the dense path contracts the gate's column legs."""
    psi = ClosedMPS(t2(), t3(), t2(), 3)
    g = CircuitGate((3, 1), np.eye(4))
    tensor_circuit(psi, g)
    assert psi.openidx == [(4, 2), (2, 2), (4, 1)]
    assert psi.contractions[-2:] == [Summation([(3, 2), (4, 3)]),
                                     Summation([(1, 1), (4, 4)])]
