"""SWAP routing on the device MPS (qtn_mps_switch, qtn_mps_permute,
qtn_mps_apply_2q_long): switch! and permute! (src/switch.jl:63-106) with
the reference's most-significant-first wire numbering, and two-qubit gates
on non-adjacent wires.  Checked against the reference's own switch case
(test/test_mps.jl:198-221), the dense statevector, and the oracle's
TebdState running the same routing (kept ranks identical).  Tolerance
1e-10 relative on amplitudes."""
import numpy as np
import pytest

from oracle import qtn_oracle as O
from qaintensor import DeviceMPS, circuits as C
from qaintensor._lib import QTN_ERR_ARG, QtnError

pytestmark = pytest.mark.gpu


def all_bits(n):
    return np.array([[(i >> w) & 1 for w in range(n)]
                     for i in range(2 ** n)], dtype=np.uint8)


def dense(dev):
    return dev.amplitudes(all_bits(dev.n))


def same_state(a, b, tol=1e-10):
    """Equal up to normalisation and a global phase (set_sites keeps the
    state's norm and phase out of the device tensors)."""
    a = a / np.linalg.norm(a)
    b = b / np.linalg.norm(b)
    ov = np.vdot(a, b)
    return np.linalg.norm(a * (ov / abs(ov)) - b) <= tol


def test_switch_reference_case(gpu_ctx):
    """switch!(mps, [(2,3), (4,6), (4,5)]) on kron(b...) gives
    kron(b[[1,3,2,5,6,4]]...), then the same by single calls."""
    rng = np.random.default_rng(11)
    N = 6
    b = [rng.standard_normal(2) + 1j * rng.standard_normal(2)
         for _ in range(N)]
    order = [1, 3, 2, 5, 6, 4]
    psi_perm = b[order[0] - 1]
    for k in order[1:]:
        psi_perm = np.kron(psi_perm, b[k - 1])
    sites = [np.asarray(b[N - w], dtype=np.complex128).reshape((1, 2, 1))
             for w in range(1, N + 1)]
    for as_list in (True, False):
        dev = DeviceMPS(N, 8, gpu_ctx)
        dev.set_sites(sites)
        if as_list:
            dev.switch([(2, 3), (4, 6), (4, 5)])
        else:
            dev.switch(2, 3)
            dev.switch(4, 6)
            dev.switch(4, 5)
        assert same_state(dense(dev), psi_perm, 1e-12)


def test_switch_neighbouring_sites(gpu_ctx):
    """test/test_mps.jl:240-259: switch!(mps, 4), (3), (2), (1) carries the
    qubit of site 1 to site 5: kron(b[[1,3,4,5,6,2]]...)."""
    rng = np.random.default_rng(12)
    N = 6
    b = [rng.standard_normal(2) + 1j * rng.standard_normal(2)
         for _ in range(N)]
    order = [1, 3, 4, 5, 6, 2]
    psi_perm = b[order[0] - 1]
    for k in order[1:]:
        psi_perm = np.kron(psi_perm, b[k - 1])
    dev = DeviceMPS(N, 8, gpu_ctx)
    dev.set_sites([np.asarray(b[N - w], dtype=np.complex128).reshape(
        (1, 2, 1)) for w in range(1, N + 1)])
    for k in (4, 3, 2, 1):
        dev.switch(k)
    assert same_state(dense(dev), psi_perm, 1e-12)


def test_permute_entangled_state_against_oracle(gpu_ctx):
    """permute!(mps, order) on an entangled 9-qubit state: the oracle runs
    the reference's loop of switch! calls (src/switch.jl:100-105)."""
    n = 9
    layers = C.brickwall_layers(n, 6, 31)
    dev = DeviceMPS(n, 32, gpu_ctx)
    dev.run_layers(layers, 0.0, 32)
    st = O.TebdState(n)
    O.run_layers(st, layers, 0.0, 0)
    order = [3, 9, 1, 5, 2, 8, 4, 7, 6]
    dev.permute(order)
    seq = list(range(1, n + 1))
    for i in range(1, n + 1):
        loc = seq.index(order[i - 1]) + 1
        st.switch(i, loc)
        seq[loc - 1], seq[i - 1] = seq[i - 1], seq[loc - 1]
    # (bond dimensions are not compared: at cutoff 0 the device drops
    # columns at the rounding floor as exact zeros, LAPACK returns them)
    assert same_state(dense(dev), st.to_statevector())


def test_qft_long_range_gates_match_dense_and_ranks(gpu_ctx):
    """The QFT on 10 qubits (controlled phases between every pair of wires,
    routed by SWAPs) from a random product state: amplitudes match the
    dense statevector at 1e-10, and every gate keeps the rank the oracle's
    routing keeps."""
    n, chi, cut = 10, 64, 1e-12
    rng = np.random.default_rng(5)
    dev = DeviceMPS(n, chi, gpu_ctx)
    st = O.TebdState(n)
    sv = O.StateVector(n)
    for w in range(1, n + 1):
        u = C.haar_unitary(rng, 2)
        dev.apply_1q(w, u)
        st.apply_1q(w, u)
        sv.apply_1q(w, u)
    ranks, ranks_o = [], []
    for op in C.qft_gates(n):
        if op[0] == "1q":
            dev.apply_1q(op[1], op[2])
            st.apply_1q(op[1], op[2])
            sv.apply_1q(op[1], op[2])
        else:
            _, i, j, g = op
            k, _ = dev.apply_2q_long(i, j, g, cut, chi)
            ranks.append(k)
            ranks_o.append(st.apply_2q_long(i, j, g, cut, chi)[0])
            ranks[-1] = (ranks[-1], tuple(dev.bond_dims()))
            ranks_o[-1] = (ranks_o[-1], tuple(st.bond_dims()))
            sv.apply_2q_any(i, j, g)
    assert ranks == ranks_o
    assert max(dev.bond_dims()) > 1
    amp = dense(dev)
    assert np.linalg.norm(amp - sv.psi) <= 1e-10 * np.linalg.norm(sv.psi)


def test_long_range_gate_orders_and_chosen_form(gpu_ctx):
    """apply_2q_long(i, j, g) and apply_2q_long(j, i, P g P) (P the SWAP)
    give the same state; the chosen form stays within 1e-6 of it."""
    n, chi = 12, 64
    layers = C.brickwall_layers(n, 6, 8)
    g = C.haar_unitary(np.random.default_rng(2), 4)
    P = O.SWAP_GATE
    states = []
    for (i, j, gg, form) in ((3, 10, g, 0), (10, 3, P @ g @ P, 0),
                             (3, 10, g, 1)):
        dev = DeviceMPS(n, chi, gpu_ctx)
        dev.run_layers(layers, 1e-12, chi)
        dev.apply_2q_long(i, j, gg, 1e-12, chi, form=form)
        states.append(dense(dev))
    assert same_state(states[0], states[1])
    a, c = states[0], states[2]
    fid = abs(np.vdot(a, c)) ** 2 / (np.vdot(a, a).real * np.vdot(c, c).real)
    assert 1 - fid <= 1e-6


def test_routing_errors_carry_reference_messages(gpu_ctx):
    dev = DeviceMPS(6, 8, gpu_ctx)
    for (i, j), msg in (((1, -2), "Wire indices `i` and `j` must be "
                                  "positive"),
                        ((7, 4), "Indices to swap `i` and `j` must be less "
                                 "than or equal to the number of open wires "
                                 "in MPS")):
        with pytest.raises(QtnError) as e:
            dev.switch(i, j)
        assert e.value.status == QTN_ERR_ARG and e.value.message == msg
    for order, msg in (([1, 2, 3], "Given permutation must be same length "
                                   "as number of Tensors in MPS"),
                       ([1, 2, 2, 4, 5, 6], "Permutation order cannot "
                                            "contain repeat values"),
                       ([1, -2, 3, 4, -5, 6], "Permutation order can only "
                                              "contain positive values"),
                       ([1, 2, 8, 4, 5, 6], "Wire numbers in permutation "
                                            "order cannot exceed number of "
                                            "wires in MPS")):
        with pytest.raises(QtnError) as e:
            dev.permute(order)
        assert e.value.message == msg
    with pytest.raises(QtnError) as e:
        dev.apply_2q_long(3, 3, np.eye(4), 1e-10, 8)
    assert e.value.status == QTN_ERR_ARG
