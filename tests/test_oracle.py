"""Pins the CPU oracle (oracle/qtn_oracle.py) before it is trusted as the
checker: the reference's own known-answer test, closed forms, invariants and
brute force on tiny inputs.  Each check is chosen so that a plausible slip
(dropped term, wrong sign or index, transposed operand) fails it."""
import itertools

import numpy as np
import pytest

from oracle import qtn_oracle as O
from qaintensor import circuits as C


def crand(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


# -- truncation rule (src/svd.jl:29-33) -------------------------------------
def test_truncation_kat_reference_spectrum():
    # test/test_svd.jl:53-83: s = exp(-x), x = 0..99, er = 1e-10
    s = np.exp(-np.arange(100.0))
    k = O.truncation_rank(s, 1e-10)
    assert k == 24
    assert np.linalg.norm(s[24:]) <= 1e-10 < np.linalg.norm(s[23:])
    n = np.linalg.norm(s)
    tail = np.linalg.norm(s[k:])
    assert abs(tail - 4.06e-11) < 1e-13
    assert abs(2 * n * tail + tail ** 2 - 8.73e-11) < 5e-13


def brute_rank(s, er):
    # fewest leading values whose discarded tail has 2-norm <= er
    for k in range(1, len(s) + 1):
        if np.sqrt(np.sum(s[k:] ** 2)) <= er:
            return k
    return len(s)


def test_truncation_matches_brute_force():
    rng = np.random.default_rng(0)
    for _ in range(300):
        n = int(rng.integers(1, 12))
        s = np.sort(rng.random(n) ** rng.integers(1, 8))[::-1]
        er = float(rng.random() * np.linalg.norm(s) * 0.9)
        assert O.truncation_rank(s, er) == brute_rank(s, er)


def test_truncation_edge_cases():
    # er = 0 drops exact zeros only
    assert O.truncation_rank(np.array([3.0, 2.0, 0.0, 0.0]), 0.0) == 2
    assert O.truncation_rank(np.array([3.0, 2.0, 1e-150]), 0.0) == 3
    # the rule squares the values: 1e-300^2 underflows to 0 and is dropped,
    # as S.^2 does in the reference
    assert O.truncation_rank(np.array([3.0, 2.0, 1e-300]), 0.0) == 2
    # every value within er is an error in the reference
    with pytest.raises(O.TruncationError):
        O.truncation_rank(np.array([1.0, 0.5]), 2.0)
    with pytest.raises(O.TruncationError):
        O.truncation_rank(np.zeros(4), 0.0)
    # relative wrapper, chi cap and discarded weight
    # er = 1e-8 * |s| = 1.118e-8: 1e-12 goes, 1e-7 stays
    s = np.array([1.0, 0.5, 1e-7, 1e-12])
    k, disc, tot = O.truncate(s, 1e-8, True, 0)
    assert k == 3 and abs(disc - 1e-24) < 1e-36
    k, disc, tot = O.truncate(s, 1e-8, True, 2)
    assert k == 2 and abs(disc - (1e-14 + 1e-24)) < 1e-26
    # absolute er keeps 1e-9 only when it exceeds er
    assert O.truncate(np.array([1.0, 1e-9]), 1e-10, False)[0] == 2
    assert O.truncate(np.array([1.0, 1e-9]), 1e-9, False)[0] == 1
    assert abs(tot - np.sum(s ** 2)) < 1e-15
    with pytest.raises(ValueError, match="Error must be positive"):
        O.truncate(s, -1.0)


# -- contraction semantics ---------------------------------------------------
def test_contract_network_brute_force():
    rng = np.random.default_rng(1)
    A = crand(rng, 2, 3, 4)
    B = crand(rng, 4, 2, 5)
    # contract A leg 3 with B leg 1 and A leg 1 with B leg 2; open legs:
    # A2, B3
    out = O.contract_network([A, B], [[(1, 3), (2, 1)], [(1, 1), (2, 2)]],
                             [(2, 3), (1, 2)])
    ref = np.zeros((5, 3), dtype=complex)
    for i, j, k, l in itertools.product(range(2), range(3), range(4),
                                        range(5)):
        ref[l, j] += A[i, j, k] * B[k, i, l]
    assert np.allclose(out, ref, rtol=0, atol=1e-13)


def test_contract_pair_outperm():
    rng = np.random.default_rng(2)
    A = crand(rng, 3, 4, 5)
    B = crand(rng, 5, 6)
    C_ = O.contract_pair(A, B, [3], [1], outperm=[3, 1, 2])
    ref = np.einsum("abk,kc->cab", A, B)
    assert np.allclose(C_, ref, rtol=0, atol=1e-13)


def test_contract_svd_closed_forms():
    rng = np.random.default_rng(3)
    t1 = crand(rng, 4, 4)
    t2 = crand(rng, 4, 4)
    assert np.allclose(O.contract_svd(t1, t2, (2, 1)), t1 @ t2, atol=1e-13)
    # test/test_svd.jl:13-25
    t1 = crand(rng, 2, 3, 4, 6)
    t2 = crand(rng, 1, 5, 2, 4)
    tn1 = O.contract_network([t1, t2], [[(1, 1), (2, 3)]],
                             [(1, 2), (1, 3), (1, 4), (2, 1), (2, 2), (2, 4)])
    tn2 = O.contract_network([t1, t2], [[(1, 3), (2, 4)]],
                             [(1, 1), (1, 2), (1, 4), (2, 1), (2, 2), (2, 3)])
    assert np.allclose(tn1, O.contract_svd(t1, t2, (1, 3)), atol=1e-12)
    assert np.allclose(tn2, O.contract_svd(t1, t2, (3, 4)), atol=1e-12)
    with pytest.raises(ValueError, match="Error must be positive"):
        O.contract_svd(t1, t2, (1, 3), er=-0.3)
    with pytest.raises(ValueError, match="do not match"):
        O.contract_svd(t1, crand(rng, 8, 8), (1, 3))


def test_contract_svd_truncation_bound():
    # test/test_svd.jl:53-83 with the reference's bound
    rng = np.random.default_rng(4)
    s = np.exp(-np.arange(100.0))
    mats = [np.linalg.qr(rng.random((100, 100)))[0] for _ in range(4)]
    T1 = mats[0] @ np.diag(s) @ mats[1]
    T2 = mats[2] @ np.diag(s) @ mats[3]
    approx = O.contract_svd(T1, T2, (2, 1), er=1e-10)
    exact = T1 @ T2
    k = O.truncation_rank(s, 1e-10)
    n, nt = np.linalg.norm(s), np.linalg.norm(s[k:])
    assert np.linalg.norm(approx - exact) < 2 * n * nt + nt ** 2


# -- MPS(psi), MPO(m) ---------------------------------------------------------
def test_mps_from_state_round_trip():
    rng = np.random.default_rng(5)
    for n in (2, 3, 5, 7):
        psi = crand(rng, 2 ** n)
        ts = O.mps_from_state(psi)
        assert len(ts) == n
        assert np.allclose(O.mps_to_state(ts), psi, atol=1e-12)
    with pytest.raises(ValueError, match="2\\^N"):
        O.mps_from_state(np.ones(6))


def test_mps_from_product_state_closed_form():
    rng = np.random.default_rng(6)
    bs = [crand(rng, 2) for _ in range(4)]
    psi = bs[0]
    for b in bs[1:]:
        psi = np.kron(b, psi)  # wire 1 is the least significant bit
    ts = O.mps_from_state(psi)
    # no truncation (src/mps.jl:55-89): bond b keeps min(2^b, 2^(n-b))
    # values even though a product state has Schmidt rank 1
    assert [t.shape[-1] for t in ts[:-1]] == [2, 4, 2]
    assert np.allclose(O.mps_to_state(ts), psi, atol=1e-12)


def test_mpo_round_trip_and_operator_schmidt():
    rng = np.random.default_rng(7)
    U = np.linalg.qr(crand(rng, 4, 4))[0]
    ts, spectra = O.mpo_from_matrix(U)
    T = np.einsum("rcx,xRC->rcRC", ts[0], ts[1])
    ref = U.reshape((2, 2, 2, 2), order="F").transpose(0, 2, 1, 3)
    assert np.allclose(T, ref, atol=1e-13)
    assert abs(np.sum(spectra[0] ** 2) - 4.0) < 1e-12
    swap = np.eye(4)[[0, 2, 1, 3]]
    assert np.allclose(O.operator_schmidt(swap), [1, 1, 1, 1], atol=1e-13)
    cnot = np.eye(4)[[0, 1, 3, 2]]
    cs = O.operator_schmidt(cnot)
    assert np.allclose(cs[:2], [np.sqrt(2), np.sqrt(2)], atol=1e-13)
    assert np.allclose(cs[2:], 0, atol=1e-13)
    # 3-qubit chain reproduces the operator
    U3 = np.linalg.qr(crand(rng, 8, 8))[0]
    ts, _ = O.mpo_from_matrix(U3)
    T = np.einsum("abx,xcdy,yef->abcdef", *ts)
    ref = U3.reshape((2,) * 6, order="F").transpose(0, 3, 1, 4, 2, 5)
    assert np.allclose(T, ref, atol=1e-12)


def test_apply_mpo_dense_wiring():
    # CNOT with iwire = (targ, cntrl), gate bit k-1 <-> iwire[k]
    # (test/test_mpo.jl:62-86)
    cnot = np.eye(4)[[0, 1, 3, 2]]
    n = 5
    for targ, cntrl in [(1, 3), (4, 2), (5, 1)]:
        for idx in range(2 ** n):
            psi = np.zeros(2 ** n, dtype=complex)
            psi[idx] = 1
            out = O.apply_mpo_dense(psi, cnot, (targ, cntrl))
            c = (idx >> (cntrl - 1)) & 1
            exp_idx = idx ^ ((1 << (targ - 1)) if c else 0)
            assert abs(out[exp_idx] - 1) < 1e-15
    # a non-symmetric unitary on wires (1,2) of 2 qubits is a plain matvec;
    # on reversed wires it is conjugated by the bit swap
    rng = np.random.default_rng(8)
    U = np.linalg.qr(crand(rng, 4, 4))[0]
    psi = crand(rng, 4)
    assert np.allclose(O.apply_mpo_dense(psi, U, (1, 2)), U @ psi)
    P = np.eye(4)[[0, 2, 1, 3]]
    assert np.allclose(O.apply_mpo_dense(psi, U, (2, 1)), P @ U @ P @ psi)
    assert not np.allclose(O.apply_mpo_dense(psi, U, (1, 2)), U.T @ psi)


# -- TEBD in the right-canonical gauge ----------------------------------------
def run_both(n, depth, seed, cutoff=0.0, chimax=0):
    layers = C.brickwall_layers(n, depth, seed)
    st = O.TebdState(n)
    sv = O.StateVector(n)
    for singles, pairs in layers:
        for w, u in singles:
            st.apply_1q(w, u)
            sv.apply_1q(w, u)
        for w, g in pairs:
            st.apply_2q(w, g, cutoff, chimax)
            sv.apply_2q(w, g)
    return st, sv


def test_tebd_untruncated_equals_statevector():
    st, sv = run_both(8, 7, 11)
    assert np.allclose(st.to_statevector(), sv.psi, atol=1e-13)
    bits = C.random_bitstrings(8, 32, 1)
    assert np.allclose(st.amplitudes(bits), sv.amplitudes(bits), atol=1e-13)
    assert np.allclose(st.expect_z(), sv.expect_z(), atol=1e-13)
    assert abs(st.norm2() - 1) < 1e-13
    # bond spectra are the Schmidt values of the dense state
    for b in range(1, 8):
        m = sv.psi.reshape((2 ** b, 2 ** (8 - b)), order="F")
        s = np.linalg.svd(m, compute_uv=False)
        k = st.lam[b].size
        assert np.allclose(st.lam[b], s[:k], atol=1e-12)
        assert np.all(s[k:] < 1e-12)


def test_tebd_planted_transpose_is_caught():
    # applying g^T instead of g (the non-decomposed tensor_circuit! path,
    # src/tensor_circuit.jl:44-51) must fail the dense comparison
    layers = C.brickwall_layers(6, 4, 12)
    st = O.TebdState(6)
    sv = O.StateVector(6)
    for singles, pairs in layers:
        for w, u in singles:
            st.apply_1q(w, u)
            sv.apply_1q(w, u)
        for w, g in pairs:
            st.apply_2q(w, g.T, 0.0, 0)
            sv.apply_2q(w, g)
    assert not np.allclose(st.to_statevector(), sv.psi, atol=1e-6)


def test_tebd_truncation_fidelity_bound():
    # with a chi cap the fidelity loss is bounded by the summed discarded
    # weights (each truncation is an orthogonal projection of a normalised
    # state in canonical form)
    n = 10
    layers = C.brickwall_layers(n, 10, 13)
    st = O.TebdState(n)
    sv = O.StateVector(n)
    disc_sum = 0.0
    for singles, pairs in layers:
        for w, u in singles:
            st.apply_1q(w, u)
            sv.apply_1q(w, u)
        for w, g in pairs:
            _, d, _ = st.apply_2q(w, g, 1e-8, 8)
            sv.apply_2q(w, g)
            disc_sum += d
    assert max(st.bond_dims()) == 8
    f = abs(np.vdot(st.to_statevector(), sv.psi)) ** 2 / st.norm2()
    assert 1 - f <= 2 * disc_sum + 1e-10
    assert disc_sum > 1e-6


def test_product_state_observables():
    n = 4
    st = O.TebdState(n)
    # |0101> via X on wires 2 and 4
    X = np.array([[0, 1], [1, 0]], dtype=complex)
    st.apply_1q(2, X)
    st.apply_1q(4, X)
    assert np.allclose(st.expect_z(), [1, -1, 1, -1])
    bits = np.array([[0, 1, 0, 1], [1, 1, 0, 1]], dtype=np.uint8)
    assert np.allclose(st.amplitudes(bits), [1, 0])
    other = st.copy()
    assert abs(st.overlap(other) - 1) < 1e-15


# -- generator ---------------------------------------------------------------
def test_generator_shapes_and_unitarity():
    layers = C.brickwall_layers(40, 20, 2)
    assert C.count_2q(layers) == 390
    l0s, l0p = layers[0]
    l1s, l1p = layers[1]
    assert [w for w, _ in l0p] == list(range(1, 40, 2)) and not l0s
    assert [w for w, _ in l1p] == list(range(2, 39, 2))
    assert [w for w, _ in l1s] == [1, 40]
    for _, g in l0p[:5]:
        assert np.allclose(g.conj().T @ g, np.eye(4), atol=1e-13)
    again = C.brickwall_layers(40, 20, 2)
    assert np.array_equal(again[7][1][3][1], layers[7][1][3][1])
    rng = np.random.default_rng(0)
    U = C.haar_unitary(rng, 4)
    q, r = np.linalg.qr(U)
    assert np.allclose(U.conj().T @ U, np.eye(4), atol=1e-14)


def test_fused_gate_equals_sequential_application():
    rng = np.random.default_rng(9)
    u1 = C.haar_unitary(rng, 2)
    u2 = C.haar_unitary(rng, 2)
    g = C.haar_unitary(rng, 4)
    psi = crand(rng, 8)
    a = O.apply_mpo_dense(psi, u1, (2,))
    a = O.apply_mpo_dense(a, u2, (3,))
    a = O.apply_mpo_dense(a, g, (2, 3))
    b = O.apply_mpo_dense(psi, g @ np.kron(u2, u1), (2, 3))
    assert np.allclose(a, b, atol=1e-14)


# -- MPO x MPS oracle (tfim_mpo, expect_mpo, canonicalize, zipup) ----------
def test_tfim_mpo_equals_kron_hamiltonian():
    for n in (2, 3, 5, 8):
        h = np.random.default_rng(n).uniform(0.5, 1.5, n)
        Hm = O.mpo_dense(O.tfim_mpo(n, 1.0, h))
        Hk = O.tfim_dense(n, 1.0, h)
        assert np.abs(Hm - Hk).max() == 0.0
        assert np.allclose(Hk, Hk.conj().T)
    # closed form: at h = 0 the all-zero state has energy -J (n-1)
    H = O.tfim_dense(4, 0.7, np.zeros(4))
    assert np.isclose(H[0, 0].real, -0.7 * 3)


def test_expect_mpo_matches_dense_and_identity():
    n = 7
    h = np.random.default_rng(2).uniform(0.5, 1.5, n)
    W = O.tfim_mpo(n, 1.0, h)
    Bs = O.random_mps(n, 8, 11)
    psi = O.mps_dense(Bs)
    assert abs(np.linalg.norm(psi) - 1) < 1e-12
    ref = np.vdot(psi, O.tfim_dense(n, 1.0, h) @ psi)
    assert abs(O.expect_mpo(Bs, W) - ref) < 1e-12
    eye = [np.eye(2, dtype=complex).reshape((1, 2, 2, 1))] * n
    assert abs(O.expect_mpo(Bs, eye) - 1.0) < 1e-12


def test_canonicalize_right_form_and_schmidt_values():
    n = 8
    Bs = O.random_mps(n, 16, 3)
    psi = O.mps_dense(Bs)
    B, lam, nrm = O.canonicalize(Bs)
    assert np.abs(O.mps_dense(B) * nrm - psi).max() < 1e-13
    for j in range(n):
        m = B[j].reshape((B[j].shape[0], -1), order="F")
        assert np.abs(m @ m.conj().T - np.eye(m.shape[0])).max() < 1e-13
    for cut in range(1, n):
        s = np.linalg.svd(psi.reshape((2 ** cut, -1), order="F"),
                          compute_uv=False)
        s = s[s > 1e-12]
        assert np.allclose(lam[cut], s / np.linalg.norm(s), atol=1e-12)


def test_zipup_exact_and_truncated():
    n = 8
    h = np.random.default_rng(4).uniform(0.5, 1.5, n)
    W = O.tfim_mpo(n, 1.0, h)
    Bs = O.random_mps(n, 8, 5)
    psi = O.mps_dense(Bs)
    Hpsi = O.tfim_dense(n, 1.0, h) @ psi
    B, lam, nrm = O.zipup(Bs, W, 0.0, 0)
    assert np.abs(O.mps_dense(B) * nrm - Hpsi).max() < 1e-12
    # enough bond dimension for H psi (3 x 8): exact again
    B, lam, nrm = O.zipup(Bs, W, 0.0, 24)
    assert np.abs(O.mps_dense(B) * nrm - Hpsi).max() < 1e-12
    # truncated: normalised, norm below |H psi|, caps respected, and the
    # overlap with the exact H psi grows with the cap
    fids = []
    for chi in (4, 6, 8, 12):
        B, lam, nrm = O.zipup(Bs, W, 1e-3, chi)
        phi = O.mps_dense(B)
        assert abs(np.linalg.norm(phi) - 1) < 1e-12
        assert nrm <= np.linalg.norm(Hpsi) * (1 + 1e-12)
        assert max(b.shape[2] for b in B) <= chi
        fids.append(abs(np.vdot(phi, Hpsi)) / np.linalg.norm(Hpsi))
    assert all(a < b for a, b in zip(fids, fids[1:])) and fids[-1] < 1


def test_expect_mpo_blas_equals_einsum_form_and_product_mpo():
    import sys, os
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..",
                                    "qaintensor.jl_amd"))
    from qaintensor import hamiltonians as Hm
    n = 9
    h = Hm.tfim_fields(n, 3)
    W = O.tfim_mpo(n, 1.0, h)
    Wp = Hm.tfim_mpo(n, 1.0, h)
    assert all(np.array_equal(a, b) for a, b in zip(W, Wp))
    Bs = O.left_canonical(Hm.random_mps_tensors(n, 16, 2))
    assert abs(O.expect_mpo_blas(Bs, W) - O.expect_mpo(Bs, W)) < 1e-12


def test_mlp_oracle_gradient_matches_finite_differences():
    nin, w = 69, 16
    rng = np.random.default_rng(0)
    p = rng.standard_normal(O.mlp_nparams(nin, w)) * 0.1
    X = np.abs(rng.standard_normal((nin, 7))) + 1e-3
    T = rng.standard_normal(7)
    _, g = O.mlp_loss_grad(p, X, T, nin, w)
    eps = 1e-6
    for i in rng.integers(0, p.size, 25):
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        lp, _ = O.mlp_loss_grad(pp, X, T, nin, w)
        lm, _ = O.mlp_loss_grad(pm, X, T, nin, w)
        assert abs(g[i] - (lp - lm) / (2 * eps)) <= 1e-7 + 1e-5 * abs(g[i])
    # forward agrees with the loss: loss = mean (y - t)^2
    y = O.mlp_forward(p, X, nin, w)[0]
    l, _ = O.mlp_loss_grad(p, X, T, nin, w)
    assert abs(l - np.mean((y - T) ** 2)) < 1e-12


def test_mlp_oracle_forward_matches_torch():
    """The oracle's forward pass against torch's own layers (nn.Linear,
    exact-erf GELU) in float64, built from the documented flat layout; the
    first 64 inputs enter as log10(max(x, 1e-16)) / 8."""
    import torch
    nin, w, B = 72, 16, 9
    rng = np.random.default_rng(3)
    p = rng.standard_normal(O.mlp_nparams(nin, w)) * 0.3
    X = np.abs(rng.standard_normal((nin, B)))
    X[5, 0] = 0.0  # a zero spectrum entry takes the 1e-16 floor
    o = 0
    W1 = p[o:o + w * nin].reshape((w, nin), order="F"); o += w * nin
    b1 = p[o:o + w]; o += w
    W2 = p[o:o + w * w].reshape((w, w), order="F"); o += w * w
    b2 = p[o:o + w]; o += w
    W3 = p[o:o + w].reshape((1, w)); o += w
    b3 = p[o:o + 1]
    net = torch.nn.Sequential(torch.nn.Linear(nin, w), torch.nn.GELU(),
                              torch.nn.Linear(w, w), torch.nn.GELU(),
                              torch.nn.Linear(w, 1)).double()
    with torch.no_grad():
        for lin, (Wt, bt) in zip([net[0], net[2], net[4]],
                                 [(W1, b1), (W2, b2), (W3, b3)]):
            lin.weight.copy_(torch.from_numpy(np.ascontiguousarray(Wt)))
            lin.bias.copy_(torch.from_numpy(np.ascontiguousarray(bt)))
        inp = X.copy()
        inp[:64] = np.log10(np.maximum(inp[:64], 1e-16)) / 8.0
        ref = net(torch.from_numpy(inp.T.copy())).numpy().T
    got = O.mlp_forward(p, X, nin, w)
    assert np.max(np.abs(got - ref)) <= 1e-12


def _product_tebd(bvecs):
    """TebdState holding kron(b_1, ..., b_N) (b_1 the most significant
    factor, so site w holds b_{N-w+1}), unnormalised product tensors."""
    n = len(bvecs)
    st = O.TebdState(n)
    for w in range(1, n + 1):
        st.B[w - 1] = np.asarray(bvecs[n - w], dtype=np.complex128).reshape(
            (1, 2, 1))
    return st


def test_switch_reference_case():
    """test/test_mps.jl:198-221: switch!(mps, [(2,3), (4,6), (4,5)]) on
    kron(b...) gives kron(b[order]...), order = [1, 3, 2, 5, 6, 4]."""
    rng = np.random.default_rng(11)
    N = 6
    b = [rng.standard_normal(2) + 1j * rng.standard_normal(2)
         for _ in range(N)]
    order = [1, 3, 2, 5, 6, 4]
    psi_perm = b[order[0] - 1]
    for k in order[1:]:
        psi_perm = np.kron(psi_perm, b[k - 1])
    st = _product_tebd(b)
    for (i, j) in [(2, 3), (4, 6), (4, 5)]:
        st.switch(i, j)
    v = st.to_statevector()
    fid = abs(np.vdot(v, psi_perm)) / (np.linalg.norm(v) *
                                      np.linalg.norm(psi_perm))
    assert abs(1 - fid) <= 1e-12
    with pytest.raises(ValueError, match="must be positive"):
        st.switch(1, -2)
    with pytest.raises(ValueError, match="less than or equal"):
        st.switch(7, 4)


def test_switch_exchanges_dense_bits():
    """switch!(i, j) on an entangled state equals exchanging bits n-i and
    n-j of the dense vector (wire w <-> site n-w+1 <-> bit n-w)."""
    n = 7
    st = O.TebdState(n)
    O.run_layers(st, C.brickwall_layers(n, 5, 4), 0.0, 0)
    for (i, j) in [(1, 7), (3, 5), (6, 2)]:
        v0 = st.to_statevector()
        st.switch(i, j)
        v = st.to_statevector()
        idx = np.arange(2 ** n)
        a, c = n - i, n - j
        bi, bj = (idx >> a) & 1, (idx >> c) & 1
        src = idx ^ ((bi ^ bj) << a) ^ ((bi ^ bj) << c)
        ref = v0[src]
        assert np.linalg.norm(v / np.linalg.norm(v) - ref /
                              np.linalg.norm(ref)) <= 1e-12


@pytest.mark.parametrize("i,j", [(2, 6), (6, 2), (1, 8), (8, 1), (4, 5)])
def test_apply_2q_long_equals_dense(i, j):
    """SWAP routing without truncation equals the gate applied on wires
    (i, j) of the dense vector, for both orders of the wires."""
    n = 8
    rng = np.random.default_rng(i * 10 + j)
    layers = C.brickwall_layers(n, 4, i + j)
    st = O.TebdState(n)
    O.run_layers(st, layers, 0.0, 0)
    sv = O.StateVector(n)
    sv.psi = st.to_statevector()
    g = C.haar_unitary(rng, 4)
    st.apply_2q_long(i, j, g, 0.0, 0)
    sv.apply_2q_any(i, j, g)
    v = st.to_statevector()
    assert np.linalg.norm(v - sv.psi) <= 1e-12 * np.linalg.norm(sv.psi)


def test_ncon_ordered_matches_einsum_and_long_chains():
    """ncon_ordered (pairwise tensordot in a label order) against the
    einsum ncon on a network with a trace, parallel legs and an unconnected
    vector, under three orders; and a 60-matrix chain (beyond einsum's
    alphabet) against the plain matrix product."""
    rng = np.random.default_rng(8)

    def cr(*d):
        return rng.standard_normal(d) + 1j * rng.standard_normal(d)
    ts = [cr(3, 4, 5, 2, 2), cr(4, 6), cr(6, 7, 8), cr(7, 3), cr(9),
          cr(5, 8, 2)]
    labels = [[4, 1, 6, -2, 5], [1, 2], [2, 3, 7], [3, 4], [-1],
              [6, 7, 5]]
    ref = O.ncon(ts, labels)
    for order in ([1, 2, 3, 4, 5, 6, 7], [7, 6, 5, 4, 3, 2, 1],
                  [3, 6, 1, 7, 2, 5, 4]):
        got = O.ncon_ordered(ts, labels, order)
        assert got.shape == ref.shape
        assert np.allclose(got, ref, rtol=1e-12, atol=1e-12 * abs(ref).max())
    tr = cr(3, 5, 3)
    assert np.allclose(O.ncon_ordered([tr], [[1, -1, 1]]),
                       np.einsum("aba->b", tr))
    mats = [cr(2, 2) for _ in range(60)]
    labels = [[-1, 1]] + [[i, i + 1] for i in range(1, 59)] + [[59, -2]]
    ref = mats[0]
    for m in mats[1:]:
        ref = ref @ m
    got = O.ncon_ordered(mats, labels)
    assert np.allclose(got, ref, rtol=1e-10, atol=1e-10 * abs(ref).max())
