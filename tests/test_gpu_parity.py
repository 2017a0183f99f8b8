"""GPU parity: libqtn (through the C ABI) against the FP64 oracle and the
committed golden fixtures.

Tolerances (complex FP64 end to end):
  - amplitudes: ||a_gpu - a_ref|| <= 1e-10 ||a_ref||
  - <Z_i>, bond spectra, truncated singular values: |diff| <= 1e-10
    (spectra are normalised, so this is relative to the largest value)
  - kept ranks: identical
"""
import os

import numpy as np
import pytest

from oracle import qtn_oracle as O
from qaintensor import DeviceMPS, circuits as C, ops
from qaintensor._lib import QtnError, QTN_ERR_ARG, QTN_ERR_NOTFINITE, \
    QTN_ERR_TRUNC

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL = 1e-10


def crand(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


# -- contraction ------------------------------------------------------------
@pytest.mark.parametrize("m,k,n", [(1, 1, 1), (5, 7, 3), (64, 64, 64),
                                   (100, 37, 130), (33, 1, 65),
                                   (512, 256, 512)])
def test_contract_pair_gemm_shapes(gpu_ctx, m, k, n):
    rng = np.random.default_rng(m * 7 + n)
    A, B = crand(rng, m, k), crand(rng, k, n)
    C_ = ops.contract_pair(A, B, [2], [1])
    ref = A @ B
    assert np.abs(C_ - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max()) * k


def test_contract_pair_transposed_operands_and_outperm(gpu_ctx):
    rng = np.random.default_rng(1)
    A = crand(rng, 6, 9, 4)
    B = crand(rng, 4, 5, 9)
    for la, lb, op in [([2, 3], [3, 1], None), ([1], [3], [3, 1, 2, 4, 5]),
                       ([3], [1], [4, 2, 1, 3]), ([2], [3], None)]:
        fa = A.ndim + B.ndim - 2 * len(la)
        if op is not None and len(op) != fa:
            continue
        got = ops.contract_pair(A, B, la, lb, op)
        ref = O.contract_pair(A, B, la, lb, op)
        assert got.shape == ref.shape
        assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def test_contract_pair_dimension_mismatch_message(gpu_ctx):
    A = np.ones((2, 3), dtype=complex)
    B = np.ones((4, 2), dtype=complex)
    with pytest.raises(QtnError, match="Dimensions of contraction legs"):
        ops.contract_pair(A, B, [2], [1])


@pytest.mark.parametrize("shape,perm", [((3, 4, 5), (3, 1, 2)),
                                        ((2, 2, 2, 2), (1, 3, 2, 4)),
                                        ((64, 2, 2, 64), (1, 3, 2, 4)),
                                        ((70, 1, 33), (3, 2, 1)),
                                        ((5, 40, 7, 3), (2, 4, 1, 3)),
                                        ((128, 96), (2, 1))])
def test_permute_matches_transpose(gpu_ctx, shape, perm):
    rng = np.random.default_rng(len(shape))
    A = crand(rng, *shape)
    got = ops.permute(A, perm)
    assert np.array_equal(got, np.transpose(A, [p - 1 for p in perm]))


# -- truncated SVD ----------------------------------------------------------
@pytest.mark.parametrize("m,n", [(1, 1), (4, 4), (8, 3), (3, 8), (64, 64),
                                 (100, 70), (65, 129), (256, 256),
                                 (512, 512)])
def test_svd_values_and_factors(gpu_ctx, m, n):
    rng = np.random.default_rng(m + 1000 * n)
    A = crand(rng, m, n)
    U, S, Vh, disc = ops.svd_trunc(A, 0.0, False, 0)
    ref = np.linalg.svd(A, compute_uv=False)
    k = S.size
    assert k == min(m, n)
    assert np.abs(S - ref).max() <= TOL * ref[0]
    assert np.abs((U * S) @ Vh - A).max() <= TOL * np.abs(A).max()
    assert np.abs(U.conj().T @ U - np.eye(k)).max() <= TOL
    assert np.abs(Vh @ Vh.conj().T - np.eye(k)).max() <= TOL


def test_svd_reference_kat_spectrum(gpu_ctx):
    # test/test_svd.jl:53-83 spectrum through the GPU: er = 1e-10 -> k = 24
    rng = np.random.default_rng(4)
    s = np.exp(-np.arange(100.0))
    Q1 = np.linalg.qr(crand(rng, 100, 100))[0]
    Q2 = np.linalg.qr(crand(rng, 100, 100))[0]
    A = (Q1 * s) @ Q2
    U, S, Vh, disc = ops.svd_trunc(A, 1e-10, False, 0)
    assert S.size == 24
    ref = np.linalg.svd(A, compute_uv=False)
    assert np.abs(S - ref[:24]).max() <= 1e-12
    assert np.abs(S - s[:24]).max() <= 1e-12
    assert abs(np.sqrt(disc) - np.linalg.norm(s[24:])) <= 1e-15


def test_svd_rank_deficient_and_cap(gpu_ctx):
    rng = np.random.default_rng(5)
    A = crand(rng, 200, 37) @ crand(rng, 37, 150)
    U, S, Vh, disc = ops.svd_trunc(A, 1e-10, True, 0)
    assert S.size == 37
    U, S, Vh, disc = ops.svd_trunc(A, 1e-10, True, 20)
    ref = np.linalg.svd(A, compute_uv=False)
    assert S.size == 20
    assert abs(disc - np.sum(ref[20:] ** 2)) <= 1e-10 * np.sum(ref ** 2)


@pytest.mark.parametrize("m,n", [(1300, 700), (700, 1300), (640, 640),
                                 (3072, 2048)])
def test_svd_qr_preconditioned_shapes(gpu_ctx, monkeypatch, m, n):
    """With QTN_SVD_PRECOND=1, splits with min(m, n) >= 512 run the Jacobi
    sweeps on R2^H of the two-QR preconditioner (sorted columns,
    X P = Q1 R1, R1^H = Q2 R2); the factors come back through Q1, Q2 and
    the column permutation.  Values, reconstruction and orthonormality
    against LAPACK at 1e-10, untruncated and under a cap."""
    monkeypatch.setenv("QTN_SVD_PRECOND", "1")
    rng = np.random.default_rng(m + 7 * n)
    A = crand(rng, m, n)
    ref = np.linalg.svd(A, compute_uv=False)
    U, S, Vh, disc = ops.svd_trunc(A, 0.0, False, 0)
    k = min(m, n)
    assert S.size == k
    assert np.abs(S - ref).max() <= TOL * ref[0]
    assert np.linalg.norm((U * S) @ Vh - A) <= TOL * np.linalg.norm(A)
    assert np.abs(U.conj().T @ U - np.eye(k)).max() <= TOL
    assert np.abs(Vh @ Vh.conj().T - np.eye(k)).max() <= TOL
    cap = k // 3
    U, S, Vh, disc = ops.svd_trunc(A, 1e-10, True, cap)
    assert S.size == cap
    assert np.abs(S - ref[:cap]).max() <= TOL * ref[0]
    assert abs(disc - np.sum(ref[cap:] ** 2)) <= TOL * np.sum(ref ** 2)
    # the best rank-cap approximation: its error is the discarded weight
    err = np.linalg.norm((U * S) @ Vh - A) ** 2
    assert abs(err - disc) <= 1e-9 * np.sum(ref ** 2)


def test_svd_qr_preconditioned_rank_deficient_graded(gpu_ctx, monkeypatch):
    """Rank 60 of 600 columns with a graded spectrum (1 .. 1e-9), through
    the QR preconditioner: the relative cutoff keeps exactly the nonzero
    values, to LAPACK's."""
    monkeypatch.setenv("QTN_SVD_PRECOND", "1")
    rng = np.random.default_rng(21)
    s = np.logspace(0, -9, 60)
    Q1 = np.linalg.qr(crand(rng, 900, 60))[0]
    Q2 = np.linalg.qr(crand(rng, 600, 60))[0]
    A = (Q1 * s) @ Q2.conj().T
    for B in (A, A.T.copy()):
        U, S, Vh, disc = ops.svd_trunc(B, 1e-12, True, 0)
        assert S.size == 60
        assert np.abs(S - s).max() <= 1e-12
        assert np.linalg.norm((U * S) @ Vh - B) <= 1e-12
        assert np.abs(U.conj().T @ U - np.eye(60)).max() <= TOL
        assert np.abs(Vh @ Vh.conj().T - np.eye(60)).max() <= TOL


@pytest.mark.parametrize("scale", [1e100, 1e-100, 1e160, 1e-160])
def test_svd_qr_preconditioned_magnitudes(gpu_ctx, monkeypatch, scale):
    """Entries within 2^+-450 go through the preconditioner unscaled,
    beyond it the rescaling engine takes the split: both exact relative to
    the scale."""
    monkeypatch.setenv("QTN_SVD_PRECOND", "1")
    rng = np.random.default_rng(13)
    A = crand(rng, 800, 600)
    ref = np.linalg.svd(A, compute_uv=False)
    U, S, Vh, _ = ops.svd_trunc(A * scale, 1e-10, True, 0)
    assert np.abs(S / scale - ref[:S.size]).max() <= 1e-12 * ref[0]
    R = (U * (S / scale)) @ Vh
    assert np.linalg.norm(R - A) <= 1e-11 * np.linalg.norm(A)


def test_svd_error_paths(gpu_ctx, monkeypatch):
    A = np.ones((4, 4), dtype=complex)
    with pytest.raises(QtnError, match="Error must be positive") as e:
        ops.svd_trunc(A, -1.0)
    assert e.value.status == QTN_ERR_ARG
    with pytest.raises(QtnError) as e:
        ops.svd_trunc(np.zeros((4, 4), dtype=complex), 0.0)
    assert e.value.status == QTN_ERR_TRUNC
    B = A.copy()
    B[1, 2] = np.nan
    with pytest.raises(QtnError) as e:
        ops.svd_trunc(B, 0.0)
    assert e.value.status == QTN_ERR_NOTFINITE
    # a split large enough for the QR preconditioner: zero and non-finite
    # inputs report as on the small ones
    monkeypatch.setenv("QTN_SVD_PRECOND", "1")
    with pytest.raises(QtnError) as e:
        ops.svd_trunc(np.zeros((700, 600), dtype=complex), 0.0)
    assert e.value.status == QTN_ERR_TRUNC
    for bad in (np.nan, np.inf):
        C = np.ones((700, 600), dtype=complex)
        C[500, 3] = bad
        with pytest.raises(QtnError) as e:
            ops.svd_trunc(C, 0.0)
        assert e.value.status == QTN_ERR_NOTFINITE


# -- two-site update / MPS ----------------------------------------------------
def load_gold(name):
    return dict(np.load(os.path.join(GOLD, name), allow_pickle=False))


@pytest.mark.parametrize("batched", [False, True])
@pytest.mark.parametrize("fixture", ["config1_chi32.npz",
                                     "config1_chi64.npz"])
def test_config1_against_golden(gpu_ctx, fixture, batched):
    g = load_gold(fixture)
    n, chimax, cutoff = int(g["n"]), int(g["chimax"]), float(g["cutoff"])
    layers = C.brickwall_layers(n, int(g["depth"]), int(g["gate_seed"]))
    d = DeviceMPS(n, chimax)
    rec = []
    d.run_layers(layers, cutoff, chimax, batched=batched, record=rec)
    assert [r[1] for r in rec] == list(g["ranks"])
    a = d.amplitudes(g["bits"])
    ref = g["amplitudes"]
    assert np.linalg.norm(a - ref) <= TOL * np.linalg.norm(ref)
    assert np.abs(d.expect_z() - g["expect_z"]).max() <= TOL
    assert list(d.bond_dims()) == list(g["bond_dims"])
    for b in range(1, n):
        k = int(g["bond_dims"][b])
        assert np.abs(d.bond_spectrum(b) - g["bond_spectra"][b, :k]).max() \
            <= TOL
    assert abs(d.norm2() - 1.0) <= TOL
    if "dense_amplitudes" in g:
        da = g["dense_amplitudes"]
        assert np.linalg.norm(a - da) <= TOL * np.linalg.norm(da)


def test_two_site_spectrum_per_gate(gpu_ctx):
    g = load_gold("config1_chi32.npz")
    n = int(g["n"])
    layers = C.brickwall_layers(n, int(g["depth"]), int(g["gate_seed"]))
    d = DeviceMPS(n, 32)
    i = 0
    for singles, pairs in layers:
        for w, u in singles:
            d.apply_1q(w, u)
        for w, gate in pairs:
            k, disc, S = d.apply_2q(w, gate, float(g["cutoff"]), 32,
                                    want_spectrum=True)
            ref = g["spectra"][i]
            ref = ref[~np.isnan(ref)]
            assert S.size == ref.size
            assert np.abs(S - ref).max() <= TOL * max(1.0, ref[0])
            assert k == g["ranks"][i]
            assert abs(disc - g["disc"][i]) <= TOL
            i += 1


def test_untruncated_matches_dense_statevector(gpu_ctx):
    n = 10
    layers = C.brickwall_layers(n, 9, 21)
    sv = O.StateVector(n)
    for singles, pairs in layers:
        for w, u in singles:
            sv.apply_1q(w, u)
        for w, gate in pairs:
            sv.apply_2q(w, gate)
    d = DeviceMPS(n, 32)
    d.run_layers(layers, 0.0, 32)
    bits = C.random_bitstrings(n, 128, 3)
    a = d.amplitudes(bits)
    ref = sv.amplitudes(bits)
    assert np.linalg.norm(a - ref) <= TOL * np.linalg.norm(ref)
    assert np.abs(d.expect_z() - sv.expect_z()).max() <= TOL


def test_config2_leading_layers_against_oracle(gpu_ctx):
    # 40 qubits, chi_max 256, cutoff 1e-8: the first 9 layers reach chi=256
    n, chimax, cutoff = 40, 256, 1e-8
    layers = C.brickwall_layers(n, 20, 2)[:9]
    st = O.TebdState(n)
    rec_o = []
    O.run_layers(st, layers, cutoff, chimax, rec_o)
    d = DeviceMPS(n, chimax)
    rec = []
    d.run_layers(layers, cutoff, chimax, record=rec)
    assert [r[1] for r in rec] == [r[1] for r in rec_o]
    assert max(d.bond_dims()) == 256
    bits = C.random_bitstrings(n, 64, 12)
    a = d.amplitudes(bits)
    ref = st.amplitudes(bits)
    assert np.linalg.norm(a - ref) <= TOL * np.linalg.norm(ref)
    assert np.abs(d.expect_z() - st.expect_z()).max() <= TOL
    for b in (10, 20, 30):
        k = st.lam[b].size
        assert np.abs(d.bond_spectrum(b) - st.lam[b]).max() <= TOL


def test_config2_full_circuit(gpu_ctx):
    # the full 40-qubit, depth-20, chi 256 circuit (390 gates).  Truncation
    # is heavy from layer 10 on (discarded weight up to ~3e-2 per gate), so
    # the gauge is only approximately canonical and <psi|psi> drifts from 1
    # (to 1.1126 at the end); the oracle runs the same algebra and drifts
    # identically, so everything is compared with it.
    n, chimax, cutoff = 40, 256, 1e-8
    layers = C.brickwall_layers(n, 20, 2)
    d = DeviceMPS(n, chimax)
    rec = []
    d.run_layers(layers, cutoff, chimax, record=rec)
    st = O.TebdState(n)
    rec_o = []
    O.run_layers(st, layers, cutoff, chimax, rec_o)
    assert len(rec) == 390
    assert [r[1] for r in rec] == [r[1] for r in rec_o]
    assert max(abs(r[2] - q[2]) for r, q in zip(rec, rec_o)) <= TOL
    bits = C.random_bitstrings(n, 64, 12)
    a, ref = d.amplitudes(bits), st.amplitudes(bits)
    assert np.linalg.norm(a - ref) <= TOL * np.linalg.norm(ref)
    assert np.abs(d.expect_z() - st.expect_z()).max() <= TOL
    assert abs(d.norm2() - st.norm2()) <= TOL * st.norm2()
    # size-independent properties: normalised, sorted bond spectra
    for b in range(1, n):
        lam = d.bond_spectrum(b)
        assert abs(np.sum(lam ** 2) - 1.0) <= 1e-12
        assert np.all(np.diff(lam) <= 0)
    # per-gate and batched runs produce the same state
    e = DeviceMPS(n, chimax)
    e.run_layers(layers, cutoff, chimax, batched=False)
    f = abs(d.overlap(e)) ** 2 / (d.norm2() * e.norm2())
    assert abs(f - 1.0) <= 1e-12


def test_qr_preconditioned_jacobi_against_oracle(gpu_ctx, monkeypatch):
    # the optional QR-preconditioned Jacobi (sorted columns, X = Q1 R1,
    # R1^H = Q2 R2, Jacobi on R2^H) must reproduce the oracle like the
    # default path
    monkeypatch.setenv("QTN_JACOBI_PRECOND", "1")
    n, chimax, cutoff = 40, 256, 1e-8
    layers = C.brickwall_layers(n, 20, 2)[:9]
    st = O.TebdState(n)
    rec_o = []
    O.run_layers(st, layers, cutoff, chimax, rec_o)
    for batched in (True, False):
        d = DeviceMPS(n, chimax)
        rec = []
        d.run_layers(layers, cutoff, chimax, batched=batched, record=rec)
        assert [r[1] for r in rec] == [r[1] for r in rec_o]
        bits = C.random_bitstrings(n, 64, 12)
        a, ref = d.amplitudes(bits), st.amplitudes(bits)
        assert np.linalg.norm(a - ref) <= TOL * np.linalg.norm(ref)
        assert np.abs(d.expect_z() - st.expect_z()).max() <= TOL


@pytest.mark.parametrize("scale", [1e160, 1e-160])
def test_svd_extreme_magnitudes(gpu_ctx, scale):
    """Squared Gram entries of X = 1e160 * A would overflow: the engine
    sweeps a power-of-two rescaled copy, so the singular values come out
    exact relative to the scale (regression for a 64-site MPS whose
    unnormalised carry reached 1e92)."""
    rng = np.random.default_rng(12)
    A = crand(rng, 300, 200)
    ref = np.linalg.svd(A, compute_uv=False)
    S = ops.svd_values(A * scale)
    assert np.abs(S / scale - ref).max() <= 1e-12 * ref[0]
    U, S2, Vh, _ = ops.svd_trunc(A * scale, 1e-10, True, 0)
    assert np.abs(S2 / scale - ref[:S2.size]).max() <= 1e-12 * ref[0]
    R = (U * (S2 / scale)) @ Vh
    assert np.linalg.norm(R - A) <= 1e-11 * np.linalg.norm(A)
