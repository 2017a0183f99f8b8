"""The chosen form of the two-site update (QTN_FORM_CHOSEN): Jacobi sweeps
in complex FP32 on MFMA, V_k re-orthogonalised by Newton-Schulz in FP64,
B_i and lambda_i recomputed in FP64.  Its bar is state infidelity
<= 1e-6 against the FP64 path (which itself matches the oracle to 1e-10);
right-canonical site tensors must stay orthonormal to FP64 rounding.
"""
import numpy as np
import pytest

from oracle import qtn_oracle as O
from qaintensor import DeviceMPS, FORM_CHOSEN, circuits as C

pytestmark = pytest.mark.gpu


def infidelity(a, b):
    return 1.0 - abs(a.overlap(b)) ** 2 / (a.norm2() * b.norm2())


def test_chosen_config1_against_oracle(gpu_ctx):
    n, depth, chi, cut = 12, 8, 32, 1e-10
    layers = C.brickwall_layers(n, depth, 1)
    st = O.TebdState(n)
    O.run_layers(st, layers, cut, chi)
    dev = DeviceMPS(n, chi, gpu_ctx)
    dev.run_layers(layers, cut, chi, form=FORM_CHOSEN, batched=True)
    bits = C.random_bitstrings(n, 64, 3)
    a_ref = st.amplitudes(bits)
    a = dev.amplitudes(bits)
    fid = abs(np.vdot(a_ref, a)) / (np.linalg.norm(a_ref) *
                                   np.linalg.norm(a))
    assert 1 - fid ** 2 <= 1e-6
    assert np.abs(dev.expect_z() - st.expect_z()).max() <= 1e-5


def test_chosen_config2_layers_infidelity_and_canonical(gpu_ctx):
    """40 qubits, chi 256, cutoff 1e-8, the first 10 layers of the
    headline circuit (bonds saturate at 256 from layer 7)."""
    n, chi, cut = 40, 256, 1e-8
    layers = C.brickwall_layers(n, 20, 2)[:10]
    ref = DeviceMPS(n, chi, gpu_ctx)
    ref.run_layers(layers, cut, chi, batched=True)
    dev = DeviceMPS(n, chi, gpu_ctx)
    dev.run_layers(layers, cut, chi, form=FORM_CHOSEN, batched=True)
    assert infidelity(ref, dev) <= 1e-6
    assert abs(dev.norm2() - ref.norm2()) <= 1e-8
    # the right site of every gate of the last layer is V_k^H: exactly
    # right-canonical, sum_{s,b} B[a,s,b] conj(B[a',s,b]) = delta
    for j in (5, 21, 33):
        B = dev.site(j)
        Bm = B.reshape((B.shape[0], -1), order="F")
        assert np.abs(Bm @ Bm.conj().T - np.eye(B.shape[0])).max() <= 1e-12
    # kept spectra agree with the FP64 ones where both keep a value
    for bond in (10, 20, 30):
        s_ref = ref.bond_spectrum(bond)
        s = dev.bond_spectrum(bond)
        k = min(s.size, s_ref.size)
        assert np.abs(s[:k] - s_ref[:k]).max() <= 1e-6


@pytest.mark.parametrize("single_form", [False, True])
def test_fused_round_kernel_opt_in(gpu_ctx, monkeypatch, single_form):
    """The fused per-pair round (QTN_JACOBI_FUSED=1) gives the same
    singular values and truncation as LAPACK, and the same state as the
    default three-launch round on a 40-qubit layer stack."""
    monkeypatch.setenv("QTN_JACOBI_FUSED", "1")
    rng = np.random.default_rng(5)
    A = rng.standard_normal((300, 200)) + 1j * rng.standard_normal((300, 200))
    A = (A * np.exp(-np.arange(200) / 20.0))
    from qaintensor import ops
    U, S, Vh, _ = ops.svd_trunc(A, 1e-10, True, 0)
    ref = np.linalg.svd(A, compute_uv=False)
    assert S.size == O.truncation_rank(ref, 1e-10 * np.linalg.norm(ref))
    assert np.abs(S - ref[:S.size]).max() <= 1e-12 * ref[0]
    n, chi, cut = 40, 256, 1e-8
    layers = C.brickwall_layers(n, 20, 2)[:8]
    form = FORM_CHOSEN if single_form else 0
    a = DeviceMPS(n, chi, gpu_ctx)
    a.run_layers(layers, cut, chi, form=form, batched=True)
    monkeypatch.setenv("QTN_JACOBI_FUSED", "0")
    b = DeviceMPS(n, chi, gpu_ctx)
    b.run_layers(layers, cut, chi, form=form, batched=True)
    assert infidelity(a, b) <= (1e-6 if single_form else 1e-12)
    if not single_form:
        assert list(a.bond_dims()) == list(b.bond_dims())
