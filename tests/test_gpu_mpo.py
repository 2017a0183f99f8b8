"""MPO x MPS on the GPU (qtn_mpo_*, qtn_mps_set_sites, qtn_mps_expect_mpo,
qtn_mps_apply_mpo) against the oracle (tfim_mpo, canonicalize, expect_mpo,
zipup) and the dense Hamiltonian.  Complex FP64; tolerance 1e-10 relative
on values, identical bond dimensions.  The BF16x3 split-precision GEMM and
the chosen-form expectation sweep and zip-up are held to FP32-level
error."""
import numpy as np
import pytest

from oracle import qtn_oracle as O
from qaintensor import DeviceMPO, DeviceMPS, FORM_CHOSEN
from qaintensor import hamiltonians as Hm

pytestmark = pytest.mark.gpu


def upload(tensors, cap, ctx):
    d = DeviceMPS(len(tensors), cap, ctx)
    d.set_sites(tensors)
    return d


def fp32_sweep_launches(ctx):
    """Launches of the FP32 sweep kernels (three-launch inner solve or fused
    round) since profiling was enabled."""
    prof = ctx.profile_read()
    return prof["jac_inner_fp32"]["count"] + prof["jac_round_fp32"]["count"]


def dense_of(dev):
    return O.mps_dense([dev.site(j) for j in range(1, dev.n + 1)])


def test_set_sites_canonical_form(gpu_ctx):
    n = 10
    Bs = O.random_mps(n, 16, 21)
    # scramble the gauge: the library must not rely on left-canonical input
    rng = np.random.default_rng(1)
    for j in range(n - 1):
        G = rng.standard_normal((Bs[j].shape[2],) * 2) + 0j
        Bs[j] = np.einsum("asb,bc->asc", Bs[j], G)
        Bs[j + 1] = np.einsum("cb,bsd->csd", np.linalg.inv(G), Bs[j + 1])
    psi = O.mps_dense(Bs)
    dev = upload(Bs, 16, gpu_ctx)
    v = dense_of(dev)
    assert abs(np.linalg.norm(v) - 1) < 1e-12
    assert np.abs(v - psi / np.linalg.norm(psi)).max() < 1e-12
    _, lam, _ = O.canonicalize(Bs)
    for j in range(1, n):
        assert np.abs(dev.bond_spectrum(j) - lam[j]).max() < 1e-12
    for j in range(1, n + 1):
        B = dev.site(j)
        m = B.reshape((B.shape[0], -1), order="F")
        assert np.abs(m @ m.conj().T - np.eye(B.shape[0])).max() < 1e-12


def test_expect_mpo_tfim_dense(gpu_ctx):
    n = 10
    h = np.random.default_rng(3).uniform(0.5, 1.5, n)
    W = O.tfim_mpo(n, 1.0, h)
    Bs = O.random_mps(n, 32, 4)
    psi = O.mps_dense(Bs)
    ref = np.vdot(psi, O.tfim_dense(n, 1.0, h) @ psi)
    dev = upload(Bs, 32, gpu_ctx)
    mpo = DeviceMPO(W, gpu_ctx)
    e = dev.expect_mpo(mpo)
    assert abs(e - ref) <= 1e-10 * abs(ref)


def test_apply_mpo_exact_equals_dense(gpu_ctx):
    n = 9
    h = np.random.default_rng(5).uniform(0.5, 1.5, n)
    W = O.tfim_mpo(n, 1.0, h)
    Bs = O.random_mps(n, 8, 6)
    psi = O.mps_dense(Bs)
    Hpsi = O.tfim_dense(n, 1.0, h) @ psi
    dev = upload(Bs, 8, gpu_ctx)
    out = DeviceMPS(n, 32, gpu_ctx)
    nrm = dev.apply_mpo(DeviceMPO(W, gpu_ctx), 0.0, 32, out)
    assert abs(nrm - np.linalg.norm(Hpsi)) <= 1e-10 * nrm
    v = dense_of(out) * nrm
    assert np.abs(v - Hpsi).max() <= 1e-10 * np.abs(Hpsi).max()


@pytest.mark.parametrize("n,chi,cap", [(16, 16, 16), (24, 32, 24)])
def test_apply_mpo_truncated_matches_oracle_zipup(gpu_ctx, n, chi, cap):
    h = np.random.default_rng(n).uniform(0.5, 1.5, n)
    W = O.tfim_mpo(n, 1.0, h)
    Bs = O.random_mps(n, chi, 7)
    cutoff = 1e-8
    # zip-up truncates against the input's gauge: both sides start from the
    # right-canonical form the device MPS holds
    Bc, _, _ = O.canonicalize(Bs)
    B_ref, lam_ref, nrm_ref = O.zipup(Bc, W, cutoff, cap)
    dev = upload(Bs, chi, gpu_ctx)
    out = DeviceMPS(n, cap, gpu_ctx)
    nrm = dev.apply_mpo(DeviceMPO(W, gpu_ctx), cutoff, cap, out)
    assert abs(nrm - nrm_ref) <= 1e-10 * nrm_ref
    dims = list(out.bond_dims())
    assert dims == [1] + [b.shape[2] for b in B_ref]
    for j in range(1, n):
        assert np.abs(out.bond_spectrum(j) - lam_ref[j]).max() <= 1e-10
    ref = upload(B_ref, cap, gpu_ctx)
    fid = abs(out.overlap(ref))
    assert abs(1 - fid) <= 1e-10
    # <psi|W|psi> from the exact environment sweep agrees with the oracle
    e = dev.expect_mpo(DeviceMPO(W, gpu_ctx))
    e_ref = O.expect_mpo(Bs, W)
    assert abs(e - e_ref) <= 1e-10 * abs(e_ref)


@pytest.mark.parametrize("n,chi,cap", [(20, 32, 24), (14, 64, 40)])
def test_apply_mpo_chosen_form_zipup(n, chi, cap):
    """The chosen-form zip-up (FP32 sweeps on the cap-decided splits, FP64
    re-orthonormalisation of the leading cap + 64 vectors and an FP64
    Rayleigh-Ritz split of the projection): the FP32 kernels run, the bond
    dimensions equal the FP64 zip-up's, and the compressed state and norm
    agree with it within the chosen form's bound (infidelity and norm
    difference <= 1e-6 relative)."""
    from qaintensor import Context
    ctx = Context(0)
    try:
        h = np.random.default_rng(n + 1).uniform(0.5, 1.5, n)
        W = O.tfim_mpo(n, 1.0, h)
        Bs = O.random_mps(n, chi, 8)
        dev = upload(Bs, chi, ctx)
        mpo = DeviceMPO(W, ctx)
        out64 = DeviceMPS(n, cap, ctx)
        nrm64 = dev.apply_mpo(mpo, 1e-8, cap, out64)
        assert fp32_sweep_launches(ctx) == 0
        ctx.profile(True)
        out32 = DeviceMPS(n, cap, ctx)
        nrm32 = dev.apply_mpo(mpo, 1e-8, cap, out32, form=FORM_CHOSEN)
        assert fp32_sweep_launches(ctx) > 0
        ctx.profile(False)
        assert list(out32.bond_dims()) == list(out64.bond_dims())
        assert max(out64.bond_dims()) == cap
        dn = abs(nrm32 - nrm64) / nrm64
        infid = abs(1 - abs(out32.overlap(out64)))
        ds = max(np.abs(out32.bond_spectrum(j) - out64.bond_spectrum(j)).max()
                 for j in range(1, n))
        print("chosen zip-up: norm %.2e infidelity %.2e spectrum %.2e"
              % (dn, infid, ds))
        assert dn <= 1e-6
        assert infid <= 1e-6
        assert ds <= 1e-4
    finally:
        ctx.close()


def _chosen_vs_fp64(ctx, dev, mpo, n, cap):
    """FP64 and chosen-form zip-ups of dev; returns (out64, out32, counters
    of the chosen run) after checking the chosen form's bound: same bond
    dimensions, norm difference and infidelity <= 1e-6."""
    out64 = DeviceMPS(n, cap, ctx)
    nrm64 = dev.apply_mpo(mpo, 1e-8, cap, out64)
    ctx.counters(reset=True)
    out32 = DeviceMPS(n, cap, ctx)
    nrm32 = dev.apply_mpo(mpo, 1e-8, cap, out32, form=FORM_CHOSEN)
    cnt = ctx.counters(reset=True)
    print("%d splits accepted, %d redone in FP64"
          % (cnt["rr_accepted"], cnt["rr_redone"]))
    assert list(out32.bond_dims()) == list(out64.bond_dims())
    assert abs(nrm32 - nrm64) <= 1e-6 * nrm64
    assert abs(1 - abs(out32.overlap(out64))) <= 1e-6
    return out64, out32, cnt


def test_apply_mpo_chosen_form_split_passes_its_check(gpu_ctx):
    """The chosen zip-up's FP64 check of an FP32-seeded Rayleigh-Ritz split
    (weight outside span(Q) of the cap + 64 vectors <= weight of the 64
    guard vectors) passes where the spectrum falls off past the cap: TFIM
    on a random 16-site state of bond 128, cap 192 (splits up to
    384 x 384, square, so the smallest third of the values carries less
    weight than the 64 past the cap)."""
    n, chi, cap = 16, 128, 192
    W = Hm.tfim_mpo(n, 1.0, Hm.tfim_fields(n, 4))
    dev = upload(O.random_mps(n, chi, 31), chi, gpu_ctx)
    out64, _, cnt = _chosen_vs_fp64(gpu_ctx, dev, DeviceMPO(W, gpu_ctx), n,
                                    cap)
    assert cnt["rr_accepted"] > 0 and cnt["rr_redone"] == 0
    assert max(out64.bond_dims()) == cap


def test_apply_mpo_chosen_form_flat_spectrum_is_redone_in_fp64(gpu_ctx):
    """A split whose spectrum is flat past the cap fails the check and is
    redone by the FP64 engine.  18 qubits, psi = sum_a L_a R_a / sqrt(384)
    with orthonormal L_a = l_(a div 2) (x) |a mod 2> on qubits 1-9 (l: a
    random 256 x 192 isometry, so bond 8 has rank 192) and R_a a random
    512 x 384 isometry on qubits 10-18: 384 equal Schmidt values at bond
    9.  The identity MPO with cap 192 splits that bond as a 384 x 384
    matrix: 128 values outside span(Q) against 64 in the guard."""
    rng = np.random.default_rng(17)

    def isometry(r, c):
        q, _ = np.linalg.qr(rng.standard_normal((r, c)) +
                            1j * rng.standard_normal((r, c)))
        return q

    n, cap = 18, 192
    lq, R = isometry(256, 192), isometry(512, 384)
    Lm = np.zeros((512, 384), dtype=np.complex128)
    for a in range(384):
        Lm[256 * (a % 2):256 * (a % 2) + 256, a] = lq[:, a // 2]
    psi = (Lm @ R.T).ravel(order="F") / np.sqrt(384.0)
    # site tensors by an SVD sweep that drops the exactly-zero values
    T, m, lb = [], psi.reshape((1, -1)), 1
    for _ in range(n - 1):
        m = m.reshape((2 * lb, -1), order="F")
        U, S, Vh = np.linalg.svd(m, full_matrices=False)
        r = int(np.sum(S > 1e-12 * S[0]))
        T.append(U[:, :r].reshape((lb, 2, r), order="F"))
        m, lb = S[:r, None] * Vh[:r], r
    T.append(m.reshape((lb, 2, 1), order="F"))
    assert [t.shape[2] for t in T][7:9] == [192, 384]
    dev = upload(T, 384, gpu_ctx)
    eye = [np.eye(2, dtype=np.complex128).reshape((1, 2, 2, 1))] * n
    out64, _, cnt = _chosen_vs_fp64(gpu_ctx, dev, DeviceMPO(eye, gpu_ctx), n,
                                    cap)
    assert cnt["rr_redone"] >= 1
    assert out64.bond_dims()[9] == cap


def test_apply_mpo_chosen_form_falls_back_below_the_cap():
    """|a> + |b> (two random product states: rank 2 at every bond) under
    W = 2 I written with bond dimension 2 (two identical paths), cap 3: the
    zip-up splits are 4 x 4 of rank 2, so each is decided by the cutoff
    below the cap, FP32 keeps 2 != 3 and the split is redone in FP64; the
    chosen zip-up equals the FP64 one bit for bit while the FP32 kernels
    ran."""
    from qaintensor import Context
    from qaintensor._lib import QTN_ERR_ARG, QtnError
    n = 10
    rng = np.random.default_rng(12)
    va = rng.standard_normal((n, 2)) + 1j * rng.standard_normal((n, 2))
    vb = rng.standard_normal((n, 2)) + 1j * rng.standard_normal((n, 2))
    Bs = []
    for j in range(n):
        dl = 1 if j == 0 else 2
        dr = 1 if j == n - 1 else 2
        B = np.zeros((dl, 2, dr), dtype=np.complex128)
        for a, v in enumerate((va, vb)):
            B[0 if dl == 1 else a, :, 0 if dr == 1 else a] = v[j]
        Bs.append(B)
    eye = np.eye(2, dtype=np.complex128)
    W = []
    for j in range(n):
        dl = 1 if j == 0 else 2
        dr = 1 if j == n - 1 else 2
        Wj = np.zeros((dl, 2, 2, dr), dtype=np.complex128)
        for d in range(2):
            Wj[0 if dl == 1 else d, :, :, 0 if dr == 1 else d] = eye
        W.append(Wj)
    ctx = Context(0)
    try:
        dev = upload(Bs, 2, ctx)
        mpo = DeviceMPO(W, ctx)
        out64 = DeviceMPS(n, 3, ctx)
        nrm64 = dev.apply_mpo(mpo, 1e-10, 3, out64)
        ctx.profile(True)
        out32 = DeviceMPS(n, 3, ctx)
        nrm32 = dev.apply_mpo(mpo, 1e-10, 3, out32, form=FORM_CHOSEN)
        assert fp32_sweep_launches(ctx) > 0
        ctx.profile(False)
        assert nrm32 == nrm64
        assert list(out32.bond_dims()) == list(out64.bond_dims())
        assert max(out64.bond_dims()) == 2
        for j in range(1, n + 1):
            assert np.array_equal(out32.site(j), out64.site(j))
        with pytest.raises(QtnError) as e:
            dev.apply_mpo(mpo, 1e-10, 3, out32, form=7)
        assert e.value.status == QTN_ERR_ARG
    finally:
        ctx.close()


@pytest.mark.parametrize("op,m,n,k", [("N", 200, 130, 300), ("C", 257, 64, 129),
                                      ("N", 640, 512, 1024)])
def test_zgemm_bf16x3_split_precision(gpu_ctx, op, m, n, k):
    """BF16x3 (three bf16 terms, six products, FP32 accumulation): entries
    within FP32-level error of the FP64 product, relative to the row and
    column norms that bound them, across tile padding; not the FP64 result
    itself (the split path ran)."""
    from qaintensor import ops
    rng = np.random.default_rng(m + n + k)
    shape = (m, k) if op == "N" else (k, m)
    A = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    B = rng.standard_normal((k, n)) + 1j * rng.standard_normal((k, n))
    C = ops.zgemm_bf16x3(A, B, op, ctx=gpu_ctx)
    opA = A if op == "N" else A.conj().T
    ref = opA @ B
    bound = np.outer(np.linalg.norm(opA, axis=1), np.linalg.norm(B, axis=0))
    rel = np.max(np.abs(C - ref) / bound)
    assert rel < 2e-6
    assert rel > 1e-12


def test_zgemm_bf16x3_leading_dimensions(gpu_ctx):
    """Leading dimensions larger than the rows: only the rows x cols
    extents are read, and the padding rows of C keep the caller's values
    (the entry point's host arrays are BLAS-style column-major views)."""
    from qaintensor import _lib as L
    rng = np.random.default_rng(5)
    m, n, k = 70, 33, 90
    lda, ldb, ldc = m + 7, k + 5, m + 11
    Ab = rng.standard_normal((lda, k)) + 1j * rng.standard_normal((lda, k))
    Bb = rng.standard_normal((ldb, n)) + 1j * rng.standard_normal((ldb, n))
    Ab, Bb = np.asfortranarray(Ab), np.asfortranarray(Bb)
    Cb = np.full((ldc, n), 3.0 - 2.0j, dtype=np.complex128, order="F")
    L.check(L.load().qtn_zgemm_bf16x3(
        gpu_ctx.handle, b"N", m, n, k, L.dptr(Ab), lda, L.dptr(Bb), ldb,
        L.dptr(Cb), ldc))
    ref = Ab[:m] @ Bb[:k]
    bound = np.outer(np.linalg.norm(Ab[:m], axis=1),
                     np.linalg.norm(Bb[:k], axis=0))
    assert np.max(np.abs(Cb[:m] - ref) / bound) < 2e-6
    assert np.all(Cb[m:] == 3.0 - 2.0j)
    # op = 'C' with a padded A (k x m stored, lda > k)
    Ac = np.asfortranarray(rng.standard_normal((k + 3, m)) +
                           1j * rng.standard_normal((k + 3, m)))
    Cc = np.full((ldc, n), -1.0, dtype=np.complex128, order="F")
    L.check(L.load().qtn_zgemm_bf16x3(
        gpu_ctx.handle, b"C", m, n, k, L.dptr(Ac), k + 3, L.dptr(Bb), ldb,
        L.dptr(Cc), ldc))
    refc = Ac[:k].conj().T @ Bb[:k]
    boundc = np.outer(np.linalg.norm(Ac[:k], axis=0),
                      np.linalg.norm(Bb[:k], axis=0))
    assert np.max(np.abs(Cc[:m] - refc) / boundc) < 2e-6
    assert np.all(Cc[m:] == -1.0)


def test_expect_mpo_bf16x3_form(gpu_ctx):
    """<psi|H|psi> of a chi = 512 state: the chosen form (BF16x3 on the
    contractions with every dimension >= 512) within FP32-level error of
    the FP64 sweep, which matches the oracle."""
    n, chi = 22, 512
    W = Hm.tfim_mpo(n, 1.0, Hm.tfim_fields(n, 9))
    raw = Hm.random_mps_tensors(n, chi, 9)
    psi = DeviceMPS(n, chi, gpu_ctx)
    psi.set_sites(raw)
    mpo = DeviceMPO(W, gpu_ctx)
    e64 = psi.expect_mpo(mpo)
    e16 = psi.expect_mpo(mpo, FORM_CHOSEN)
    ref = O.expect_mpo_blas(O.left_canonical(raw), W)
    assert abs(e64 - ref) <= 1e-10 * abs(ref)
    rel = abs(e16 - e64) / abs(e64)
    assert rel <= 1e-5
    assert rel > 0.0
