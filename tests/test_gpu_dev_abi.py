"""The device-pointer entry points (qtn_zgemm_dev, qtn_cgemm_dev,
qtn_permute_dev, qtn_contract_pair_dev) and qtn_mps_expect_mpo, called
through the C ABI on column-major interleaved buffers held in HBM (torch
only allocates them), against numpy on the host.  The _dev calls are
enqueued on the context's stream without a host sync, so each test waits
for torch's stream before the call and for the context's after it.
Tolerances: 1e-12 relative (FP64, permutes bit-exact), 2e-5 (FP32)."""
import ctypes

import numpy as np
import pytest

from oracle import qtn_oracle as O
from qaintensor import DeviceMPO, DeviceMPS, _lib as L

pytestmark = pytest.mark.gpu


def to_dev(a, single=False):
    """Column-major interleaved copy of a on cuda:0 (a flat real tensor)."""
    import torch
    cdt, rdt = (np.complex64, np.float32) if single else (np.complex128,
                                                          np.float64)
    flat = np.asarray(a, dtype=cdt).ravel(order="F")
    t = torch.from_numpy(flat.view(rdt).copy()).to("cuda:0")
    torch.cuda.synchronize()
    return t


def from_dev(t, shape, single=False):
    cdt = np.complex64 if single else np.complex128
    return t.cpu().numpy().view(cdt).reshape(shape, order="F")


def vp(t):
    return ctypes.c_void_p(t.data_ptr())


def crandn(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def op(X, o):
    return X if o == "N" else (X.T if o == "T" else X.conj().T)


@pytest.mark.parametrize("single", [False, True])
@pytest.mark.parametrize("oa,ob,shared_b", [("N", "N", False),
                                            ("T", "C", False),
                                            ("C", "N", True),
                                            ("N", "T", True)])
def test_gemm_dev_batched(gpu_ctx, single, oa, ob, shared_b):
    """C_b = alpha op(A_b) op(B_b) + beta C_b for 3 problems with element
    strides (stride 0: B shared by every problem)."""
    rng = np.random.default_rng(ord(oa) + 3 * ord(ob) + 7 * shared_b
                                + 13 * single)
    m, n, k, batch = 70, 45, 33, 3
    ra, ca = (m, k) if oa == "N" else (k, m)
    rb, cb = (k, n) if ob == "N" else (n, k)
    A = crandn(rng, ra, ca, batch)
    B = crandn(rng, rb, cb, 1 if shared_b else batch)
    C0 = crandn(rng, m, n, batch)
    alpha, beta = 0.7 - 0.2j, 0.3 + 0.5j
    dA, dB, dC = to_dev(A, single), to_dev(B, single), to_dev(C0, single)
    lib = L.load()
    if single:
        fa = np.array([alpha.real, alpha.imag], dtype=np.float32)
        fb = np.array([beta.real, beta.imag], dtype=np.float32)
        fp = ctypes.POINTER(ctypes.c_float)
        st = lib.qtn_cgemm_dev(
            gpu_ctx.handle, oa.encode(), ob.encode(), m, n, k,
            fa.ctypes.data_as(fp), vp(dA), ra, ra * ca, vp(dB), rb,
            0 if shared_b else rb * cb, fb.ctypes.data_as(fp), vp(dC), m,
            m * n, batch)
    else:
        da = np.array([alpha.real, alpha.imag])
        db = np.array([beta.real, beta.imag])
        st = lib.qtn_zgemm_dev(
            gpu_ctx.handle, oa.encode(), ob.encode(), m, n, k, L.dptr(da),
            vp(dA), ra, ra * ca, vp(dB), rb, 0 if shared_b else rb * cb,
            L.dptr(db), vp(dC), m, m * n, batch)
    L.check(st)
    gpu_ctx.sync()
    C = from_dev(dC, (m, n, batch), single)
    tol = 2e-5 if single else 1e-12
    for b in range(batch):
        Bb = B[:, :, 0 if shared_b else b]
        ref = alpha * op(A[:, :, b], oa) @ op(Bb, ob) + beta * C0[:, :, b]
        scale = (np.abs(op(A[:, :, b], oa)) @ np.abs(op(Bb, ob))).max()
        assert np.abs(C[:, :, b] - ref).max() <= tol * scale


def test_gemm_dev_rejects_unknown_op(gpu_ctx):
    one = np.array([1.0, 0.0])
    st = L.load().qtn_zgemm_dev(gpu_ctx.handle, b"X", b"N", 1, 1, 1,
                                L.dptr(one), None, 1, 0, None, 1, 0,
                                L.dptr(one), None, 1, 0, 1)
    assert st == L.QTN_ERR_ARG


@pytest.mark.parametrize("oa,ob,lda,ldb,ldc", [
    ("N", "N", 69, 33, 70),   # lda < m for op(A) = A (m x k stored)
    ("T", "N", 32, 33, 70),   # lda < k for op(A) = A^T (k x m stored)
    ("N", "N", 70, 32, 70),   # ldb < k for op(B) = B (k x n stored)
    ("N", "C", 70, 44, 70),   # ldb < n for op(B) = B^H (n x k stored)
    ("N", "N", 70, 33, 69),   # ldc < m
])
def test_gemm_dev_rejects_short_leading_dimension(gpu_ctx, oa, ob, lda, ldb,
                                                  ldc):
    """BLAS's leading-dimension rule (m = 70, n = 45, k = 33): a stored
    operand shorter than its rows is an argument error, not a read out of
    bounds; nothing is launched (the device pointers are never touched)."""
    one = np.array([1.0, 0.0])
    for fn in ("qtn_zgemm_dev", "qtn_cgemm_dev"):
        f = getattr(L.load(), fn)
        if fn == "qtn_cgemm_dev":
            o = np.array([1.0, 0.0], dtype=np.float32)
            al = o.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        else:
            al = L.dptr(one)
        st = f(gpu_ctx.handle, oa.encode(), ob.encode(), 70, 45, 33, al,
               None, lda, 0, None, ldb, 0, al, None, ldc, 0, 1)
        assert st == L.QTN_ERR_ARG
        assert "leading dimension" in L.load().qtn_last_error().decode()


def test_permute_dev(gpu_ctx):
    rng = np.random.default_rng(3)
    dims = (3, 4, 5, 2)
    perm = (3, 1, 4, 2)  # 1-based, out = permutedims(in, perm)
    a = crandn(rng, *dims)
    ref = np.transpose(a, [p - 1 for p in perm])
    din = to_dev(a)
    dout = to_dev(np.zeros(ref.shape, dtype=np.complex128))
    d = np.asarray(dims, dtype=np.int64)
    pm = np.asarray(perm, dtype=np.int32)
    L.check(L.load().qtn_permute_dev(gpu_ctx.handle, vp(din), len(dims),
                                     L.i64ptr(d), L.i32ptr(pm), vp(dout)))
    gpu_ctx.sync()
    assert np.array_equal(from_dev(dout, ref.shape), ref)


def test_contract_pair_dev(gpu_ctx):
    """A (3,4,5) and B (5,6,4) over (A2,B3) and (A3,B1); natural output
    (3,6), returned as (6,3) by outperm (2,1)."""
    rng = np.random.default_rng(4)
    A = crandn(rng, 3, 4, 5)
    B = crandn(rng, 5, 6, 4)
    ref = np.einsum("ijk,klj->il", A, B).T
    dA, dB = to_dev(A), to_dev(B)
    dC = to_dev(np.zeros(ref.shape, dtype=np.complex128))
    da = np.asarray(A.shape, dtype=np.int64)
    db = np.asarray(B.shape, dtype=np.int64)
    la = np.asarray([2, 3], dtype=np.int32)
    lb = np.asarray([3, 1], dtype=np.int32)
    po = np.asarray([2, 1], dtype=np.int32)
    L.check(L.load().qtn_contract_pair_dev(
        gpu_ctx.handle, vp(dA), 3, L.i64ptr(da), vp(dB), 3, L.i64ptr(db), 2,
        L.i32ptr(la), L.i32ptr(lb), L.i32ptr(po), vp(dC)))
    gpu_ctx.sync()
    C = from_dev(dC, ref.shape)
    assert np.abs(C - ref).max() <= 1e-12 * np.abs(ref).max()


def test_expect_mpo_plain_entry_point(gpu_ctx):
    """qtn_mps_expect_mpo equals the plain form of qtn_mps_expect_mpo_form
    bit for bit and the oracle's environment sweep to 1e-10."""
    n = 10
    h = np.random.default_rng(6).uniform(0.5, 1.5, n)
    W = O.tfim_mpo(n, 1.0, h)
    Bs = O.random_mps(n, 16, 6)
    dev = DeviceMPS(n, 16, gpu_ctx)
    dev.set_sites(Bs)
    mpo = DeviceMPO(W, gpu_ctx)
    out = np.zeros(2)
    L.check(L.load().qtn_mps_expect_mpo(dev.handle, mpo.handle,
                                        L.dptr(out)))
    e = complex(out[0], out[1])
    assert e == dev.expect_mpo(mpo)
    ref = O.expect_mpo(Bs, W)
    assert abs(e - ref) <= 1e-10 * abs(ref)
