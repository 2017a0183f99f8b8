"""qtn_contract_pair_dev_noalloc (the allocation-free pairwise contraction,
scratch from a fixed per-GPU device array) against qtn_contract_pair_dev
(the same permute + GEMM plan with the context's growable scratch), through
the C ABI on device buffers: identical results (the same kernels on the
same data), on two calls in a row, the second smaller than the first, and
refusal of a contraction beyond the scratch limit."""
import ctypes

import numpy as np
import pytest

from qaintensor import _lib as L

pytestmark = pytest.mark.gpu


def to_dev(a):
    import torch
    flat = np.asarray(a, dtype=np.complex128).ravel(order="F")
    t = torch.from_numpy(flat.view(np.float64).copy()).to("cuda:0")
    torch.cuda.synchronize()
    return t


def from_dev(t, shape):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy().view(np.complex128).reshape(shape, order="F")


def vp(t):
    return ctypes.c_void_p(t.data_ptr())


def call(fn, ctx, A, B, legsA, legsB, outperm):
    """fn(ctx, ...) on device copies of A and B; returns C on the host."""
    import torch
    dA, dB = to_dev(A), to_dev(B)
    la = [l - 1 for l in legsA]
    lb = [l - 1 for l in legsB]
    fa = [A.shape[i] for i in range(A.ndim) if i not in la]
    fb = [B.shape[i] for i in range(B.ndim) if i not in lb]
    nat = fa + fb
    shape = tuple(nat[p - 1] for p in outperm)
    dC = torch.zeros(2 * int(np.prod(shape)), dtype=torch.float64,
                     device="cuda:0")
    torch.cuda.synchronize()
    dimsA = np.asarray(A.shape, dtype=np.int64)
    dimsB = np.asarray(B.shape, dtype=np.int64)
    a32 = np.asarray(legsA, dtype=np.int32)
    b32 = np.asarray(legsB, dtype=np.int32)
    o32 = np.asarray(outperm, dtype=np.int32)
    L.check(fn(ctx.handle, vp(dA), A.ndim, L.i64ptr(dimsA), vp(dB), B.ndim,
               L.i64ptr(dimsB), len(legsA), L.i32ptr(a32), L.i32ptr(b32),
               L.i32ptr(o32), vp(dC)))
    ctx.sync()
    return from_dev(dC, shape)


def test_noalloc_matches_contract_pair_dev_twice(gpu_ctx):
    """Two calls in a row (the second smaller), each with permutes of A, B
    and the result (every scratch piece in use): the allocation-free call
    returns exactly what qtn_contract_pair_dev returns, and both match the
    einsum definition."""
    lib = L.load()
    rng = np.random.default_rng(9)
    for (da, db, legsA, legsB, outperm, spec) in (
            ((6, 40, 7, 30), (30, 9, 40, 5), [2, 4], [3, 1], [3, 1, 4, 2],
             "axby,ycxd->cadb"),
            ((3, 5, 4), (4, 2, 5), [2, 3], [3, 1], [2, 1],
             "axy,ybx->ba")):
        A = rng.standard_normal(da) + 1j * rng.standard_normal(da)
        B = rng.standard_normal(db) + 1j * rng.standard_normal(db)
        ref = call(lib.qtn_contract_pair_dev, gpu_ctx, A, B, legsA, legsB,
                   outperm)
        got = call(lib.qtn_contract_pair_dev_noalloc, gpu_ctx, A, B, legsA,
                   legsB, outperm)
        assert np.array_equal(got, ref)
        dense = np.einsum(spec, A, B)
        assert np.max(np.abs(got - dense)) <= 1e-12 * np.max(np.abs(dense))


def test_noalloc_refuses_beyond_scratch(gpu_ctx):
    """A contraction whose permuted operands need more than the 32 MiB is
    refused before anything is launched (the buffers are never touched)."""
    lib = L.load()
    dimsA = np.asarray([512, 1024, 3], dtype=np.int64)   # permuted: 24 MiB
    dimsB = np.asarray([3, 1024, 600], dtype=np.int64)   # permuted: 28 MiB
    a32 = np.asarray([2], dtype=np.int32)
    b32 = np.asarray([2], dtype=np.int32)
    st = lib.qtn_contract_pair_dev_noalloc(
        gpu_ctx.handle, None, 3, L.i64ptr(dimsA), None, 3, L.i64ptr(dimsB), 1,
        L.i32ptr(a32), L.i32ptr(b32), None, None)
    assert st == L.QTN_ERR_ARG
    assert "32 MiB" in lib.qtn_last_error().decode()
