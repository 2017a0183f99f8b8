"""Array-level wrappers of libqtn's compute entry points.

Arrays follow the reference's (Julia) index order and are passed to the
library in column-major (Fortran) memory order, so dims go across the ABI in
Julia order unchanged.
"""
import ctypes

import numpy as np

from . import _lib as L


def _ctx(ctx):
    return ctx if ctx is not None else L.default_context()


def _dims(a):
    return np.ascontiguousarray(a.shape, dtype=np.int64)


def contract_pair(A, B, legsA, legsB, outperm=None, ctx=None):
    """One pairwise ncon step on the GPU (qtn_contract_pair): contract leg
    legsA[i] of A with legsB[i] of B (1-based); output axes are the free
    legs of A then of B, permuted by outperm (1-based) if given."""
    ctx = _ctx(ctx)
    A = L.as_c128(A, "F")
    B = L.as_c128(B, "F")
    la = np.ascontiguousarray(legsA, dtype=np.int32)
    lb = np.ascontiguousarray(legsB, dtype=np.int32)
    if la.size != lb.size:
        raise ValueError("contract_pair: legsA and legsB differ in length")
    fa = [A.shape[i] for i in range(A.ndim) if (i + 1) not in set(la)]
    fb = [B.shape[i] for i in range(B.ndim) if (i + 1) not in set(lb)]
    nat = fa + fb
    if outperm is not None:
        op = np.ascontiguousarray(outperm, dtype=np.int32)
        if sorted(op.tolist()) != list(range(1, len(nat) + 1)):
            raise ValueError("contract_pair: outperm is not a permutation")
        out_shape = [nat[p - 1] for p in op]
        op_ptr = L.i32ptr(op)
    else:
        out_shape = nat
        op_ptr = None
    C = np.zeros(out_shape, dtype=np.complex128, order="F")
    dA, dB = _dims(A), _dims(B)
    L.check(L.load().qtn_contract_pair(
        ctx.handle, L.dptr(A), A.ndim, L.i64ptr(dA), L.dptr(B), B.ndim,
        L.i64ptr(dB), la.size, L.i32ptr(la), L.i32ptr(lb), op_ptr,
        L.dptr(C)))
    return C


def contract_network(tensors, labels, order, ctx=None):
    """Contract a whole network on the GPU (qtn_contract_network): arrays
    (Julia index order), ncon labels per tensor, and the contraction order
    (every positive label once).  Returns the result with its legs ordered
    by open label -1, -2, ... (a 0-d array for a scalar)."""
    ctx = _ctx(ctx)
    arrs = [L.as_c128(t, "F") for t in tensors]
    if len(arrs) != len(labels):
        raise ValueError("one label list per tensor")
    ndims = np.asarray([a.ndim for a in arrs], dtype=np.int32)
    dims = np.asarray([d for a in arrs for d in a.shape], dtype=np.int64)
    labs = np.asarray([x for l in labels for x in l], dtype=np.int32)
    if labs.size != dims.size:
        raise ValueError("one label per leg")
    size = {}
    for a, l in zip(arrs, labels):
        for x, d in zip(l, a.shape):
            size[x] = d
    open_ = sorted((x for x in size if x < 0), reverse=True)
    shape = tuple(size[x] for x in open_)
    out = np.zeros(shape, dtype=np.complex128, order="F")
    ordr = np.ascontiguousarray(order, dtype=np.int32)
    ptrs = (ctypes.POINTER(ctypes.c_double) * len(arrs))(
        *[L.dptr(a) for a in arrs])
    L.check(L.load().qtn_contract_network(
        ctx.handle, len(arrs), L.i32ptr(ndims), L.i64ptr(dims),
        L.i32ptr(labs), ptrs, ordr.size, L.i32ptr(ordr),
        L.dptr(out), int(out.size)))
    return out


def permute(A, perm, ctx=None):
    """permutedims(A, perm) (perm 1-based) on the GPU."""
    ctx = _ctx(ctx)
    A = L.as_c128(A, "F")
    p = np.ascontiguousarray(perm, dtype=np.int32)
    if sorted(p.tolist()) != list(range(1, A.ndim + 1)):
        raise ValueError("permute: perm is not a permutation")
    out = np.zeros([A.shape[i - 1] for i in p], dtype=np.complex128,
                   order="F")
    dA = _dims(A)
    L.check(L.load().qtn_permute(ctx.handle, L.dptr(A), A.ndim,
                                 L.i64ptr(dA), L.i32ptr(p), L.dptr(out)))
    return out


def svd_trunc(A, cutoff=0.0, relative=False, chimax=0, ctx=None):
    """Truncated SVD on the GPU: (U, S, Vh, discarded weight) with the
    truncation rule of src/svd.jl:29-33 (er = cutoff, or cutoff*||s|| when
    relative), then k <= chimax when chimax > 0."""
    ctx = _ctx(ctx)
    A = L.as_c128(A, "F")
    if A.ndim != 2:
        raise ValueError("svd_trunc: matrix expected")
    m, n = A.shape
    kk = min(m, n)
    U = np.zeros((m, kk), dtype=np.complex128, order="F")
    S = np.zeros(kk, dtype=np.float64)
    Vh = np.zeros((kk, n), dtype=np.complex128, order="F")
    k = ctypes.c_int64(0)
    disc = ctypes.c_double(0.0)
    L.check(L.load().qtn_svd_trunc(
        ctx.handle, m, n, L.dptr(A), m, float(cutoff), int(bool(relative)),
        int(chimax), ctypes.byref(k), L.dptr(U), L.dptr(S), L.dptr(Vh),
        ctypes.byref(disc)))
    k = k.value
    # the library writes U with ld m and Vh with ld k
    Vh_k = np.asfortranarray(
        Vh.reshape(-1, order="F")[:k * n].reshape((k, n), order="F"))
    return U[:, :k].copy(order="F"), S[:k].copy(), Vh_k, disc.value


def svd_values(A, ctx=None):
    """All min(m, n) singular values of A (descending) on the GPU."""
    ctx = _ctx(ctx)
    A = L.as_c128(A, "F")
    m, n = A.shape
    S = np.zeros(min(m, n), dtype=np.float64)
    L.check(L.load().qtn_svd_values(ctx.handle, m, n, L.dptr(A), m,
                                    L.dptr(S)))
    return S


def zgemm_bf16x3(A, B, opA="N", ctx=None):
    """op(A) @ B in BF16x3 split precision on the GPU (FP32-level error;
    qtn_zgemm_bf16x3); opA "N" or "C"."""
    ctx = _ctx(ctx)
    A = L.as_c128(A, "F")
    B = L.as_c128(B, "F")
    if opA not in ("N", "C"):
        raise ValueError("opA must be 'N' or 'C'")
    m = A.shape[0] if opA == "N" else A.shape[1]
    k = A.shape[1] if opA == "N" else A.shape[0]
    if B.shape[0] != k:
        raise ValueError("inner dimensions differ")
    n = B.shape[1]
    C = np.zeros((m, n), dtype=np.complex128, order="F")
    L.check(L.load().qtn_zgemm_bf16x3(
        ctx.handle, opA.encode(), m, n, k, L.dptr(A), max(1, A.shape[0]),
        L.dptr(B), max(1, k), L.dptr(C), max(1, m)))
    return C


def svd(A, absorb=False, ctx=None):
    """Thin SVD without truncation on the GPU (qtn_svd): (U, S, Vh) with all
    min(m, n) singular values, descending; absorb=True returns diag(S) @ Vh
    in place of Vh.  A zero singular value comes with a zero vector."""
    ctx = _ctx(ctx)
    A = L.as_c128(A, "F")
    if A.ndim != 2:
        raise ValueError("svd: matrix expected")
    m, n = A.shape
    k = min(m, n)
    U = np.zeros((m, k), dtype=np.complex128, order="F")
    S = np.zeros(k, dtype=np.float64)
    Vh = np.zeros((k, n), dtype=np.complex128, order="F")
    L.check(L.load().qtn_svd(ctx.handle, m, n, L.dptr(A), m,
                             int(bool(absorb)), L.dptr(U), L.dptr(S),
                             L.dptr(Vh)))
    return U, S, Vh


def contract_svd(A, B, legA, legB, er=0.0, ctx=None):
    """contract_svd on the GPU (qtn_contract_svd): leg legA of A with leg
    legB of B (1-based), each operand first replaced by its SVD truncated at
    the absolute tail norm er."""
    ctx = _ctx(ctx)
    A = L.as_c128(A, "F")
    B = L.as_c128(B, "F")
    if A.ndim < 1 or B.ndim < 1:
        raise ValueError("contract_svd: tensors of rank >= 1 expected")
    dA, dB = _dims(A), _dims(B)
    shape = [A.shape[i] for i in range(A.ndim) if i != legA - 1] + \
            [B.shape[i] for i in range(B.ndim) if i != legB - 1]
    C = np.zeros(shape, dtype=np.complex128, order="F")
    L.check(L.load().qtn_contract_svd(
        ctx.handle, L.dptr(A), A.ndim, L.i64ptr(dA), L.dptr(B), B.ndim,
        L.i64ptr(dB), int(legA), int(legB), float(er), L.dptr(C)))
    return C
