"""Autotuner of the Jacobi SVD engine's block-pair tiling.

For every column bucket (the padded width of the two-site matrix, a
multiple of 64), sweep precision and batch class (floor(log2) of the number
of problems one batched SVD solves, capped at 6) the library can run three
tilings:
32x32 Gram blocks in three launches a round (QTN_TILING_W32), the same in
one fused launch (QTN_TILING_W32_FUSED), or 64x64 Gram blocks
(QTN_TILING_W64).  qtn_tune_jacobi times each on a batch of synthetic
complex Gaussian problems of the bucket's size with HIP events and keeps
the fastest on the context; every later SVD of that bucket and class on
the context uses it.  Which tiling wins depends on the batch: one
workgroup per block pair (the fused round) hides its latency only when many
problems run at once.  The table can be saved as JSON and set again on
another context (qtn_set_jacobi_tiling) without re-timing.
"""
import ctypes
import json

import numpy as np

from . import _lib as L

TILINGS = {0: "w32", 1: "w32_fused", 2: "w64"}
_CODES = {v: k for k, v in TILINGS.items()}


def batch_class(n):
    """floor(log2(n)) capped at 6, the library's batch class."""
    c = 0
    while c < 6 and (2 << c) <= n:
        c += 1
    return c


def tune_jacobi(ctx=None, buckets=((128, 128), (256, 256), (512, 512)),
                batches=(16,), precisions=("fp64", "fp32")):
    """Time the tilings for each (rows, cols) bucket, batch size and
    precision; the fastest is kept on ctx.  Returns a table
    {"fp64:512:b16": {"best": "w32", "us_per_svd": {...}, ...}, ...} whose
    keys name the bucket and the smallest batch of the class."""
    ctx = ctx if ctx is not None else L.default_context()
    table = {}
    for prec in precisions:
        single = 1 if prec == "fp32" else 0
        for batch in batches:
            for rows, cols in buckets:
                best = ctypes.c_int32(0)
                us = np.full(3, -1.0)
                L.check(L.load().qtn_tune_jacobi(
                    ctx.handle, int(rows), int(cols), int(batch), single,
                    ctypes.byref(best), L.dptr(us)))
                cols_p = ((int(cols) + 63) // 64) * 64
                key = "%s:%d:b%d" % (prec, cols_p, 1 << batch_class(batch))
                table[key] = {
                    "best": TILINGS[best.value],
                    "us_per_svd": {TILINGS[i]: (float(us[i]) if us[i] >= 0
                                                else None) for i in range(3)},
                    "rows": int(rows), "batch": int(batch)}
    return table


def apply_table(table, ctx=None):
    """Set a saved table's tilings on ctx (no timing)."""
    ctx = ctx if ctx is not None else L.default_context()
    for key, entry in table.items():
        prec, cols, b = key.split(":")
        L.check(L.load().qtn_set_jacobi_tiling(
            ctx.handle, int(cols), 1 if prec == "fp32" else 0,
            int(b.lstrip("b")), _CODES[entry["best"]]))


def save_table(table, path):
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)


def load_table(path):
    with open(path) as f:
        return json.load(f)
