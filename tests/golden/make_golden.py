"""Writes the golden fixtures under tests/golden/ from the FP64 oracle.

The reference (pure Julia, with its arithmetic in TensorOperations and
LAPACK) ships no vectors for this path and cannot run in this image, so the
fixtures are the oracle's outputs on the seeded brick-wall inputs; the
oracle itself is pinned by tests/test_oracle.py.  Run from the repo root:
    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "qaintensor.jl_amd"))

from oracle import qtn_oracle as O  # noqa: E402
from qaintensor import circuits as C  # noqa: E402


def config1(chimax, with_dense):
    n, depth, cutoff, gseed, bseed = 12, 8, 1e-10, 1, 11
    layers = C.brickwall_layers(n, depth, gseed)
    bits = C.random_bitstrings(n, 64, bseed)
    st = O.TebdState(n)
    rec = []
    O.run_layers(st, layers, cutoff, chimax, rec)
    spec = np.full((len(rec), 64), np.nan)
    for i, (_, _, _, S) in enumerate(rec):
        spec[i, :S.size] = S
    bonds = np.full((n + 1, 64), np.nan)
    for b in range(n + 1):
        bonds[b, :st.lam[b].size] = st.lam[b]
    out = dict(n=n, depth=depth, chimax=chimax, cutoff=cutoff,
               gate_seed=gseed, bits_seed=bseed, bits=bits,
               amplitudes=st.amplitudes(bits), expect_z=st.expect_z(),
               sites=np.array([r[0] for r in rec]),
               ranks=np.array([r[1] for r in rec]),
               disc=np.array([r[2] for r in rec]), spectra=spec,
               bond_spectra=bonds, bond_dims=np.array(st.bond_dims()))
    if with_dense:
        sv = O.StateVector(n)
        for singles, pairs in layers:
            for w, u in singles:
                sv.apply_1q(w, u)
            for w, g in pairs:
                sv.apply_2q(w, g)
        out["dense_amplitudes"] = sv.amplitudes(bits)
        out["dense_expect_z"] = sv.expect_z()
    return out


def main():
    np.savez_compressed(os.path.join(HERE, "config1_chi32.npz"),
                        **config1(32, False))
    np.savez_compressed(os.path.join(HERE, "config1_chi64.npz"),
                        **config1(64, True))
    print("wrote", os.listdir(HERE))


if __name__ == "__main__":
    main()
