"""CPU: the oracle reproduces the committed golden fixtures, and the
uncapped config-1 run agrees with the dense statevector."""
import os

import numpy as np

from oracle import qtn_oracle as O
from qaintensor import circuits as C

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return dict(np.load(os.path.join(GOLD, name), allow_pickle=False))


def rerun(g):
    n, depth = int(g["n"]), int(g["depth"])
    layers = C.brickwall_layers(n, depth, int(g["gate_seed"]))
    bits = C.random_bitstrings(n, 64, int(g["bits_seed"]))
    assert np.array_equal(bits, g["bits"])
    st = O.TebdState(n)
    rec = []
    O.run_layers(st, layers, float(g["cutoff"]), int(g["chimax"]), rec)
    return st, rec, bits


def test_config1_capped_fixture():
    g = load("config1_chi32.npz")
    st, rec, bits = rerun(g)
    assert [r[1] for r in rec] == list(g["ranks"])
    assert max(g["ranks"]) == 32  # the chi cap binds at the centre
    a = st.amplitudes(bits)
    assert np.linalg.norm(a - g["amplitudes"]) <= 1e-13 * np.linalg.norm(a)
    assert np.abs(st.expect_z() - g["expect_z"]).max() <= 1e-13


def test_config1_uncapped_matches_dense():
    g = load("config1_chi64.npz")
    st, rec, bits = rerun(g)
    a = st.amplitudes(bits)
    d = g["dense_amplitudes"]
    assert np.linalg.norm(a - d) <= 1e-12 * np.linalg.norm(d)
    assert np.abs(st.expect_z() - g["dense_expect_z"]).max() <= 1e-12
    # cutoff 1e-10 only removes numerically zero Schmidt values here
    assert max(g["disc"]) < 1e-20
