"""Many circuits in one batch (qtn_mps_apply_layer_multi /
qtn_mps_apply_1q_multi): every circuit of a mixed batch (different qubit
counts, depths and bond caps) must end in the state its own FP64 oracle run
reaches (amplitudes 1e-10 relative, <Z> 1e-10, identical bond dimensions),
and the chosen form within 1e-6 infidelity."""
import numpy as np
import pytest

from oracle import qtn_oracle as O
from qaintensor import (DeviceMPS, FORM_CHOSEN, circuits as C,
                        run_circuits)
from qaintensor._lib import QtnError

pytestmark = pytest.mark.gpu


def oracle_state(n, chi, layers, cutoff):
    st = O.TebdState(n)
    O.run_layers(st, layers, cutoff, chi)
    return st


def test_mixed_batch_matches_each_oracle(gpu_ctx):
    cutoff = 1e-10
    circs = C.batch_circuits(10, 5, n_range=(6, 14), depth_range=(3, 9),
                             chis=(8, 16, 32))
    states = [DeviceMPS(n, chi, gpu_ctx) for (n, _, chi, _) in circs]
    run_circuits(states, [c[3] for c in circs], [c[2] for c in circs],
                 cutoff)
    for dev, (n, depth, chi, layers) in zip(states, circs):
        ref = oracle_state(n, chi, layers, cutoff)
        assert list(dev.bond_dims()) == ref.bond_dims(), (n, depth, chi)
        bits = C.random_bitstrings(n, 16, n)
        a, r = dev.amplitudes(bits), ref.amplitudes(bits)
        assert np.linalg.norm(a - r) <= 1e-10 * np.linalg.norm(r)
        assert np.abs(dev.expect_z() - ref.expect_z()).max() <= 1e-10


def test_mixed_batch_chosen_form(gpu_ctx):
    cutoff = 1e-8
    circs = C.batch_circuits(6, 9, n_range=(16, 24), depth_range=(8, 12),
                             chis=(32, 64))
    a = [DeviceMPS(n, chi, gpu_ctx) for (n, _, chi, _) in circs]
    b = [DeviceMPS(n, chi, gpu_ctx) for (n, _, chi, _) in circs]
    run_circuits(a, [c[3] for c in circs], [c[2] for c in circs], cutoff)
    run_circuits(b, [c[3] for c in circs], [c[2] for c in circs], cutoff,
                 form=FORM_CHOSEN)
    for x, y in zip(a, b):
        inf = 1 - abs(x.overlap(y)) ** 2 / (x.norm2() * y.norm2())
        assert inf <= 1e-6


def test_multi_rejects_overlap(gpu_ctx):
    from qaintensor import apply_layer_multi
    d = DeviceMPS(6, 8, gpu_ctx)
    g = np.eye(4, dtype=complex)
    with pytest.raises(QtnError, match="overlap"):
        apply_layer_multi([d, d], [2, 3], [g, g], [8, 8], 1e-10)
