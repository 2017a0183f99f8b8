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


def test_1q_multi_rejects_two_gates_on_one_site(gpu_ctx):
    """One launch applies each site's gate once: a second gate on the same
    (state, site) is refused before anything runs; DeviceMPS.run_layers
    then applies such gates one by one, in order (X then H on wire 1 gives
    H X |0> = (|0> - |1>) / sqrt(2))."""
    from qaintensor.device import apply_1q_multi
    d = DeviceMPS(3, 4, gpu_ctx)
    x = np.array([[0, 1], [1, 0]], dtype=complex)
    h = np.array([[1, 1], [1, -1]], dtype=complex) / np.sqrt(2.0)
    with pytest.raises(QtnError, match="two gates on one site"):
        apply_1q_multi([d, d], [1, 1], [x, h])
    d.run_layers([([(1, x), (1, h)], [])], 1e-10, 4)
    bits = np.array([[0, 0, 0], [1, 0, 0]], dtype=np.uint8)
    amp = d.amplitudes(bits)
    assert np.allclose(amp, [1 / np.sqrt(2.0), -1 / np.sqrt(2.0)],
                       atol=1e-14)


def test_jacobi_autotuner_table_and_results(gpu_ctx):
    """qtn_tune_jacobi times the tilings of a bucket and keeps the fastest;
    states run under each forced tiling agree with the default run to the
    plain form's bar.  Private contexts keep the tuned tables away from the
    other tests."""
    from qaintensor import Context, autotune as A
    tctx = Context(0)
    table = A.tune_jacobi(tctx, buckets=((128, 128), (256, 192)),
                          batches=(1, 5), precisions=("fp64", "fp32"))
    assert set(table) == {"%s:%d:b%d" % (p, c, b) for p in ("fp64", "fp32")
                          for c in (128, 192) for b in (1, 4)}
    for entry in table.values():
        us = entry["us_per_svd"]
        assert us["w32"] > 0 and us["w64"] > 0
        assert entry["best"] == min((v, k) for k, v in us.items()
                                    if v is not None)[1]
    n, chi, cutoff = 14, 64, 1e-10
    layers = C.brickwall_layers(n, 10, 21)
    for tiling in ("w32", "w32_fused", "w64"):
        c = Context(0)
        ref = DeviceMPS(n, chi, c)  # default tilings
        ref.run_layers(layers, cutoff, chi, batched=True)
        # every batch class of the 128-column bucket
        A.apply_table({"fp64:128:b%d" % (1 << k): {"best": tiling}
                       for k in range(7)}, c)
        dev = DeviceMPS(n, chi, c)
        dev.run_layers(layers, cutoff, chi, batched=True)
        assert list(dev.bond_dims()) == list(ref.bond_dims())
        ov = dev.overlap(ref)
        assert abs(1 - abs(ov) ** 2 / (dev.norm2() * ref.norm2())) <= 1e-10
        dev.free()
        ref.free()
