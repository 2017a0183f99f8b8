"""The learned form's randomized range finder with its exact check
(qtn_mps_apply_layer_learned) and the harvest / training / inference loop.

Cases where the sketch must succeed use two-site tensors of known low rank:
a chi = 64 random MPS hit by product gates u (x) v (theta = lambda Phi has
rank chi_M = 64 of 128), so hint 64 (sketch 80) captures it exactly and the
state equals the full-SVD one; a too-small hint must fail the check and
fall back, again giving the full-SVD state."""
import numpy as np
import pytest

from oracle import qtn_oracle as O
from qaintensor import DeviceMPS, circuits as C, hamiltonians as Hm
from qaintensor import learned as Lrn

pytestmark = pytest.mark.gpu


def product_layer(n, seed):
    rng = np.random.default_rng(seed)
    pairs = []
    for w in range(2, n - 1, 2):
        u = C.haar_unitary(rng, 2)
        v = C.haar_unitary(rng, 2)
        pairs.append((w, np.kron(v, u)))
    return pairs


def start_state(ctx, n=16, chi=64, seed=3):
    raw = Hm.random_mps_tensors(n, chi, seed)
    a = DeviceMPS(n, 256, ctx)
    a.set_sites(raw)
    return a


def infid(a, b):
    return 1 - abs(a.overlap(b)) ** 2 / (a.norm2() * b.norm2())


@pytest.mark.parametrize("hint,expect_accept", [(64, True), (8, False),
                                                (0, False)])
def test_range_finder_exact_check(gpu_ctx, hint, expect_accept):
    n, cutoff = 16, 1e-10
    pairs = product_layer(n, 7)
    ref = start_state(gpu_ctx)
    dev = start_state(gpu_ctx)
    sites = [w for (w, _) in pairs]
    gates = [g for (_, g) in pairs]
    d0 = list(ref.bond_dims())
    k_ref, _ = ref.apply_layer(sites, gates, cutoff, 256)
    hints = np.full(len(pairs), hint)
    k, _, acc = dev.apply_layer_learned(sites, gates, cutoff, 256, hints)
    # gates whose sketch (hint + 16) is at most 3/4 of min(M, N)
    big = [i for i, (w, _) in enumerate(pairs)
           if 4 * (hint + 16) <= 3 * min(2 * d0[w - 1], 2 * d0[w + 1])]
    if expect_accept:
        assert len(big) > 0 and all(acc[i] for i in big)
        # exact low rank: the same kept ranks as the full SVD
        assert list(k) == list(k_ref)
    else:
        assert acc.sum() == 0
        assert list(k) == list(k_ref)
    assert infid(dev, ref) <= 1e-10
    assert list(dev.bond_dims()) == list(ref.bond_dims())


def test_harvest_train_predict_roundtrip(gpu_ctx):
    """Harvest (features, kept rank) from a plain run, train the MLP on the
    GPU, and run the learned form with it: the state matches the plain
    run (every randomized decision passed the exact check)."""
    n, chi, cutoff = 16, 64, 1e-8
    layers = C.brickwall_layers(n, 10, 4)
    X, T = Lrn.harvest(layers, n, chi, cutoff, ctx=gpu_ctx)
    assert X.shape == (Lrn.NFEAT, C.count_2q(layers))
    mlp = Lrn.RankMLP(Lrn.init_params(1), ctx=gpu_ctx)
    perm = np.random.default_rng(0).permutation(T.size)
    hist = mlp.train(X[:, perm], T[perm], steps=300, minibatch=64, lr=2e-3)
    assert hist[-10:].mean() < hist[:10].mean()
    ref = DeviceMPS(n, chi, gpu_ctx)
    ref.run_layers(layers, cutoff, chi, batched=True)
    dev = DeviceMPS(n, chi, gpu_ctx)
    stats = {}
    Lrn.run_layers_learned(dev, layers, cutoff, chi, mlp, stats=stats)
    assert stats["gates"] == C.count_2q(layers)
    assert infid(dev, ref) <= 1e-8
    mlp.free()
