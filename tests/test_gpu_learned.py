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
    # gates whose sketch (hint + 16) is narrower than the cap and at most
    # 3/4 of min(M, N)
    big = [i for i, (w, _) in enumerate(pairs)
           if hint + 16 < 256
           and 4 * (hint + 16) <= 3 * min(2 * d0[w - 1], 2 * d0[w + 1])]
    if expect_accept:
        assert len(big) > 0 and all(acc[i] == 1 for i in big)
        # exact low rank: the same kept ranks as the full SVD
        assert list(k) == list(k_ref)
    else:
        assert (acc == 1).sum() == 0
        # a sketch was tried exactly where the hint is usable
        tried = [i for i in range(len(pairs)) if acc[i] == -1]
        assert tried == (big if hint > 0 else [])
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
    # gates at the cap take the FP32 sweeps of the chosen form
    assert infid(dev, ref) <= 1e-7
    mlp.free()


def test_trained_network_beats_rank_sizing_without_it(gpu_ctx):
    """What the network gains.  On the weakly entangling brick-wall (tau
    0.5, kept ranks cutoff-bound at about half the full rank and growing
    ~10% per layer) the sketch sized by the bond's previous rank misses the
    growth and falls back, while the trained network predicts it: more
    gates are decided by the sketch and fewer sketches are wasted.  Both
    end at the plain state."""
    n, depth, chi, cutoff = 24, 20, 256, 1e-8
    Xs, Ts = [], []
    for seed in (1, 2, 3):
        x, t = Lrn.harvest(C.weak_brickwall_layers(n, depth, seed), n, chi,
                           cutoff, ctx=gpu_ctx)
        Xs.append(x)
        Ts.append(t)
    X, T = np.concatenate(Xs, axis=1), np.concatenate(Ts)
    mlp = Lrn.RankMLP(Lrn.init_params(2), ctx=gpu_ctx)
    perm = np.random.default_rng(0).permutation(T.size)
    mlp.train(X[:, perm], T[perm], steps=1500, minibatch=128, lr=2e-3)
    layers = C.weak_brickwall_layers(n, depth, 9)
    ref = DeviceMPS(n, chi, gpu_ctx)
    ref.run_layers(layers, cutoff, chi, batched=True)
    res = {}
    for name, net, rule in (("mlp", mlp, None), ("previous", None,
                                                 "previous"),
                            ("none", None, None)):
        dev = DeviceMPS(n, chi, gpu_ctx)
        st = {}
        Lrn.run_layers_learned(dev, layers, cutoff, chi, net, margin=0.02,
                               stats=st, hint_rule=rule)
        assert infid(dev, ref) <= 1e-7, name
        res[name] = st
    print(res)
    assert res["none"]["tried"] == 0 and res["none"]["accepted"] == 0
    m, p = res["mlp"], res["previous"]
    assert m["accepted"] > p["accepted"]
    assert m["accepted"] >= 0.1 * m["gates"]
    fb_m = (m["tried"] - m["accepted"]) / max(1, m["tried"])
    fb_p = (p["tried"] - p["accepted"]) / max(1, p["tried"])
    assert fb_m < fb_p
    mlp.free()


def test_gate_schmidt_matches_numpy(gpu_ctx):
    rng = np.random.default_rng(4)
    gates = [C.haar_unitary(rng, 4) for _ in range(40)]
    gates += [np.kron(C.haar_unitary(rng, 2), C.haar_unitary(rng, 2))
              for _ in range(10)]
    gates += [p for (_, pairs) in C.weak_brickwall_layers(12, 3, 5)
              for (_, p) in pairs]
    gates.append(np.eye(4))
    got = Lrn.gate_schmidt_batch(gates, gpu_ctx)
    for g, row in zip(gates, got):
        m = g.reshape((2, 2, 2, 2), order="F").transpose(0, 2, 1, 3) \
            .reshape((4, 4), order="F")
        ref = np.linalg.svd(m, compute_uv=False)
        assert np.max(np.abs(row - ref)) <= 1e-12
    # a product gate has one coefficient; the identity's is 2
    assert np.allclose(got[40:50, 1:], 0, atol=1e-12)
    assert np.allclose(got[-1], [2, 0, 0, 0], atol=1e-12)


def test_gate_features_match_host(gpu_ctx):
    """The device-gathered predictor input equals the host restatement
    (learned.features) column by column."""
    n, chi = 14, 48
    layers = C.brickwall_layers(n, 7, 6)
    dev = DeviceMPS(n, chi, gpu_ctx)
    dev.run_layers(layers, 1e-8, chi, batched=True)
    sites = [1, 4, 7, 13]
    sch = Lrn.gate_schmidt_batch([C.haar_unitary(np.random.default_rng(i),
                                                 4) for i in range(4)],
                                 gpu_ctx)
    got = dev.gate_features(sites, sch, chi)
    d = dev.bond_dims()
    for i, w in enumerate(sites):
        ref = Lrn.features(dev.bond_spectrum(w), sch[i], w, n, d, chi)
        assert np.array_equal(got[:, i], ref)


def test_apply_layer_predicted_matches_host_hints(gpu_ctx):
    """qtn_mps_apply_layer_predicted = host features -> qtn_mlp_forward ->
    hints -> qtn_mps_apply_layer_learned."""
    n, chi, cutoff, margin = 16, 64, 1e-8, 0.02
    layers = C.weak_brickwall_layers(n, 12, 3)
    a = DeviceMPS(n, chi, gpu_ctx)
    a.run_layers(layers[:-1], cutoff, chi, batched=True)
    b = DeviceMPS(n, chi, gpu_ctx)
    b.copy_from(a)
    mlp = Lrn.RankMLP(Lrn.init_params(3), ctx=gpu_ctx)
    singles, pairs = layers[-1]
    for (w, u) in singles:
        a.apply_1q(w, u)
        b.apply_1q(w, u)
    sites = [w for (w, _) in pairs]
    gates = [g for (_, g) in pairs]
    sch = Lrn.gate_schmidt_batch(gates, gpu_ctx)
    X = np.stack([Lrn.features(a.bond_spectrum(w), sch[i], w, n,
                               a.bond_dims(), chi)
                  for i, w in enumerate(sites)], axis=1)
    y = mlp.forward(X).astype(np.float64)
    hints = np.clip(np.ceil((y + margin) * chi), 0, chi).astype(np.int64)
    ka, _, acc_a = a.apply_layer_learned(sites, gates, cutoff, chi, hints)
    kb, _, acc_b, hb = b.apply_layer_predicted(sites, gates, sch, cutoff, chi,
                                               mlp, margin)
    assert np.array_equal(hb, hints)
    # the same gates try a sketch (each call draws its own Gaussian sketch,
    # so borderline accept decisions may differ, the states may not)
    assert np.array_equal(acc_a != 0, acc_b != 0)
    assert infid(a, b) <= 1e-9
    mlp.free()
