"""The rank-predictor MLP's forward, backward and Adam kernels (FP32)
against the FP64 oracle (oracle.mlp_forward / mlp_loss_grad)."""
import numpy as np
import pytest

from oracle import qtn_oracle as O
from qaintensor import learned as Lrn

pytestmark = pytest.mark.gpu


def data(seed, nb):
    rng = np.random.default_rng(seed)
    X = np.zeros((Lrn.NFEAT, nb), dtype=np.float32)
    # spectra: descending, normalised, some zero padding
    for b in range(nb):
        k = int(rng.integers(1, 65))
        s = np.sort(rng.random(k))[::-1] ** 3
        X[:k, b] = s / np.linalg.norm(s)
    X[64:68] = rng.random((4, nb)) * 2
    X[68] = rng.random(nb)
    T = rng.random(nb).astype(np.float32)
    return X, T


def test_mlp_forward_and_gradient_match_oracle(gpu_ctx):
    p = Lrn.init_params(3)
    m = Lrn.RankMLP(p, ctx=gpu_ctx)
    X, T = data(1, 300)
    y = m.forward(X)
    y_ref = O.mlp_forward(p.astype(np.float64), X.astype(np.float64),
                          Lrn.NFEAT, Lrn.WIDTH)[0]
    assert np.abs(y - y_ref).max() <= 1e-4 * max(1.0, np.abs(y_ref).max())
    loss, g = m.loss_grad(X, T)
    l_ref, g_ref = O.mlp_loss_grad(p.astype(np.float64), X.astype(np.float64),
                                   T.astype(np.float64), Lrn.NFEAT, Lrn.WIDTH)
    assert abs(loss - l_ref) <= 1e-4 * l_ref
    assert np.linalg.norm(g - g_ref) <= 1e-3 * np.linalg.norm(g_ref)
    m.free()


def test_mlp_training_fits_a_rank_function(gpu_ctx):
    """Adam on the GPU learns t = (length of the spectrum) / 64 from the
    spectrum features (zero padding enters as log10(1e-30))."""
    X, _ = data(2, 4096)
    T = (X[:64] > 0).sum(axis=0).astype(np.float32) / 64.0
    m = Lrn.RankMLP(Lrn.init_params(5), ctx=gpu_ctx)
    base = float(np.mean((T - T.mean()) ** 2))
    hist = m.train(X, T, steps=600, minibatch=512, lr=2e-3)
    assert hist[-20:].mean() < 0.2 * base
    # the trained parameters reproduce the final loss through the oracle
    p = m.params().astype(np.float64)
    y = O.mlp_forward(p, X.astype(np.float64), Lrn.NFEAT, Lrn.WIDTH)[0]
    assert np.mean((y - T) ** 2) < 0.2 * base
    m.free()
