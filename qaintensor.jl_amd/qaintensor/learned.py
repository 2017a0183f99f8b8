"""Learned rank predictor: a 3-layer FP32 MLP (width 128, GELU) trained and
evaluated on the GPU by libqtn's own kernels (qtn_mlp_*).

Features of one two-site update on bond i (column of 72 floats):
  0..63  the bond's spectrum before the gate (lambda_i, normalised,
         descending, zero padded; the library takes log10(max(x,1e-16))/8)
  64..67 the gate's operator-Schmidt coefficients (singular values of the
         4x4 gate reshaped as (s_i s_i') x (s_{i+1} s_{i+1}'), src/mpo.jl:
         40-56 split, computed on the GPU by qtn_svd)
  68     the bond position i / n
  69..71 the bond dimensions chi_{i-1}, chi_i, chi_{i+1} / chi_max (the
         shape of the two-site matrix and the bond's rank before the gate;
         the 64 spectrum values alone do not show ranks above 64)
Target: the kept rank at the cutoff divided by chi_max.
"""
import ctypes

import numpy as np

from . import _lib as L
from . import ops

NFEAT = 72
WIDTH = 128


def nparams(nin=NFEAT, width=WIDTH):
    return width * nin + width + width * width + width + width + 1


def init_params(seed, nin=NFEAT, width=WIDTH):
    """He-style Gaussian weights, zero biases (float32)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    W1 = rng.standard_normal((width, nin)) * np.sqrt(2.0 / nin)
    W2 = rng.standard_normal((width, width)) * np.sqrt(2.0 / width)
    W3 = rng.standard_normal((1, width)) * np.sqrt(1.0 / width)
    p = np.concatenate([W1.reshape(-1, order="F"), np.zeros(width),
                        W2.reshape(-1, order="F"), np.zeros(width),
                        W3.reshape(-1, order="F"), np.zeros(1)])
    return p.astype(np.float32)


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class RankMLP:
    def __init__(self, params, nin=NFEAT, width=WIDTH, ctx=None):
        self.ctx = ctx if ctx is not None else L.default_context()
        self.nin, self.width = nin, width
        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.size != nparams(nin, width):
            raise ValueError("wrong number of parameters")
        h = ctypes.c_void_p()
        L.check(L.load().qtn_mlp_create(self.ctx.handle, nin, width, _fp(p),
                                        ctypes.byref(h)))
        self._h = h
        self.ctx.adopt(self)

    @property
    def handle(self):
        return self._h

    def params(self):
        out = np.zeros(nparams(self.nin, self.width), dtype=np.float32)
        L.check(L.load().qtn_mlp_get_params(self._h, _fp(out)))
        return out

    def forward(self, X):
        X = np.asfortranarray(X, dtype=np.float32)
        nb = X.shape[1]
        Y = np.zeros(nb, dtype=np.float32)
        L.check(L.load().qtn_mlp_forward(self._h, nb, _fp(X), _fp(Y)))
        return Y

    def loss_grad(self, X, T):
        X = np.asfortranarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        g = np.zeros(nparams(self.nin, self.width), dtype=np.float32)
        L.check(L.load().qtn_mlp_loss_grad(self._h, X.shape[1], _fp(X),
                                           _fp(T), _fp(loss), _fp(g)))
        return float(loss[0]), g

    def train(self, X, T, steps, minibatch=1024, lr=1e-3):
        X = np.asfortranarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        hist = np.zeros(max(1, steps), dtype=np.float32)
        L.check(L.load().qtn_mlp_train(self._h, X.shape[1], _fp(X), _fp(T),
                                       int(steps), int(minibatch),
                                       ctypes.c_float(lr), _fp(hist)))
        return hist[:steps]

    def free(self):
        if self._h is not None:
            L.check(L.load().qtn_mlp_destroy(self._h))
            self._h = None


def schmidt_coefficients(g):
    """Operator-Schmidt coefficients of a 4x4 gate (index s_i + 2 s_{i+1})
    on the GPU: singular values of the (s_i, s_i') x (s_{i+1}, s_{i+1}')
    rearrangement."""
    t = np.asarray(g, dtype=np.complex128).reshape((2, 2, 2, 2), order="F")
    m = ops.permute(t, [1, 3, 2, 4]).reshape((4, 4), order="F")
    return ops.svd_values(m)


def features(spectrum, schmidt, bond, n, dims, chimax):
    """Feature column of the gate on wires (bond, bond + 1) from host data;
    dims = the state's bond dimensions chi_0..chi_n.  The library gathers
    the same column on the device (qtn_mps_gate_features)."""
    x = np.zeros(NFEAT, dtype=np.float32)
    s = np.asarray(spectrum, dtype=np.float64)[:64]
    x[:s.size] = s
    x[64:68] = np.asarray(schmidt, dtype=np.float64)[:4]
    x[68] = bond / float(n)
    x[69:72] = np.asarray(dims[bond - 1:bond + 2], dtype=np.float64) / chimax
    return x


def gate_schmidt_batch(gates, ctx=None):
    """Operator-Schmidt coefficients (len(gates) x 4) of many 4x4 gates in
    one launch (qtn_gate_schmidt)."""
    ctx = ctx if ctx is not None else L.default_context()
    n = len(gates)
    if n == 0:
        return np.zeros((0, 4))
    g = np.ascontiguousarray(np.stack(
        [np.asarray(x, dtype=np.complex128).reshape(-1, order="F")
         for x in gates]))
    out = np.zeros((n, 4), dtype=np.float64)
    L.check(L.load().qtn_gate_schmidt(
        ctx.handle, n,
        g.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
        L.dptr(out)))
    return out


def gate_schmidt_table(layers, ctx=None):
    """Operator-Schmidt coefficients of every two-qubit gate of a circuit,
    computed once on the GPU in one launch (the circuit's compile step):
    per layer an array (gates x 4)."""
    flat = [g for (_, pairs) in layers for (_, g) in pairs]
    allc = gate_schmidt_batch(flat, ctx)
    out, i = [], 0
    for (_, pairs) in layers:
        out.append(allc[i:i + len(pairs)])
        i += len(pairs)
    return out


def harvest(layers, n, chimax, cutoff, ctx=None, dev=None):
    """(X, T) training pairs from one plain FP64 run: for every gate the
    features of its bond before the gate's layer and the kept rank /
    chimax.  The gates of a layer are disjoint, so each layer is one batched
    qtn_mps_apply_layer and the bond spectra read before it are the ones
    each gate meets."""
    from .device import DeviceMPS
    dev = dev if dev is not None else DeviceMPS(n, chimax, ctx)
    sch = gate_schmidt_table(layers)
    X, T = [], []
    for li, (singles, pairs) in enumerate(layers):
        for (w, u) in singles:
            dev.apply_1q(w, u)
        if not pairs:
            continue
        X.append(dev.gate_features([w for (w, _) in pairs],
                                   sch[li], chimax))
        k, _ = dev.apply_layer([w for (w, _) in pairs],
                               [g for (_, g) in pairs], cutoff, chimax)
        T.extend((k / float(chimax)).tolist())
    return np.concatenate(X, axis=1), np.asarray(T, dtype=np.float32)


def run_layers_learned(dev, layers, cutoff, chimax, mlp, sch=None,
                       margin=0.0, stats=None, hint_rule=None):
    """Run a circuit in the learned form: per layer, the features of every
    gate's bond go through the MLP on the GPU, the predicted rank
    (ceil((y + margin) * chimax)) sizes the randomized range finder, and
    gates whose check fails fall back to the full SVD.

    mlp None: no predictor; hint_rule (None or "previous") then sizes the
    sketch without the network: None tries no sketch (the chosen form),
    "previous" uses the bond's rank before the gate.  stats (dict, may be
    None) accumulates gates / tried (sketches run) / accepted counts."""
    if sch is None and mlp is not None:
        sch = gate_schmidt_table(layers)
    for li, (singles, pairs) in enumerate(layers):
        for (w, u) in singles:
            dev.apply_1q(w, u)
        if not pairs:
            continue
        sites = [w for (w, _) in pairs]
        gates = [g for (_, g) in pairs]
        if mlp is not None:
            # features, inference and hints inside the library
            _, _, acc, _ = dev.apply_layer_predicted(
                sites, gates, sch[li], cutoff, chimax, mlp, margin)
        elif hint_rule == "previous":
            d = dev.bond_dims()
            hints = np.asarray([d[w] for w in sites], dtype=np.int64)
        else:
            hints = np.zeros(len(pairs), dtype=np.int64)
        if mlp is None:
            _, _, acc = dev.apply_layer_learned(sites, gates, cutoff, chimax,
                                                hints)
        if stats is not None:
            stats["gates"] = stats.get("gates", 0) + len(pairs)
            stats["tried"] = stats.get("tried", 0) + int((acc != 0).sum())
            stats["accepted"] = stats.get("accepted", 0) + int(
                (acc == 1).sum())
