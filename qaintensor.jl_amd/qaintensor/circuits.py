"""Seeded Haar-random brick-wall circuits (the synthetic inputs of every
benchmark configuration).

Haar unitaries follow Mezzadri, "How to generate random matrices from the
classical compact groups", Notices AMS 54 (2007): Q R = Z for a complex
Ginibre matrix Z = (X + iY)/sqrt(2), then U = Q diag(r_ii / |r_ii|).

Layer l (0-based) applies a Haar U(2) on every wire 1..n, then a Haar U(4) on
the pairs (1,2),(3,4),... when l is even and (2,3),(4,5),... when l is odd.
Draw order per layer: the n single-qubit gates (wire order), then the
two-qubit gates (pair order), all from one numpy PCG64 generator seeded
once.  Each single-qubit gate is fused into the two-qubit gate that follows
it on its wire; unpaired wires keep their single-qubit gate.  A 4x4 gate
acts on the index s_w + 2 s_{w+1} (wire w <-> bit w-1), so the fused gate
is U4 @ kron(u_{w+1}, u_w).
"""
import numpy as np


def haar_unitary(rng, d):
    z = (rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))) \
        / np.sqrt(2.0)
    q, r = np.linalg.qr(z)
    ph = np.diag(r) / np.abs(np.diag(r))
    return q * ph[None, :]


def brickwall_layers(n, depth, seed):
    """[(singles, pairs)] per layer: singles = [(wire, u 2x2)] for wires not
    covered by a pair, pairs = [(wire, fused 4x4 gate on (wire, wire+1))]."""
    rng = np.random.default_rng(seed)
    layers = []
    for layer in range(depth):
        us = [haar_unitary(rng, 2) for _ in range(n)]
        start = 1 if layer % 2 == 0 else 2
        pair_wires = list(range(start, n, 2))
        gs = [haar_unitary(rng, 4) for _ in pair_wires]
        covered = set()
        pairs = []
        for w, g in zip(pair_wires, gs):
            pairs.append((w, g @ np.kron(us[w], us[w - 1])))
            covered.add(w)
            covered.add(w + 1)
        singles = [(w, us[w - 1]) for w in range(1, n + 1)
                   if w not in covered]
        layers.append((singles, pairs))
    return layers


def weak_brickwall_layers(n, depth, seed, tau=0.5):
    """The brick-wall pattern with weakly entangling pair gates: each pair
    gate is exp(-i tau H) for a GUE Hamiltonian H (4x4, (Z + Z^H)/2 with Z
    complex Ginibre) scaled to spectral norm 1, fused with Haar single-qubit
    gates as in brickwall_layers.  Its entanglement grows slowly, so at a
    relative cutoff the kept ranks stay well below the full rank (the
    regime where a rank-sized randomized SVD pays)."""
    rng = np.random.default_rng(seed)
    layers = []
    for layer in range(depth):
        us = [haar_unitary(rng, 2) for _ in range(n)]
        start = 1 if layer % 2 == 0 else 2
        pair_wires = list(range(start, n, 2))
        pairs = []
        covered = set()
        for w in pair_wires:
            z = (rng.standard_normal((4, 4)) + 1j * rng.standard_normal(
                (4, 4))) / np.sqrt(2.0)
            h = (z + z.conj().T) / 2
            e, v = np.linalg.eigh(h)
            e = e / np.max(np.abs(e))
            g = (v * np.exp(-1j * tau * e)[None, :]) @ v.conj().T
            pairs.append((w, g @ np.kron(us[w], us[w - 1])))
            covered.add(w)
            covered.add(w + 1)
        singles = [(w, us[w - 1]) for w in range(1, n + 1)
                   if w not in covered]
        layers.append((singles, pairs))
    return layers


def count_2q(layers):
    return sum(len(p) for _, p in layers)


def random_bitstrings(n, count, seed):
    """count x n array of 0/1 (uint8); column w-1 is wire w."""
    rng = np.random.default_rng(seed)
    return rng.integers(0, 2, size=(count, n), dtype=np.uint8)


def batch_circuits(count, seed, n_range=(16, 48), depth_range=(8, 24),
                   chis=(32, 64, 128)):
    """count independent seeded circuits for many-circuit batching:
    [(n, depth, chi_max, layers)], n and depth uniform on the inclusive
    ranges, chi_max uniform over chis, circuit i seeded by (seed, i)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(rng.integers(n_range[0], n_range[1] + 1))
        depth = int(rng.integers(depth_range[0], depth_range[1] + 1))
        chi = int(chis[int(rng.integers(0, len(chis)))])
        out.append((n, depth, chi, brickwall_layers(n, depth,
                                                    seed * 100003 + i)))
    return out


H_GATE = np.array([[1, 1], [1, -1]], dtype=np.complex128) / np.sqrt(2.0)


def qft_gates(n):
    """The quantum Fourier transform on wires 1..n (wire n the most
    significant) as a gate list, without the final bit reversal:
    [("1q", w, u)] and [("2q", i, j, g)] with g = diag(1, 1, 1, e^{i pi /
    2^(k-m)}) the controlled phase of wires k > m (symmetric in its two
    qubits, index s_i + 2 s_j).  Wire k meets every lower wire: the
    long-range gates of the circuits SWAP routing serves."""
    ops = []
    for k in range(n, 0, -1):
        ops.append(("1q", k, H_GATE))
        for m in range(k - 1, 0, -1):
            ph = np.exp(1j * np.pi / 2 ** (k - m))
            ops.append(("2q", k, m, np.diag([1, 1, 1, ph]).astype(
                np.complex128)))
    return ops
