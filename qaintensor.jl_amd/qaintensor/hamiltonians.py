"""Operators and synthetic states for the MPO x MPS path.

tfim_mpo builds the transverse-field Ising Hamiltonian
    H = -J sum_i Z_i Z_{i+1} - sum_i h_i X_i
as an MPO of bond dimension 3 in the layout of qtn_mpo_create: site tensors
(D_j, 2 out, 2 in, D_{j+1}), bulk operator matrix (row = left bond)
    [[I, 0, 0], [Z, 0, 0], [-h X, -J Z, I]],
the first site its last row and the last site its first column.
"""
import numpy as np

_I2 = np.eye(2, dtype=np.complex128)
_X = np.array([[0, 1], [1, 0]], dtype=np.complex128)
_Z = np.diag([1.0, -1.0]).astype(np.complex128)


def tfim_mpo(n, J, h):
    h = np.asarray(h, dtype=np.float64)
    if h.shape != (n,):
        raise ValueError("need one field per site")
    out = []
    for j in range(n):
        W = np.zeros((3, 2, 2, 3), dtype=np.complex128)
        W[0, :, :, 0] = _I2
        W[1, :, :, 0] = _Z
        W[2, :, :, 0] = -h[j] * _X
        W[2, :, :, 1] = -J * _Z
        W[2, :, :, 2] = _I2
        if j == 0:
            W = W[2:3]
        if j == n - 1:
            W = W[:, :, :, 0:1]
        out.append(np.asfortranarray(W))
    return out


def tfim_fields(n, seed, lo=0.5, hi=1.5):
    """Seeded transverse fields h_i ~ U[lo, hi] (numpy PCG64)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return rng.uniform(lo, hi, n)


def random_mps_tensors(n, chimax, seed):
    """Complex Gaussian site tensors (chi_j, 2, chi_{j+1}) with bonds
    min(chimax, 2^j, 2^(n-j)), entries (N(0,1) + i N(0,1)) / sqrt(2);
    DeviceMPS.set_sites brings them to canonical form on the GPU."""
    rng = np.random.Generator(np.random.PCG64(seed))
    dims = [min(chimax, 2 ** min(j, 62), 2 ** min(n - j, 62))
            for j in range(n + 1)]
    out = []
    for j in range(n):
        a, b = dims[j], dims[j + 1]
        t = (rng.standard_normal((a, 2, b)) +
             1j * rng.standard_normal((a, 2, b))) / np.sqrt(2.0)
        out.append(np.asfortranarray(t))
    return out
