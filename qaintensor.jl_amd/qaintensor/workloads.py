"""Synthetic networks of the contract() workloads.

qft_expectation_network builds the network of test/test_treewidth.jl:352-368
(`expectation_value`): a random ClosedMPS T0 of bond dimension 2 on N wires,
the circuit appended by tensor_circuit!, and a second copy of T0 contracted
with the circuit's open legs (no conjugation), so the network has no open
leg and contracts to the scalar T0^T U T0.  The circuit is the quantum
Fourier transform (circuits.qft_gates), the circuit behind the reference's
published contract() timings (examples/expectation_value_optimization_
example.ipynb).
"""
import numpy as np

from . import circuits as C
from .mpo import CircuitGate, tensor_circuit
from .mps import ClosedMPS
from .network import GeneralTensorNetwork, Summation, Tensor, shift_summation


def crand(rng, *dims):
    """rand(ComplexF64, dims...): real and imaginary parts uniform on
    [0, 1)."""
    return rng.random(dims) + 1j * rng.random(dims)


def qft_circuit(n):
    """The QFT gate list as CircuitGates (H, and controlled phases on two
    wires)."""
    out = []
    for op in C.qft_gates(n):
        if op[0] == "1q":
            out.append(CircuitGate((op[1],), op[2]))
        else:
            out.append(CircuitGate((op[1], op[2]), op[3]))
    return out


def expectation_network(n, gates, seed, is_decompose=False):
    """expectation_value(cgc) of test/test_treewidth.jl:354-368 for the
    CircuitGates `gates` on n wires, with seeded random T0 tensors.
    Returns (network, T0 tensors)."""
    rng = np.random.default_rng(seed)
    tensors = [Tensor(crand(rng, 2, 2))] + \
        [Tensor(crand(rng, 2, 2, 2)) for _ in range(2, n)] + \
        [Tensor(crand(rng, 2, 2))]
    T0 = ClosedMPS(tensors)
    T = GeneralTensorNetwork(list(T0.tensors), list(T0.contractions),
                             list(T0.openidx))
    tensor_circuit(T, gates, is_decompose=is_decompose)
    L0 = len(T.tensors)
    T.contractions = T.contractions + [shift_summation(c, L0)
                                       for c in T0.contractions]
    for i in range(n):
        T.tensors.append(T0.tensors[i])
        T.contractions.append(Summation([T.openidx[0],
                                         (len(T.tensors),
                                          T0.openidx[i][1])]))
        T.openidx.pop(0)
    return T, [t.data for t in tensors]


def qft_expectation_network(n, seed, is_decompose=False):
    return expectation_network(n, qft_circuit(n), seed, is_decompose)
