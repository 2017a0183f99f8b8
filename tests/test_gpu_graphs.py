"""Jacobi sweeps replayed from cached hipGraphs (kernel nodes only) give
bit-identical results to the same sweeps launched one by one
(QTN_SWEEP_GRAPHS=0): same kernels, same order.  Each case runs graphs off,
then on twice (the first call captures, the second replays the cached
graph), with the program's own synchronisations between the calls."""
import numpy as np
import pytest

from qaintensor import DeviceMPO, DeviceMPS, FORM_CHOSEN, circuits as C, ops
from qaintensor import hamiltonians as Hm

pytestmark = pytest.mark.gpu


def run_three(monkeypatch, fn):
    monkeypatch.setenv("QTN_SWEEP_GRAPHS", "0")
    ref = fn()
    monkeypatch.setenv("QTN_SWEEP_GRAPHS", "1")
    first = fn()
    second = fn()
    return ref, first, second


def same(a, b):
    if isinstance(a, tuple):
        return all(same(x, y) for x, y in zip(a, b))
    return np.array_equal(np.asarray(a), np.asarray(b))


@pytest.mark.parametrize("form", [0, FORM_CHOSEN])
def test_batched_circuit_graphs_bit_identical(gpu_ctx, monkeypatch, form):
    n, chi = 14, 64
    layers = C.brickwall_layers(n, 10, 21)
    bits = C.random_bitstrings(n, 32, 3)

    def run():
        dev = DeviceMPS(n, chi, gpu_ctx)
        dev.run_layers(layers, 1e-10, chi, batched=True, form=form)
        return dev.amplitudes(bits), tuple(dev.bond_dims())
    ref, first, second = run_three(monkeypatch, run)
    assert same(ref, first) and same(ref, second)


def test_svd_graphs_bit_identical(gpu_ctx, monkeypatch):
    rng = np.random.default_rng(4)
    A = rng.standard_normal((300, 200)) + 1j * rng.standard_normal((300, 200))
    ref, first, second = run_three(
        monkeypatch, lambda: ops.svd_trunc(A, 1e-10, True, 120)[:3])
    assert same(ref, first) and same(ref, second)


def test_zipup_graphs_bit_identical(gpu_ctx, monkeypatch):
    n, chi = 12, 32
    W = Hm.tfim_mpo(n, 1.0, Hm.tfim_fields(n, 3))
    raw = Hm.random_mps_tensors(n, chi, 3)
    psi = DeviceMPS(n, chi, gpu_ctx)
    psi.set_sites(raw)
    mpo = DeviceMPO(W, gpu_ctx)
    bits = C.random_bitstrings(n, 16, 5)

    def run():
        out = DeviceMPS(n, chi, gpu_ctx)
        nrm = psi.apply_mpo(mpo, 1e-10, 24, out)
        return out.amplitudes(bits), np.float64(nrm)
    ref, first, second = run_three(monkeypatch, run)
    assert same(ref, first) and same(ref, second)
