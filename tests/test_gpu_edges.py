"""Edge cases of the device MPS path (qtn_mps_*): the smallest chain, the
boundary sites, a bond cap of 1, an empty layer, and the error statuses
for a zero gate, a non-finite gate, a site out of range and a cap above
the state's capacity — each error leaves the state as it was — and the
teardown order of a context closed while its objects are alive."""
import numpy as np
import pytest

from oracle import qtn_oracle as O
from qaintensor import DeviceMPS, FORM_CHOSEN, circuits as C
from qaintensor._lib import (QTN_ERR_ARG, QTN_ERR_NOTFINITE, QTN_ERR_TRUNC,
                             QtnError)

pytestmark = pytest.mark.gpu


def all_bits(n):
    return np.array([[(i >> w) & 1 for w in range(n)]
                     for i in range(2 ** n)], dtype=np.uint8)


def assert_same_state(dev, st, tol=1e-12):
    bits = all_bits(dev.n)
    a = dev.amplitudes(bits)
    b = st.amplitudes(bits)
    assert np.max(np.abs(a - b)) <= tol * max(1.0, np.max(np.abs(b)))


def test_two_site_chain(gpu_ctx):
    rng = np.random.default_rng(1)
    dev = DeviceMPS(2, 4, gpu_ctx)
    st = O.TebdState(2)
    for _ in range(3):
        g = C.haar_unitary(rng, 4)
        k, _ = dev.apply_2q(1, g, 1e-12, 4)
        k_ref, _, _ = st.apply_2q(1, g, 1e-12, 4)
        assert k == k_ref
    assert_same_state(dev, st)
    assert list(dev.bond_dims()) == st.bond_dims()


def test_boundary_sites_and_cap_one(gpu_ctx):
    """Gates on the first and last pair of a 6-site chain, then a layer
    truncated to bond 1 (kept rank 1 everywhere, product state)."""
    rng = np.random.default_rng(2)
    n = 6
    dev = DeviceMPS(n, 8, gpu_ctx)
    st = O.TebdState(n)
    for w in (1, n - 1, 1, n - 1):
        g = C.haar_unitary(rng, 4)
        dev.apply_2q(w, g, 1e-12, 8)
        st.apply_2q(w, g, 1e-12, 8)
    assert_same_state(dev, st)
    gates = [C.haar_unitary(rng, 4) for _ in range(3)]
    k, _ = dev.apply_layer([1, 3, 5], gates, 1e-12, 1)
    for w, g in zip([1, 3, 5], gates):
        st.apply_2q(w, g, 1e-12, 1)
    assert list(k) == [1, 1, 1]
    assert_same_state(dev, st, 1e-11)


def test_empty_layer_is_a_no_op(gpu_ctx):
    dev = DeviceMPS(4, 4, gpu_ctx)
    dev.apply_2q(2, C.haar_unitary(np.random.default_rng(3), 4), 1e-12, 4)
    before = dev.amplitudes(all_bits(4))
    k, d = dev.apply_layer([], [], 1e-10, 4)
    assert k.size == 0 and d.size == 0
    assert np.array_equal(dev.amplitudes(all_bits(4)), before)


@pytest.mark.parametrize("form", [0, FORM_CHOSEN])
def test_zero_gate_reports_truncation_and_keeps_state(gpu_ctx, form):
    """Every singular value of a zero two-site tensor is within the cutoff:
    QTN_ERR_TRUNC (the reference fails in findfirst there), state intact."""
    dev = DeviceMPS(4, 4, gpu_ctx)
    dev.apply_2q(2, C.haar_unitary(np.random.default_rng(4), 4), 1e-12, 4)
    before = dev.amplitudes(all_bits(4))
    with pytest.raises(QtnError) as e:
        dev.apply_2q(2, np.zeros((4, 4)), 1e-10, 4, form=form)
    assert e.value.status == QTN_ERR_TRUNC
    assert np.array_equal(dev.amplitudes(all_bits(4)), before)


def test_non_finite_gate_is_rejected(gpu_ctx):
    dev = DeviceMPS(4, 4, gpu_ctx)
    before = dev.amplitudes(all_bits(4))
    g = np.eye(4, dtype=np.complex128)
    g[1, 2] = np.nan
    with pytest.raises(QtnError) as e:
        dev.apply_2q(1, g, 1e-10, 4)
    assert e.value.status == QTN_ERR_NOTFINITE
    assert np.array_equal(dev.amplitudes(all_bits(4)), before)


@pytest.mark.parametrize("form", [0, FORM_CHOSEN])
def test_non_finite_gate_in_a_mixed_layer_leaves_state(gpu_ctx, form):
    """A layer whose gates fall in two size groups (two-site matrices of
    at most 64 and of 128 columns), one gate non-finite: refused with
    QTN_ERR_NOTFINITE before any group runs, so no gate of the layer is
    applied."""
    rng = np.random.default_rng(6)
    n = 14
    dev = DeviceMPS(n, 64, gpu_ctx)
    for layer in range(8):
        ws = list(range(1 + layer % 2, n, 2))
        dev.apply_layer(ws, [C.haar_unitary(rng, 4) for _ in ws], 1e-12, 64)
    dims = list(dev.bond_dims())
    assert dims[1] == 2 and dims[7] == 64
    before = dev.amplitudes(all_bits(n))
    ws = list(range(1, n, 2))
    gates = [C.haar_unitary(rng, 4) for _ in ws]
    gates[3] = gates[3].copy()
    gates[3][0, 1] = np.inf
    with pytest.raises(QtnError) as e:
        dev.apply_layer(ws, gates, 1e-12, 64, form=form)
    assert e.value.status == QTN_ERR_NOTFINITE
    assert list(dev.bond_dims()) == dims
    assert np.array_equal(dev.amplitudes(all_bits(n)), before)


@pytest.mark.parametrize("site", [0, 4, -1])
def test_site_out_of_range(gpu_ctx, site):
    dev = DeviceMPS(4, 4, gpu_ctx)
    with pytest.raises(QtnError) as e:
        dev.apply_2q(site, np.eye(4), 1e-10, 4)
    assert e.value.status == QTN_ERR_ARG


def test_cap_above_capacity_is_rejected(gpu_ctx):
    dev = DeviceMPS(6, 4, gpu_ctx)
    with pytest.raises(QtnError) as e:
        dev.apply_layer([1, 3], [np.eye(4), np.eye(4)], 1e-10, 8)
    assert e.value.status == QTN_ERR_ARG


def test_context_close_frees_live_objects():
    """Closing a context while states, an MPO and an MLP made on it are
    alive frees them first (their destroy calls use the context); their
    handles are then gone, and a new context works normally."""
    from qaintensor import Context, DeviceMPO, learned
    from qaintensor import hamiltonians as Hm
    ctx = Context(0)
    dev = DeviceMPS(6, 8, ctx)
    dev.apply_2q(3, C.haar_unitary(np.random.default_rng(5), 4), 1e-12, 8)
    mpo = DeviceMPO(Hm.tfim_mpo(6, 1.0, Hm.tfim_fields(6, 1)), ctx)
    mlp = learned.RankMLP(learned.init_params(0), ctx=ctx)
    ctx.close()
    for obj in (dev, mpo):
        with pytest.raises(QtnError) as e:
            obj.handle
        assert e.value.status == QTN_ERR_ARG
    assert mlp.handle is None
    dev.free()  # already freed: no-op
    ctx2 = Context(0)
    try:
        fresh = DeviceMPS(4, 4, ctx2)
        fresh.apply_2q(2, np.eye(4), 1e-12, 4)
        assert list(fresh.bond_dims()) == [1, 1, 1, 1, 1]
        fresh.free()
    finally:
        ctx2.close()
