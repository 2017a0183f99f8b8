"""Device-resident MPS driven through libqtn (qtn_mps_* in include/qtn.h).

DeviceMPS keeps the state on the GPU in the right-canonical gauge with the
bond spectra beside the site tensors.  Only 2x2 / 4x4 gates and scalars
cross the ABI per gate.  This is the hot path behind the build's
apply!/expect/amplitude methods (the reference has the two-site split only
as switch!, src/switch.jl:18-56, and no truncating update).
"""
import ctypes

import numpy as np

from . import _lib as L

FORM_PLAIN = L.QTN_FORM_PLAIN
FORM_CHOSEN = L.QTN_FORM_CHOSEN


class DeviceMPS:
    """|0...0> of `nsites` qubits with per-bond capacity `chimax_cap`."""

    def __init__(self, nsites, chimax_cap, ctx=None):
        self.ctx = ctx if ctx is not None else L.default_context()
        self.n = int(nsites)
        self.cap = int(chimax_cap)
        h = ctypes.c_void_p()
        L.check(L.load().qtn_mps_create(self.ctx.handle, self.n, self.cap,
                                        ctypes.byref(h)))
        self._h = h
        self.ctx.adopt(self)

    @property
    def handle(self):
        if self._h is None:
            raise L.QtnError(L.QTN_ERR_ARG, "DeviceMPS already freed")
        return self._h

    def free(self):
        if self._h is not None:
            L.check(L.load().qtn_mps_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            if self._h is not None and L._lib is not None:
                L._lib.qtn_mps_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def copy_from(self, other):
        L.check(L.load().qtn_mps_copy(other.handle, self.handle))

    # -- gates -------------------------------------------------------------
    def apply_1q(self, site, u):
        u = L.as_c128(u, "F")
        if u.shape != (2, 2):
            raise ValueError("single-qubit gate must be 2x2")
        L.check(L.load().qtn_mps_apply_1q(self.handle, int(site), L.dptr(u)))

    def apply_1q_layer(self, singles):
        """Single-qubit gates [(site, u)]: one launch when the sites are
        distinct (qtn_mps_apply_1q_multi), else one call per gate in order
        (gates on one site do not commute)."""
        if not singles:
            return
        sites = [int(w) for w, _ in singles]
        if len(set(sites)) == len(sites):
            apply_1q_multi([self] * len(sites), sites, [u for _, u in singles])
            return
        for (w, u) in singles:
            self.apply_1q(w, u)

    def apply_2q(self, site, g, cutoff, chimax, form=FORM_PLAIN,
                 want_spectrum=False):
        """Gate g (4x4, index s_site + 2 s_{site+1}) on (site, site+1).
        Returns (kept rank, discarded weight[, full spectrum])."""
        g = L.as_c128(g, "F")
        if g.shape != (4, 4):
            raise ValueError("two-qubit gate must be 4x4")
        kept = ctypes.c_int64(0)
        disc = ctypes.c_double(0.0)
        spec = None
        slen = ctypes.c_int64(0)
        if want_spectrum:
            spec = np.zeros(2 * self.cap + 2, dtype=np.float64)
        L.check(L.load().qtn_mps_apply_2q(
            self.handle, int(site), L.dptr(g), float(cutoff), int(chimax),
            int(form), ctypes.byref(kept), ctypes.byref(disc),
            L.dptr(spec) if spec is not None else None,
            ctypes.byref(slen)))
        if want_spectrum:
            return kept.value, disc.value, spec[:slen.value].copy()
        return kept.value, disc.value

    SWAP = np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]],
                    dtype=np.complex128)

    def apply_2q_long(self, i, j, g, cutoff, chimax, form=FORM_PLAIN):
        """Gate g (4x4, index s_i + 2 s_j) on any two sites i != j: SWAP
        routing of the farther qubit next to the nearer one, the gate, and
        the way back, each step a truncated two-site update
        (qtn_mps_apply_2q_long).  Returns (kept rank, discarded weight) of
        the gate's own update."""
        g = L.as_c128(g, "F")
        if g.shape != (4, 4):
            raise ValueError("two-qubit gate must be 4x4")
        kept = ctypes.c_int64(0)
        disc = ctypes.c_double(0.0)
        L.check(L.load().qtn_mps_apply_2q_long(
            self.handle, int(i), int(j), L.dptr(g), float(cutoff),
            int(chimax), int(form), ctypes.byref(kept), ctypes.byref(disc)))
        return kept.value, disc.value

    def switch(self, i, j=None):
        """switch! on the device state (src/switch.jl): switch(i, j) swaps
        wires i and j, numbered from the most significant as the reference
        does (qtn_mps_switch); switch([(i, j), ...]) swaps pair by pair
        (src/switch.jl:6-11); switch(i) swaps sites i and i + 1
        (src/switch.jl:18-56).  Untruncated but for exact zeros, up to the
        state's capacity."""
        if j is None and isinstance(i, (list, tuple)):
            for (a, b) in i:
                L.check(L.load().qtn_mps_switch(self.handle, int(a), int(b)))
            return self
        if j is None:
            self.apply_2q(int(i), self.SWAP, 0.0, self.cap)
            return self
        L.check(L.load().qtn_mps_switch(self.handle, int(i), int(j)))
        return self

    def permute(self, order):
        """Base.permute!(mps, order) on the device state
        (src/switch.jl:94-106; qtn_mps_permute)."""
        o = np.ascontiguousarray(order, dtype=np.int64)
        L.check(L.load().qtn_mps_permute(self.handle, L.i64ptr(o), o.size))
        return self

    def apply_layer(self, sites, gates, cutoff, chimax, form=FORM_PLAIN):
        """Gates on disjoint pairs in one batched pass.  Returns (kept[],
        disc[])."""
        ng = len(sites)
        if ng == 0:
            return np.zeros(0, np.int64), np.zeros(0)
        s = np.ascontiguousarray(sites, dtype=np.int32)
        g = np.empty((ng, 16), dtype=np.complex128)
        for i, gi in enumerate(gates):
            gi = np.asarray(gi, dtype=np.complex128)
            if gi.shape != (4, 4):
                raise ValueError("two-qubit gate must be 4x4")
            g[i] = gi.reshape(-1, order="F")
        kept = np.zeros(ng, dtype=np.int64)
        disc = np.zeros(ng, dtype=np.float64)
        L.check(L.load().qtn_mps_apply_layer(
            self.handle, ng, L.i32ptr(s), L.dptr(g), float(cutoff),
            int(chimax), int(form), L.i64ptr(kept), L.dptr(disc)))
        return kept, disc

    def run_layers(self, layers, cutoff, chimax, form=FORM_PLAIN,
                   batched=True, record=None):
        """Apply brick-wall layers ([(singles, pairs)]).  batched=False
        applies the two-qubit gates one at a time (the plain per-gate
        form)."""
        for singles, pairs in layers:
            self.apply_1q_layer(singles)
            if not pairs:
                continue
            if batched:
                k, d = self.apply_layer([w for w, _ in pairs],
                                        [g for _, g in pairs], cutoff,
                                        chimax, form)
                if record is not None:
                    for (w, _), ki, di in zip(pairs, k, d):
                        record.append((w, int(ki), float(di)))
            else:
                for (w, g) in pairs:
                    ki, di = self.apply_2q(w, g, cutoff, chimax, form)
                    if record is not None:
                        record.append((w, ki, di))
        return self

    # -- observables -------------------------------------------------------
    def bond_dims(self):
        d = np.zeros(self.n + 1, dtype=np.int64)
        L.check(L.load().qtn_mps_bond_dims(self.handle, L.i64ptr(d)))
        return d

    def bond_spectrum(self, bond):
        S = np.zeros(self.cap + 1, dtype=np.float64)
        k = ctypes.c_int64(0)
        L.check(L.load().qtn_mps_bond_spectrum(self.handle, int(bond),
                                               L.dptr(S), ctypes.byref(k)))
        return S[:k.value].copy()

    def site(self, j):
        dims = np.zeros(3, dtype=np.int64)
        L.check(L.load().qtn_mps_get_site(self.handle, int(j), None,
                                          L.i64ptr(dims)))
        out = np.zeros(int(np.prod(dims)), dtype=np.complex128)
        L.check(L.load().qtn_mps_get_site(self.handle, int(j), L.dptr(out),
                                          L.i64ptr(dims)))
        return out.reshape(tuple(dims), order="F")

    def amplitudes(self, bits):
        bits = np.ascontiguousarray(bits, dtype=np.uint8)
        if bits.ndim != 2 or bits.shape[1] != self.n:
            raise ValueError("bits must be (count, nsites)")
        out = np.zeros(bits.shape[0], dtype=np.complex128)
        L.check(L.load().qtn_mps_amplitudes(
            self.handle, bits.shape[0],
            bits.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
            L.dptr(out)))
        return out

    def expect_z(self):
        out = np.zeros(self.n, dtype=np.float64)
        L.check(L.load().qtn_mps_expect_z(self.handle, L.dptr(out)))
        return out

    def overlap(self, other):
        """<self|other>."""
        out = np.zeros(1, dtype=np.complex128)
        L.check(L.load().qtn_mps_overlap(self.handle, other.handle,
                                         L.dptr(out)))
        return complex(out[0])

    def norm2(self):
        out = ctypes.c_double(0.0)
        L.check(L.load().qtn_mps_norm2(self.handle, ctypes.byref(out)))
        return out.value


def _set_sites(self, tensors):
    """Replace the state by site tensors (chi_j, 2, chi_{j+1}) of any gauge
    (qtn_mps_set_sites: brought to the right-canonical gauge on the GPU,
    normalised)."""
    if len(tensors) != self.n:
        raise ValueError("need one tensor per site")
    arrs = [L.as_c128(t, "F") for t in tensors]
    dims = np.zeros(3 * self.n, dtype=np.int64)
    for j, t in enumerate(arrs):
        if t.ndim != 3:
            raise ValueError("site tensors must have 3 legs")
        dims[3 * j:3 * j + 3] = t.shape
    ptrs = (ctypes.POINTER(ctypes.c_double) * self.n)(
        *[L.dptr(t) for t in arrs])
    L.check(L.load().qtn_mps_set_sites(self.handle, L.i64ptr(dims), ptrs))


def _expect_mpo(self, mpo, form=FORM_PLAIN):
    """<psi|W|psi> (complex) by the environment sweep; form FORM_CHOSEN runs
    the bond-sized contractions (all dimensions >= 512) as BF16x3
    split-precision GEMMs (qtn_mps_expect_mpo_form)."""
    out = np.zeros(2, dtype=np.float64)
    L.check(L.load().qtn_mps_expect_mpo_form(self.handle, mpo.handle,
                                             int(form), L.dptr(out)))
    return complex(out[0], out[1])


def _apply_mpo(self, mpo, cutoff, chimax, out, form=0):
    """out <- W|self> by zip-up and recompression; returns |W psi| of the
    compressed state (out is normalised).  form FORM_CHOSEN runs the
    cap-decided zip-up splits in FP32 with FP64 re-orthonormalisation
    (qtn_mps_apply_mpo_form)."""
    nrm = ctypes.c_double(0.0)
    lib = L.load()
    if form:
        L.check(lib.qtn_mps_apply_mpo_form(self.handle, mpo.handle,
                                           float(cutoff), int(chimax),
                                           int(form), out.handle,
                                           ctypes.byref(nrm)))
    else:
        L.check(lib.qtn_mps_apply_mpo(self.handle, mpo.handle,
                                      float(cutoff), int(chimax),
                                      out.handle, ctypes.byref(nrm)))
    return nrm.value


DeviceMPS.set_sites = _set_sites
DeviceMPS.expect_mpo = _expect_mpo
DeviceMPS.apply_mpo = _apply_mpo


class DeviceMPO:
    """An MPO on the GPU: site tensors (D_j, 2 out, 2 in, D_{j+1})."""

    def __init__(self, tensors, ctx=None):
        self.ctx = ctx if ctx is not None else L.default_context()
        arrs = [L.as_c128(t, "F") for t in tensors]
        self.n = len(arrs)
        D = np.zeros(self.n + 1, dtype=np.int64)
        for j, t in enumerate(arrs):
            if t.ndim != 4 or t.shape[1:3] != (2, 2):
                raise ValueError("MPO tensors must be (D, 2, 2, D')")
            if j > 0 and t.shape[0] != D[j]:
                raise ValueError("MPO bond dimensions do not match")
            D[j] = t.shape[0]
            D[j + 1] = t.shape[3]
        ptrs = (ctypes.POINTER(ctypes.c_double) * self.n)(
            *[L.dptr(t) for t in arrs])
        h = ctypes.c_void_p()
        L.check(L.load().qtn_mpo_create(self.ctx.handle, self.n,
                                        L.i64ptr(D), ptrs, ctypes.byref(h)))
        self._h = h
        self.ctx.adopt(self)

    @property
    def handle(self):
        if self._h is None:
            raise L.QtnError(L.QTN_ERR_ARG, "DeviceMPO already freed")
        return self._h

    def free(self):
        if self._h is not None:
            L.check(L.load().qtn_mpo_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            if self._h is not None and L._lib is not None:
                L._lib.qtn_mpo_destroy(self._h)
                self._h = None
        except Exception:
            pass


def _state_array(states):
    arr = (ctypes.c_void_p * len(states))()
    for i, st in enumerate(states):
        arr[i] = st.handle.value if isinstance(st.handle, ctypes.c_void_p) \
            else st.handle
    return arr


def apply_1q_multi(states, sites, us):
    """Single-qubit gates us[i] on (states[i], sites[i]) in one launch
    (qtn_mps_apply_1q_multi)."""
    n = len(states)
    if n == 0:
        return
    st = np.ascontiguousarray(sites, dtype=np.int32)
    u = np.ascontiguousarray(
        np.stack([L.as_c128(x, "F").reshape(-1, order="F") for x in us]))
    L.check(L.load().qtn_mps_apply_1q_multi(
        n, _state_array(states), L.i32ptr(st),
        u.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double))))


def apply_layer_multi(states, sites, gates, chimax, cutoff,
                      form=FORM_PLAIN):
    """Two-qubit gates on many states in batched passes grouped by size
    (qtn_mps_apply_layer_multi).  Returns (kept, discarded weight)."""
    n = len(states)
    kept = np.zeros(n, dtype=np.int64)
    disc = np.zeros(n, dtype=np.float64)
    if n == 0:
        return kept, disc
    st = np.ascontiguousarray(sites, dtype=np.int32)
    g = np.ascontiguousarray(
        np.stack([L.as_c128(x, "F").reshape(-1, order="F") for x in gates]))
    cm = np.ascontiguousarray(chimax, dtype=np.int64)
    L.check(L.load().qtn_mps_apply_layer_multi(
        n, _state_array(states), L.i32ptr(st),
        g.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
        L.i64ptr(cm), float(cutoff), int(form), L.i64ptr(kept),
        L.dptr(disc)))
    return kept, disc


def run_circuits(states, circuits, chimax, cutoff, form=FORM_PLAIN,
                 step_times=None):
    """Run many brick-wall circuits (circuits[i] = list of (singles, pairs)
    layers for states[i]) step by step: at step t the singles of layer t of
    every circuit that has one go in one launch, then all its pairs in one
    qtn_mps_apply_layer_multi call.  step_times, when a list, receives the
    wall time of every step (synchronised)."""
    import time
    depth = max(len(c) for c in circuits)
    for t in range(depth):
        t0 = time.perf_counter()
        s1, w1, u1, s2, w2, g2, c2 = [], [], [], [], [], [], []
        for st, circ, cm in zip(states, circuits, chimax):
            if t >= len(circ):
                continue
            singles, pairs = circ[t]
            for (w, u) in singles:
                s1.append(st)
                w1.append(w)
                u1.append(u)
            for (w, gm) in pairs:
                s2.append(st)
                w2.append(w)
                g2.append(gm)
                c2.append(cm)
        apply_1q_multi(s1, w1, u1)
        apply_layer_multi(s2, w2, g2, c2, cutoff, form)
        if step_times is not None:
            step_times.append(time.perf_counter() - t0)


def _apply_layer_learned(self, sites, gates, cutoff, chimax, hints):
    """A layer in the learned form (qtn_mps_apply_layer_learned): returns
    (kept, discarded weight, accepted) per gate; accepted is 1 where the
    sketch decided the gate, -1 where it failed the check and the full SVD
    ran, 0 where no sketch was tried."""
    n = len(sites)
    st = np.ascontiguousarray(sites, dtype=np.int32)
    g = np.ascontiguousarray(
        np.stack([L.as_c128(x, "F").reshape(-1, order="F") for x in gates]))
    h = np.ascontiguousarray(hints, dtype=np.int64)
    kept = np.zeros(n, dtype=np.int64)
    disc = np.zeros(n, dtype=np.float64)
    acc = np.zeros(n, dtype=np.int32)
    L.check(L.load().qtn_mps_apply_layer_learned(
        self.handle, n, L.i32ptr(st),
        g.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
        float(cutoff), int(chimax), L.i64ptr(h), L.i64ptr(kept),
        L.dptr(disc), L.i32ptr(acc)))
    return kept, disc, acc


def _gate_features(self, sites, schmidt, chimax):
    """Predictor features (72 x len(sites) float32) of the gates on
    (sites[i], sites[i] + 1), gathered on the device
    (qtn_mps_gate_features); schmidt is len(sites) x 4."""
    n = len(sites)
    out = np.zeros((72, n), dtype=np.float32, order="F")
    st = np.ascontiguousarray(sites, dtype=np.int32)
    sc = np.ascontiguousarray(schmidt, dtype=np.float64).reshape(n, 4)
    L.check(L.load().qtn_mps_gate_features(
        self.handle, n, L.i32ptr(st), L.dptr(sc), int(chimax),
        out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
    return out


def _apply_layer_predicted(self, sites, gates, schmidt, cutoff, chimax, mlp,
                           margin=0.0):
    """A layer in the learned form with the predictor inside the library
    (qtn_mps_apply_layer_predicted): returns (kept, discarded weight,
    accepted, hints) per gate."""
    n = len(sites)
    st = np.ascontiguousarray(sites, dtype=np.int32)
    g = np.ascontiguousarray(
        np.stack([L.as_c128(x, "F").reshape(-1, order="F") for x in gates]))
    sc = np.ascontiguousarray(schmidt, dtype=np.float64).reshape(n, 4)
    kept = np.zeros(n, dtype=np.int64)
    disc = np.zeros(n, dtype=np.float64)
    acc = np.zeros(n, dtype=np.int32)
    hints = np.zeros(n, dtype=np.int64)
    L.check(L.load().qtn_mps_apply_layer_predicted(
        self.handle, n, L.i32ptr(st),
        g.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
        L.dptr(sc), float(cutoff), int(chimax), mlp.handle, float(margin),
        L.i64ptr(kept), L.dptr(disc), L.i32ptr(acc), L.i64ptr(hints)))
    return kept, disc, acc, hints


DeviceMPS.apply_layer_learned = _apply_layer_learned
DeviceMPS.gate_features = _gate_features
DeviceMPS.apply_layer_predicted = _apply_layer_predicted
