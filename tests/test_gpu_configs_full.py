"""The hot path's workloads at their full sizes, on the GPU.

  40 qubits, Haar brick-wall depth 20, chi 256, cutoff 1e-8 (390 two-qubit
      gates), chosen form (FP32 sweeps, FP64 core): against the FP64 oracle
      run of the same gates -- state infidelity <= 1e-6, identical bond
      dimensions, kept (normalised) singular values within 1e-6.
  256 seeded circuits, n ~ U[16, 48], depth ~ U[8, 24], chi ~ {32, 64, 128},
      batched across circuits: every circuit's caps, normalised sorted
      spectra and chosen-vs-plain infidelity <= 1e-6; the largest circuit of
      each chi class against its own oracle run (1e-10, identical ranks).
  64-site TFIM MPO on a random chi 1024 MPS: <psi|H|psi> from the FP64
      environment sweep within 1e-10 of the oracle's; BF16x3 within 1e-4
      of FP64; the zip-up + recompression to 1024 (FP64 and chosen form)
      with full-size properties (bonds <= 1024, normalised, right-canonical)
      and the chosen zip-up within 1e-6 of the FP64 one.
  64 qubits, Haar depth 24, chi 256, learned form (MLP-sized range finder
      with exact check and full-SVD fallback): within 1e-6 of the plain
      state, with the fallback accounting consistent.
The reference (Qaintensor.jl) functions these replace: svd + truncation
(src/svd.jl:26-33, src/switch.jl:39) and apply_MPO / contract
(src/mpo.jl:184-252)."""
import time

import numpy as np
import pytest

from oracle import qtn_oracle as O
from qaintensor import (DeviceMPO, DeviceMPS, FORM_CHOSEN, FORM_PLAIN,
                        circuits as C, run_circuits)
from qaintensor import hamiltonians as Hm

pytestmark = pytest.mark.gpu


def infidelity(a, b):
    return 1.0 - abs(a.overlap(b)) ** 2 / (a.norm2() * b.norm2())


def _log(msg):
    print(msg, flush=True)


def test_config2_chosen_form_full_circuit(gpu_ctx):
    n, depth, chi, cut = 40, 20, 256, 1e-8
    layers = C.brickwall_layers(n, depth, 2)
    t0 = time.perf_counter()
    st = O.TebdState(n)
    O.run_layers(st, layers, cut, chi)
    _log("config 2 oracle %.1f s" % (time.perf_counter() - t0))
    dev = DeviceMPS(n, chi, gpu_ctx)
    dev.run_layers(layers, cut, chi, form=FORM_CHOSEN, batched=True)
    assert list(dev.bond_dims()) == st.bond_dims()
    ref = DeviceMPS(n, chi, gpu_ctx)
    ref.set_sites(st.B)  # the oracle's state, canonical on the device
    inf = infidelity(dev, ref)
    err_abs, err_rel = 0.0, 0.0
    for b in range(1, n):
        s, r = dev.bond_spectrum(b), st.lam[b]
        assert s.size == r.size
        err_abs = max(err_abs, float(np.abs(s - r).max()))
        err_rel = max(err_rel, float((np.abs(s - r) / r).max()))
    _log("config 2 chosen: infidelity %.2e, kept singular values max abs "
         "error %.2e, max relative error %.2e" % (inf, err_abs, err_rel))
    assert inf <= 1e-6
    assert err_abs <= 1e-6


def test_config3_all_256_circuits(gpu_ctx):
    cutoff = 1e-8
    circs = C.batch_circuits(256, 3)
    layers = [c[3] for c in circs]
    chis = [c[2] for c in circs]
    t0 = time.perf_counter()
    a = [DeviceMPS(n, chi, gpu_ctx) for (n, _, chi, _) in circs]
    run_circuits(a, layers, chis, cutoff)
    b = [DeviceMPS(n, chi, gpu_ctx) for (n, _, chi, _) in circs]
    run_circuits(b, layers, chis, cutoff, form=FORM_CHOSEN)
    _log("config 3 both forms %.1f s" % (time.perf_counter() - t0))
    worst = 0.0
    for x, y, (n, depth, chi, _) in zip(a, b, circs):
        dx = list(x.bond_dims())
        assert max(dx) <= chi and dx[0] == dx[-1] == 1
        assert list(y.bond_dims()) == dx, (n, depth, chi)
        for bond in range(1, n, 5):
            lam = x.bond_spectrum(bond)
            assert abs(np.sum(lam ** 2) - 1.0) <= 1e-12
            assert np.all(np.diff(lam) <= 0)
        worst = max(worst, infidelity(x, y))
    _log("config 3 chosen vs plain: max infidelity %.2e" % worst)
    assert worst <= 1e-6
    # the largest circuit of each chi class against its own oracle run
    for chi in (32, 64, 128):
        i = max((c[0], c[1], j) for j, c in enumerate(circs)
                if c[2] == chi)[2]
        n, depth, _, lay = circs[i]
        st = O.TebdState(n)
        O.run_layers(st, lay, cutoff, chi)
        assert list(a[i].bond_dims()) == st.bond_dims(), (n, depth, chi)
        bits = C.random_bitstrings(n, 64, i)
        amp, ref = a[i].amplitudes(bits), st.amplitudes(bits)
        assert np.linalg.norm(amp - ref) <= 1e-10 * np.linalg.norm(ref)
        assert np.abs(a[i].expect_z() - st.expect_z()).max() <= 1e-10
        _log("config 3 circuit %d (n %d, depth %d, chi %d) matches the "
             "oracle" % (i, n, depth, chi))


def _right_canonical_error(dev, j):
    B = dev.site(j)
    m = B.reshape((B.shape[0], -1), order="F")
    return float(np.abs(m @ m.conj().T - np.eye(B.shape[0])).max())


def test_config4_tfim64_chi1024(gpu_ctx):
    n, chi, seed, cutoff = 64, 1024, 4, 1e-10
    W = Hm.tfim_mpo(n, 1.0, Hm.tfim_fields(n, seed))
    raw = Hm.random_mps_tensors(n, chi, seed)
    psi = DeviceMPS(n, chi, gpu_ctx)
    psi.set_sites(raw)
    mpo = DeviceMPO(W, gpu_ctx)
    e64 = psi.expect_mpo(mpo)
    e16 = psi.expect_mpo(mpo, FORM_CHOSEN)
    t0 = time.perf_counter()
    e_ref = O.expect_mpo_blas(O.left_canonical(raw), W)
    _log("config 4 oracle sweep %.1f s" % (time.perf_counter() - t0))
    rel64 = abs(e64 - e_ref) / abs(e_ref)
    rel16 = abs(e16 - e64) / abs(e64)
    _log("config 4 <psi|H|psi>: FP64 %.2e vs oracle, BF16x3 %.2e vs FP64"
         % (rel64, rel16))
    assert rel64 <= 1e-10
    assert rel16 <= 1e-4
    # zip-up and recompression at full size
    t0 = time.perf_counter()
    out = DeviceMPS(n, chi, gpu_ctx)
    nrm = psi.apply_mpo(mpo, cutoff, chi, out)
    t64 = time.perf_counter() - t0
    t0 = time.perf_counter()
    outc = DeviceMPS(n, chi, gpu_ctx)
    nrmc = psi.apply_mpo(mpo, cutoff, chi, outc, form=FORM_CHOSEN)
    tc = time.perf_counter() - t0
    inf = 1.0 - abs(outc.overlap(out))
    _log("config 4 zip-up FP64 %.1f s, chosen %.1f s, infidelity %.2e, "
         "norm %.2e" % (t64, tc, inf, abs(nrmc - nrm) / nrm))
    for o in (out, outc):
        d = list(o.bond_dims())
        assert max(d) == chi and d[0] == d[-1] == 1
        assert abs(o.norm2() - 1.0) <= 1e-10
        for j in (2, 17, 32, 48, 63):
            assert _right_canonical_error(o, j) <= 1e-10
    assert list(outc.bond_dims()) == list(out.bond_dims())
    # chosen form: BF16x3 contractions (FP32-level products) at each of
    # the 64 sites: the normalised state within 1e-6 infidelity, its norm
    # within 1e-5 relative
    assert inf <= 1e-6
    assert abs(nrmc - nrm) <= 1e-5 * nrm


def test_config5_learned_64q(gpu_ctx):
    from qaintensor import learned as Lrn
    cutoff, chi = 1e-8, 256
    x, t = Lrn.harvest(C.brickwall_layers(40, 20, 101), 40, chi, cutoff,
                       ctx=gpu_ctx)
    mlp = Lrn.RankMLP(Lrn.init_params(1), ctx=gpu_ctx)
    try:
        perm = np.random.default_rng(0).permutation(t.size)
        mlp.train(x[:, perm], t[perm], steps=600, minibatch=256, lr=2e-3)
        layers = C.brickwall_layers(64, 24, 7)
        n2q = C.count_2q(layers)
        plain = DeviceMPS(64, chi, gpu_ctx)
        plain.run_layers(layers, cutoff, chi, FORM_PLAIN, batched=True)
        for name, net, rule in (("learned", mlp, None),
                                ("previous_rank", None, "previous")):
            dev = DeviceMPS(64, chi, gpu_ctx)
            stats = {}
            Lrn.run_layers_learned(dev, layers, cutoff, chi, net,
                                   margin=0.02, stats=stats, hint_rule=rule)
            inf = infidelity(plain, dev)
            tried, acc = stats["tried"], stats["accepted"]
            _log("config 5 %s: %d sketches tried, %d accepted, %d fell "
                 "back, infidelity %.2e" % (name, tried, acc, tried - acc,
                                            inf))
            assert stats["gates"] == n2q
            assert 0 <= acc <= tried <= n2q
            assert inf <= 1e-6
            assert list(dev.bond_dims()) == list(plain.bond_dims())
            if rule == "previous":
                # the bond's previous rank is a futile sketch on a Haar
                # brick-wall once the bonds saturate: fallbacks happen and
                # still give the plain state
                assert tried > acc
    finally:
        mlp.free()
