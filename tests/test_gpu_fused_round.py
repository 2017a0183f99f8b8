"""The FP32 fused Jacobi round with the block pair resident in LDS (the
chosen form's default for 320-512 rows) against the three-launch FP32
round (QTN_JACOBI_FUSED=0) and the FP64 engine, on two-site updates of a
random chi-256 state: a square 512 x 512 split (theta^H swept, no
accumulated rotations) and a 512 x 320 split (theta swept with its 320 x
320 rotation accumulator, the W tile of the fused round).  Both splits are
capped at 256, so the FP32 result is the one kept.  The reference for the
state is the plain FP64 update (src/switch.jl:18-56 through the device
engine, itself checked against the oracle elsewhere)."""
import numpy as np
import pytest

from qaintensor import FORM_CHOSEN, FORM_PLAIN, DeviceMPS
from qaintensor import circuits as C
from qaintensor import hamiltonians as Hm

pytestmark = pytest.mark.gpu

N, CHI = 20, 256


def infidelity(a, b):
    return 1.0 - abs(a.overlap(b)) ** 2 / (a.norm2() * b.norm2())


def prepared(ctx, raw, tall):
    psi = DeviceMPS(N, CHI, ctx)
    psi.set_sites(raw)
    if tall:
        # bond 11 cut to 160, so the gate on sites 10, 11 splits 512 x 320
        g = C.haar_unitary(np.random.default_rng(11), 4)
        psi.apply_layer([11], [g], 0.0, 160, form=FORM_PLAIN)
    return psi


@pytest.mark.parametrize("tall", [False, True])
def test_fp32_fused_round_matches_three_launch_and_fp64(gpu_ctx, monkeypatch,
                                                        tall):
    raw = Hm.random_mps_tensors(N, CHI, 21)
    g = C.haar_unitary(np.random.default_rng(7), 4)
    dims = prepared(gpu_ctx, raw, tall).bond_dims()
    assert (2 * dims[9], 2 * dims[11]) == ((512, 320) if tall else (512, 512))

    ref = prepared(gpu_ctx, raw, tall)
    kept_ref, _ = ref.apply_layer([10], [g], 1e-8, CHI, form=FORM_PLAIN)

    out = {}
    for mode in ("default", "0"):
        if mode == "default":
            monkeypatch.delenv("QTN_JACOBI_FUSED", raising=False)
        else:
            monkeypatch.setenv("QTN_JACOBI_FUSED", mode)
        psi = prepared(gpu_ctx, raw, tall)
        gpu_ctx.profile(True)
        kept, _ = psi.apply_layer([10], [g], 1e-8, CHI, form=FORM_CHOSEN)
        prof = gpu_ctx.profile_read()
        gpu_ctx.profile(False)
        out[mode] = (psi, kept, prof)
    monkeypatch.delenv("QTN_JACOBI_FUSED", raising=False)

    fused, kept_f, prof_f = out["default"]
    three, kept_t, prof_t = out["0"]
    # the default ran the fused round, the forced setting the three launches
    assert prof_f["jac_round_fp32"]["count"] > 0
    assert prof_f["jac_inner_fp32"]["count"] == 0
    assert prof_t["jac_inner_fp32"]["count"] > 0
    assert prof_t["jac_round_fp32"]["count"] == 0
    assert int(kept_ref[0]) == int(kept_f[0]) == int(kept_t[0]) == CHI
    assert list(fused.bond_dims()) == list(ref.bond_dims())
    assert infidelity(fused, ref) <= 1e-6
    assert infidelity(three, ref) <= 1e-6
    assert infidelity(fused, three) <= 1e-6
