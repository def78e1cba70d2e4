"""The mixed-precision mode (SYMGW_PREC_MIXED: Sigma's q x q x k1 k2 product
on INT8 slices with exact int32 accumulation, everything else FP64) against
the FP64 oracle, through the C ABI.  Tolerance: max|dSigma| <= 1e-8 Ha and
max|dG| <= 1e-8 after a full iteration, absolute, as the mixed-mode target
states.  The INT8 path is forced onto every aux block (SYMGW_Z8_MIN_Q=0) so
that the small systems exercise it; the default threshold is checked on the
large-block system."""
import os

import numpy as np
import pytest

from mbsymmetry_amd import _lib
from mbsymmetry_amd import system as S
from mbsymmetry_amd.transforms import chebyshev
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def systems(lih, config1):
    out = {
        "tiny_rot_ns2": S.tiny(ns=2, n=2, ink=4, ao=(3, 2), aux=(4, 3), ni=8, beta=10.0, seed=11),
        "tr_square": S.time_reversal_square(ns=2, ao=(5, 3, 2), aux=(12, 8, 6), ni=10, beta=12.0, seed=13),
        "config1": config1,
        "lih": lih,
        "big": S.synthetic("big_blocks", 2, 4, [66, 10], [100, 24], 1, chebyshev(8, 6.0), 31),
    }
    # synthetic systems at the benchmark systems' Coulomb scale (1/sqrt(naux))
    for name in ("tiny_rot_ns2", "tr_square", "big"):
        out[name].v_scale = S.coulomb_scale(out[name].aux_sizes[0])
    return out


def _ctx(sysm, precision, min_q):
    old = os.environ.get("SYMGW_Z8_MIN_Q")
    try:
        if min_q is None:
            os.environ.pop("SYMGW_Z8_MIN_Q", None)
        else:
            os.environ["SYMGW_Z8_MIN_Q"] = str(min_q)
        ctx = _lib.Context(sysm, 0, precision)
    finally:
        if old is None:
            os.environ.pop("SYMGW_Z8_MIN_Q", None)
        else:
            os.environ["SYMGW_Z8_MIN_Q"] = old
    if sysm.V is not None:
        ctx.set_V(sysm.V)
    else:
        ctx.generate_V(sysm.v_seed, sysm.v_scale, sysm.decay_table())
    ctx.set_reps(sysm.orep, sysm.auxrep)
    ctx.set_transforms(sysm.transforms)
    ctx.set_fock(sysm.F, sysm.S)
    return ctx


@pytest.mark.parametrize("name,min_q", [("tiny_rot_ns2", 0), ("tr_square", 0), ("config1", 0), ("lih", 0),
                                        ("big", None), ("big", 0)])
def test_mixed_sigma_and_iteration_within_1e8(name, min_q, systems):
    sysm = systems[name]
    _say("mixed", name, "min_q", min_q)
    orc = Oracle(sysm)
    G0 = orc.dyson(sysm.zeros_ts_kij(), 0.0)
    V = sysm.host_V()
    Gf = orc.rotate_G(G0)
    P0 = orc.P0(V, Gf)
    P = orc.W(P0)
    Sref = orc.Sigma(V, Gf, orc.rotate_P(P))
    G1ref = orc.dyson(Sref, 0.0)
    ctx = _ctx(sysm, _lib.PREC_MIXED, min_q)
    # Sigma alone from the oracle's P
    ctx.upload_G(G0)
    ctx.upload_P(P)
    ctx.solve_Sigma()
    Sg = ctx.copy_Sigma_to_host(sysm.zeros_ts_kij()) / sysm.nk
    err_s = float(np.abs(Sg - Sref).max())
    # a whole iteration (P0 and W FP64, Sigma mixed, Dyson)
    ctx.upload_G(G0)
    ctx.gw_iteration(0.0, use_graph=False)
    ctx.sync()
    G1 = ctx.download_G(G0.shape)
    err_g = float(np.abs(G1 - G1ref).max())
    _say(f"  max|dSigma| {err_s:.2e} (max|Sigma| {np.abs(Sref).max():.3e}), max|dG| {err_g:.2e}")
    assert err_s <= 1e-8, name
    assert err_g <= 1e-8, name
    # normwise accuracy of 5 slices per operand, independent of the V scale
    assert err_s <= 1e-9 * np.abs(Sref).max(), name
    # the INT8 path must actually have run: not bitwise equal to FP64 Sigma
    if min_q == 0:
        assert err_s > 0.0
    ctx.close()


def test_mixed_is_deterministic(systems):
    """Same bits on repeated Sigma sweeps (exact integer sums, fixed-order
    reconstruction)."""
    sysm = systems["config1"]
    ctx = _ctx(sysm, _lib.PREC_MIXED, 0)
    orc = Oracle(sysm)
    G0 = orc.dyson(sysm.zeros_ts_kij(), 0.0)
    ctx.upload_G(G0)
    ctx.solve_P0()
    ctx.solve_W()
    outs = []
    for _ in range(2):
        ctx.solve_Sigma()
        outs.append(ctx.copy_Sigma_to_host(sysm.zeros_ts_kij()))
    assert np.array_equal(outs[0], outs[1])
    ctx.close()


def test_mixed_footprint_includes_slices(systems):
    sysm = systems["big"]
    fp64 = _lib.layout_footprint(sysm, _lib.PREC_FP64)
    mixed = _lib.layout_footprint(sysm, _lib.PREC_MIXED)
    assert mixed > fp64
