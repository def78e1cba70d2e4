"""GPU checks of the uneven-block system (config 3: 64 irreducible points with
their own 2..10 irreps of sizes 1..256) and of the machinery config 4 needs
(V streamed per batch of k-pairs, shards of iBZ points), through the C ABI.

Config 4 itself (8x8x8 mesh, V of 5.5 TB) is only timed on a sample; its
block structure (nao 256 in 6 blocks, naux 1200 in 6 blocks) is checked here
on a 2x2x2 mesh, where V also fits resident so the streamed sweeps can be
compared with the resident ones bit for bit, and against the oracle."""
import os

import numpy as np
import pytest

from mbsymmetry_amd import _lib
from mbsymmetry_amd import system as S
from mbsymmetry_amd.transforms import chebyshev
from oracle.oracle import Oracle
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _ctx(sysm, precision, batch_bytes=None):
    """Context with V generated; batch_bytes overrides the streamed-batch size
    (1: one iBZ point per batch)."""
    old = os.environ.get("SYMGW_STREAM_BATCH_BYTES")
    try:
        if batch_bytes is not None:
            os.environ["SYMGW_STREAM_BATCH_BYTES"] = str(batch_bytes)
        ctx = _lib.Context(sysm, 0, precision)
    finally:
        if old is None:
            os.environ.pop("SYMGW_STREAM_BATCH_BYTES", None)
        else:
            os.environ["SYMGW_STREAM_BATCH_BYTES"] = old
    ctx.generate_V(sysm.v_seed, sysm.v_scale, sysm.decay_table())
    ctx.set_reps(sysm.orep, sysm.auxrep)
    ctx.set_transforms(sysm.transforms)
    ctx.set_fock(sysm.F, sysm.S)
    return ctx


def _sweeps(ctx, sysm, G0):
    """P0, P, Sigma (both as the device holds them) and the new G."""
    ctx.upload_G(G0)
    ctx.solve_P0()
    P0 = ctx.copy_P0_to_host(sysm.zeros_q_t_QP())
    ctx.solve_W()
    P = ctx.copy_P0_to_host(sysm.zeros_q_t_QP())
    ctx.solve_Sigma()
    Sg = ctx.copy_Sigma_to_host(sysm.zeros_ts_kij())
    ctx.dyson(0.0, 1.0 / sysm.nk)
    G1 = ctx.download_G(G0.shape)
    return P0, P, Sg, G1


@pytest.fixture(scope="module")
def c4small():
    ao, aux = S.config4_sizes()
    return S.synthetic("config4-2x2x2", 2, 4, ao, aux, 1, chebyshev(6, 200.0), S.SEED_BASE + 4,
                       v_scale=S.coulomb_scale(aux))


def test_streamed_V_equals_resident_V(c4small):
    sysm = c4small
    orc = Oracle(sysm)
    G0 = orc.dyson(sysm.zeros_ts_kij(), 0.0)
    for prec in (_lib.PREC_FP64, _lib.PREC_FP32, _lib.PREC_MIXED):
        _say("stream vs resident, config-4 blocks on 2x2x2, precision", prec)
        a = _ctx(sysm, prec)
        ra = _sweeps(a, sysm, G0)
        a.close()
        # all iBZ points in one batch (the default size): the same launch
        # groups as the resident sweep, so the same bits; one point per
        # batch: the planner splits items per launch group, so items can be
        # cut into other parts and the sums differ in rounding only
        for bb in (None, 1):
            b = _ctx(sysm, prec | _lib.FLAG_STREAM_V, bb)
            assert b.info().device_bytes < _lib.layout_footprint(sysm, prec)
            rb = _sweeps(b, sysm, G0)
            b.close()
            for x, y, name in zip(ra, rb, ("P0", "P", "Sigma", "G")):
                if bb is None:
                    assert np.array_equal(x, y), (prec, bb, name)
                else:
                    tol = 1e-12 if prec != _lib.PREC_FP32 else 1e-5
                    assert np.abs(x - y).max() <= tol * np.abs(x).max(), (prec, bb, name)


def test_config4_blocks_against_the_oracle(c4small):
    """Sampled slices of the streamed FP64 sweeps against the oracle."""
    sysm = c4small
    nk, nts = sysm.nk, sysm.nts
    orc = Oracle(sysm)
    G0 = orc.dyson(sysm.zeros_ts_kij(), 0.0)
    ctx = _ctx(sysm, _lib.PREC_FP64 | _lib.FLAG_STREAM_V)
    P0, P, Sg, _ = _sweeps(ctx, sysm, G0)
    ctx.close()
    P0 /= nk
    Sg /= nk
    Gf = orc.rotate_G(G0)
    iqr, t = 1, 2
    q = int(sysm.irre_list[iqr])
    Vs, off = sysm.host_V_pairs([k * nk + int(sysm.kpq[k, q]) for k in range(nk)])
    ref = orc.P0_at(Vs, off, Gf, iqr, t, nthreads=16)
    qs, qo = sysm.qslice(q), sysm.qaux_off_irre()[iqr]
    err = float(np.abs(P0[nts * qo + t * qs: nts * qo + (t + 1) * qs] - ref).max())
    _say(f"config-4 blocks P0 slice: {err:.2e} (max {np.abs(ref).max():.2e})")
    assert err < 1e-10
    ikr, t = 2, 3
    k = int(sysm.irre_list[ikr])
    Vs, off = sysm.host_V_pairs([k * nk + int(sysm.kmq[k, qq]) for qq in range(nk)])
    ref = orc.Sigma_at(Vs, off, Gf, orc.rotate_P_t(P, t), ikr, t, 0, nthreads=16)
    ko, ks = sysm.kao_off_irre()[ikr], sysm.kslice(k)
    err = float(np.abs(Sg[t, 0, ko:ko + ks] - ref).max())
    _say(f"config-4 blocks Sigma slice: {err:.2e} (max {np.abs(ref).max():.2e})")
    assert err < 1e-10


@pytest.mark.parametrize("flags,batch_bytes", [(0, None), (_lib.FLAG_STREAM_V, None), (_lib.FLAG_STREAM_V, 1),
                                               (_lib.FLAG_STREAM_V, 5 << 20)])
def test_shard_recomputes_its_rows(config1, flags, batch_bytes):
    """A shard's P0 and Sigma rows equal those of the full sweep bit for bit;
    the rows outside it are left untouched.  Resident V, and V streamed in
    batches of all four, one, or two iBZ points (a shard cutting a batch)."""
    sysm = config1
    orc = Oracle(sysm)
    G0 = orc.dyson(sysm.zeros_ts_kij(), 0.0)
    ctx = _ctx(sysm, _lib.PREC_FP64 | flags, batch_bytes)
    P0, P, Sg, _ = _sweeps(ctx, sysm, G0)
    ctx.set_shard(1, 2)
    ctx.upload_G(G0)
    ctx.upload_P(np.zeros_like(P0))
    ctx.solve_P0()
    P0s = ctx.copy_P0_to_host(sysm.zeros_q_t_QP())
    ctx.upload_P(P)
    ctx.solve_Sigma()
    Sgs = ctx.copy_Sigma_to_host(sysm.zeros_ts_kij())
    ctx.close()
    qo = sysm.qaux_off_irre() + [sysm.flat_aux]
    ko = sysm.kao_off_irre() + [sysm.flat_ao]
    nts = sysm.nts
    for i in range(sysm.ink):
        a, b = nts * qo[i], nts * qo[i + 1]
        if 1 <= i < 3:
            assert np.array_equal(P0s[a:b], P0[a:b]), i
            assert np.array_equal(Sgs[:, :, ko[i]:ko[i + 1]], Sg[:, :, ko[i]:ko[i + 1]]), i
        else:
            assert np.all(P0s[a:b] == 0), i


@pytest.mark.timeout(900)
def test_config3_uneven_blocks():
    """Config 3 in one launch per contraction: FP64 against sampled oracle
    slices (1e-10), complex64 storage against complex128 (relative 1e-3 of
    the largest element), the mixed mode against FP64 (1e-8 Ha)."""
    sysm = S.config3()
    nk, nts = sysm.nk, sysm.nts
    orc = Oracle(sysm)
    G0 = orc.dyson(sysm.zeros_ts_kij(), 0.0)
    res = {}
    for name, prec in (("fp64", _lib.PREC_FP64), ("fp32", _lib.PREC_FP32), ("mixed", _lib.PREC_MIXED)):
        _say("config3", name)
        ctx = _ctx(sysm, prec)
        ctx.upload_G(G0)
        ctx.solve_P0()
        P0 = ctx.copy_P0_to_host(sysm.zeros_q_t_QP()) / nk
        ctx.solve_W()
        P = ctx.copy_P0_to_host(sysm.zeros_q_t_QP())
        ctx.solve_Sigma()
        Sg = ctx.copy_Sigma_to_host(sysm.zeros_ts_kij()) / nk
        ctx.close()
        res[name] = (P0, P, Sg)
    P0, P, Sg = res["fp64"]
    assert H.p0_hermiticity_error(sysm, P0) < 1e-11
    assert H.sigma_hermiticity_error(sysm, Sg) < 1e-11
    Gf = orc.rotate_G(G0)
    qo, ko = sysm.qaux_off_irre(), sysm.kao_off_irre()
    big_q = int(np.argmax([max(a) for a in sysm.aux_sizes]))
    for iqr, t in ((big_q, 1), (5, 3)):
        q = int(sysm.irre_list[iqr])
        Vs, off = sysm.host_V_pairs([k * nk + int(sysm.kpq[k, q]) for k in range(nk)])
        ref = orc.P0_at(Vs, off, Gf, iqr, t, nthreads=16)
        qs = sysm.qslice(q)
        err = float(np.abs(P0[nts * qo[iqr] + t * qs: nts * qo[iqr] + (t + 1) * qs] - ref).max())
        _say(f"config3 P0 slice q{iqr} t{t}: {err:.2e} (max {np.abs(ref).max():.2e})")
        assert err < 1e-10
    big_k = int(np.argmax([max(a) for a in sysm.ao_sizes]))
    for ikr, t in ((big_k, 2), (9, 6)):
        k = int(sysm.irre_list[ikr])
        Vs, off = sysm.host_V_pairs([k * nk + int(sysm.kmq[k, qq]) for qq in range(nk)])
        ref = orc.Sigma_at(Vs, off, Gf, orc.rotate_P_t(P, t), ikr, t, 0, nthreads=16)
        ks = sysm.kslice(k)
        err = float(np.abs(Sg[t, 0, ko[ikr]:ko[ikr] + ks] - ref).max())
        _say(f"config3 Sigma slice k{ikr} t{t}: {err:.2e} (max {np.abs(ref).max():.2e})")
        assert err < 1e-10
    for name, tol in (("fp32", 1e-3), ("mixed", None)):
        P0x, _, Sgx = res[name]
        dp = float(np.abs(P0x - P0).max())
        ds = float(np.abs(Sgx - Sg).max())
        _say(f"config3 {name} vs fp64: max|dP0| {dp:.2e}, max|dSigma| {ds:.2e} (max|Sigma| {np.abs(Sg).max():.2e})")
        if tol is None:
            assert dp == 0.0       # P0 runs in FP64 in the mixed mode
            assert ds <= 1e-8
        else:
            assert dp <= tol * np.abs(P0).max()
            assert ds <= tol * np.abs(Sg).max()
