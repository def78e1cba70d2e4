"""The BLAS timing twin of the oracle's sampled slices (oracle/blas.py)
against the C oracle's orc_P0_at / orc_Sigma_at, on systems with rotations,
two spins and time-reversal pairs, serially and split over threads."""
import numpy as np
import pytest

from mbsymmetry_amd import system as S
from oracle import blas as B
from oracle.oracle import Oracle


@pytest.mark.parametrize("name", ["config1", "tiny_rot_ns2", "tr_square"])
def test_blas_slices_equal_the_oracle(name, config1):
    sysm = {"config1": config1,
            "tiny_rot_ns2": S.tiny(ns=2, n=2, ink=4, ao=(3, 2), aux=(4, 3), ni=8, beta=10.0, seed=11),
            "tr_square": S.time_reversal_square(ns=2, ao=(5, 3, 2), aux=(12, 8, 6), ni=10, beta=12.0, seed=13)}[name]
    orc = Oracle(sysm)
    G0 = orc.dyson(sysm.zeros_ts_kij(), 0.0)
    Gf = orc.rotate_G(G0)
    V = sysm.host_V()
    off = {p: o for p, (o, _) in sysm.kpair_offsets().items()}
    P = orc.W(orc.P0(V, Gf))
    T = B.Tables(sysm)
    nk = sysm.nk
    for iqr in range(sysm.ink):
        for t in (0, sysm.nts // 2 - 1):
            ref = orc.P0_at(V, off, Gf, iqr, t)
            for w in (1, 3):
                got = B.p0_at(T, V, off, Gf, iqr, t, workers=w)
                assert np.abs(got - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max()), (iqr, t, w)
    for ikr in range(sysm.ink):
        for t in (0, sysm.nts - 1):
            Pt = orc.rotate_P_t(P, t)
            for sp in range(sysm.ns):
                ref = orc.Sigma_at(V, off, Gf, Pt, ikr, t, sp)
                got = B.sigma_at(T, V, off, Gf, Pt, ikr, t, sp, workers=2)
                assert np.abs(got - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max()), (ikr, t, sp)
    # a partial k sum is the matching part of the full one
    q_half = list(range(0, nk, 2))
    a = B.p0_at(T, V, off, Gf, 0, 1, ks=q_half)
    b = B.p0_at(T, V, off, Gf, 0, 1, ks=[k for k in range(nk) if k % 2])
    assert np.allclose(a + b, B.p0_at(T, V, off, Gf, 0, 1), rtol=0, atol=1e-13)


def test_flop_counts_match_the_device_model(config1):
    """Summed over every slice, the slice FLOP counts are the layout's P0 and
    Sigma FLOP totals."""
    from mbsymmetry_amd import _lib
    sysm = config1
    T = B.Tables(sysm)
    q = _lib.layout_query(sysm)
    p0 = sum(B.flops_p0_at(T, i) for i in range(sysm.ink)) * (sysm.nts // 2)
    sg = sum(B.flops_sigma_at(T, i) for i in range(sysm.ink)) * sysm.nts * sysm.ns
    print(p0, q.flop_P0, sg, q.flop_Sigma)
    assert p0 == pytest.approx(q.flop_P0, rel=1e-12)
    assert sg == pytest.approx(q.flop_Sigma, rel=1e-12)
