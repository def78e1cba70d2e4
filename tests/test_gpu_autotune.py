"""The work-list plan: every plan the autotuner may pick gives the oracle's
P0 and Sigma (to rounding: the plan orders the sums), replan refuses rules
out of range, and a tuned context iterates like the oracle."""
import numpy as np
import pytest

from mbsymmetry_amd import _lib, autotune
from mbsymmetry_amd import system as S
from mbsymmetry_amd.transforms import chebyshev
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def _ctx(sysm, precision=_lib.PREC_FP64):
    ctx = _lib.Context(sysm, 0, precision)
    ctx.generate_V(sysm.v_seed, sysm.v_scale, sysm.decay_table())
    ctx.set_reps(sysm.orep, sysm.auxrep)
    ctx.set_transforms(sysm.transforms)
    ctx.set_fock(sysm.F, sysm.S)
    return ctx


@pytest.mark.parametrize("name", ["config1", "permuted", "big"])
def test_every_candidate_plan_matches_the_oracle(name, config1):
    sysm = {"config1": config1, "permuted": S.permuted_blocks(n=3, ns=2),
            "big": S.synthetic("big_blocks", 2, 4, [66, 10], [100, 24], 1, chebyshev(8, 6.0), 31)}[name]
    orc = Oracle(sysm)
    G0 = orc.dyson(sysm.zeros_ts_kij(), 0.0)
    V = sysm.host_V()
    Gf = orc.rotate_G(G0)
    P0_ref = orc.P0(V, Gf)
    P = orc.W(P0_ref)
    S_ref = orc.Sigma(V, Gf, orc.rotate_P(P))
    ctx = _ctx(sysm)
    ctx.upload_G(G0)
    plans = [{}] + autotune.SIG_CANDIDATES + autotune.P0_CANDIDATES
    for plan in plans:
        print("[case]", name, plan, flush=True)
        ctx.replan(**plan)
        assert ctx.plan() == dict(_lib.plan_defaults().as_dict(), **plan)
        ctx.solve_P0()
        P0 = ctx.copy_P0_to_host(sysm.zeros_q_t_QP()) / sysm.nk
        ctx.upload_P(P)
        ctx.solve_Sigma()
        Sg = ctx.copy_Sigma_to_host(sysm.zeros_ts_kij()) / sysm.nk
        assert np.abs(P0 - P0_ref).max() <= 1e-12 * np.abs(P0_ref).max(), plan
        assert np.abs(Sg - S_ref).max() <= 1e-12 * np.abs(S_ref).max(), plan
    ctx.close()


def test_replan_refuses_rules_out_of_range(config1):
    ctx = _ctx(config1)
    for bad in (dict(sig_rows=0), dict(sig_rows=65), dict(sig_max_times=3), dict(p0_band_cap=17),
                dict(sig_chunk_cost=-1.0)):
        with pytest.raises(_lib.SymGWError):
            ctx.replan(**bad)
    ctx.close()


def test_tuned_context_iterates_like_the_oracle(config1, tmp_path, monkeypatch):
    monkeypatch.setenv("SYMGW_TUNE_CACHE", str(tmp_path / "plans.json"))
    sysm = config1
    orc = Oracle(sysm)
    G0 = orc.dyson(sysm.zeros_ts_kij(), 0.0)
    for precision in (_lib.PREC_FP64, _lib.PREC_MIXED):
        ctx = _ctx(sysm, precision)
        ctx.upload_G(G0)
        rec = autotune.tune(ctx, sysm, precision, sample_s=0.05)
        assert not rec["cached"] and set(rec["plan"]) == set(_lib.plan_defaults().as_dict())
        assert rec["Sigma"]["best_ms"] <= rec["Sigma"]["default_ms"]
        ctx.upload_G(G0)
        ctx.gw_iteration(0.0, use_graph=True)
        ctx.sync()
        G1 = ctx.download_G(G0.shape)
        _, _, _, ref = orc.gw_iteration(G0, 0.0)
        assert np.abs(G1 - ref).max() < (1e-10 if precision == _lib.PREC_FP64 else 1e-8)
        ctx.close()
        # the same block structure finds the cached plan
        ctx = _ctx(sysm, precision)
        rec2 = autotune.tune(ctx, sysm, precision)
        assert rec2["cached"] and ctx.plan() == rec["plan"]
        ctx.close()
