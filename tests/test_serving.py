"""Serving loop: admission policy on the CPU (stand-in runner), and on the
GPU the same bits for a system alone and in a concurrent mix."""
import threading
import time

import numpy as np
import pytest

from mbsymmetry_amd import serving as SV


def _fake(requests, runtime):
    lock = threading.Lock()
    running = {"n": 0, "max": 0}

    def runner(r):
        with lock:
            running["n"] += 1
            running["max"] = max(running["max"], running["n"])
        time.sleep(runtime[r.rid])
        with lock:
            running["n"] -= 1
        return None
    return runner, running


def test_scheduler_budget_fifo_and_rejection():
    rng = np.random.default_rng(0)
    n = 24
    sizes = rng.integers(1, 6, n) * 100
    sizes[5] = 10_000                       # larger than the whole budget: rejected
    arrivals = SV.poisson_arrivals(n, 200.0, seed=1)
    reqs = [SV.Request(i, None, float(arrivals[i]), iterations=1) for i in range(n)]
    runtime = {i: 0.01 + 0.002 * int(sizes[i] // 100) for i in range(n)}
    runner, running = _fake(reqs, runtime)
    sch = SV.Scheduler(runner, lambda r: int(sizes[r.rid]), hbm_budget=1000, max_active=3)
    res = sch.run(reqs)
    assert len(res) == n
    assert res[5].error is not None and "exceeds" in res[5].error
    ok = [r for r in res if r.error is None]
    assert len(ok) == n - 1
    assert sch.peak_used <= 1000 and sch.peak_active <= 3 and running["max"] <= 3
    # FIFO: admission times follow arrival order
    order = sorted(ok, key=lambda r: (r.arrival, r.rid))
    adm = [r.admitted for r in order]
    assert all(a <= b + 1e-9 for a, b in zip(adm, adm[1:]))
    for r in ok:
        assert r.admitted >= r.arrival - 1e-6
        assert r.latency >= runtime[r.rid] * 0.9
    s = SV.summarize(res, 1.0)
    assert s["completed"] == n - 1 and s["failed"] == 1 and s["p99_latency_s"] >= s["p50_latency_s"]


def test_scheduler_runner_error_is_reported():
    reqs = [SV.Request(i, None, 0.0) for i in range(3)]

    def runner(r):
        if r.rid == 1:
            raise RuntimeError("boom")
        return None
    res = SV.Scheduler(runner, lambda r: 1, hbm_budget=10, max_active=2).run(reqs)
    assert res[1].error and "boom" in res[1].error
    assert res[0].error is None and res[2].error is None


def test_poisson_arrivals_rate():
    a = SV.poisson_arrivals(20000, 50.0, seed=3)
    assert np.all(np.diff(a) > 0)
    assert abs(a[-1] / 20000 - 1 / 50.0) < 0.002


@pytest.mark.gpu
def test_same_bits_alone_and_in_the_mix():
    from mbsymmetry_amd import _lib
    from mbsymmetry_amd import system as S
    systems = [S.config1(), S.time_reversal_square(ns=2, ao=(5, 3, 2), aux=(12, 8, 6), ni=10, beta=12.0, seed=13),
               S.tiny(ns=2, n=2, ink=4, ao=(3, 2), aux=(4, 3), ni=8, beta=10.0, seed=11)]
    reqs = [SV.Request(i, s, 0.0, iterations=2, mu=0.02) for i, s in enumerate(systems)]
    print("[case] serving alone", flush=True)
    alone = [SV.run_system(r, 0, _lib.PREC_FP64, keep_G=True) for r in reqs]
    # the mix: all three at once plus a second copy of the first system
    mix = reqs + [SV.Request(3, systems[0], 0.0, iterations=2, mu=0.02)]
    print("[case] serving mix", flush=True)
    out = SV.serve(mix, device=0, max_active=4, keep_G=True)
    assert out["summary"]["completed"] == len(mix)
    assert out["peak_active"] >= 2
    for r in out["results"]:
        assert np.array_equal(r.G, alone[r.rid % 3]), r.rid


@pytest.mark.gpu
def test_footprint_model_matches_allocation():
    """The footprint model (admission control) against what a context really
    takes from the device, in both precisions."""
    from mbsymmetry_amd import _lib
    from mbsymmetry_amd import system as S
    sysm = S.config1()

    def settled_free():
        # the driver can report a release a little after hipFree returns
        last = None
        for _ in range(50):
            f, _ = _lib.device_memory(0)
            if f == last:
                return f
            last = f
            time.sleep(0.01)
        return last

    for prec in (_lib.PREC_FP64, _lib.PREC_FP32):
        print("[case] footprint", prec, flush=True)
        model = _lib.layout_footprint(sysm, prec)
        ctx = _lib.Context(sysm, 0, prec)
        f_in = settled_free()
        ctx.close()
        f_out = settled_free()
        used = f_out - f_in   # what destroying the context gives back
        assert abs(used - model) <= 0.1 * model + (32 << 20), (prec, used, model)


def test_results_archive_round_trip(tmp_path):
    """save_results writes the parameters under the reference's results-file
    names, G, Sigma and the history; load_results reads them back without
    unpickling anything."""
    from mbsymmetry_amd import system as S
    from mbsymmetry_amd.scloop import SCResult, save_results, load_results
    sysm = S.tiny()
    rng = np.random.default_rng(3)
    shape = (sysm.nts, sysm.ns, sysm.flat_ao)
    G = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    Sig = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    hist = [{"iteration": 1, "max_abs_dG": 0.5, "mu": 0.1, "nel": None},
            {"iteration": 2, "max_abs_dG": 1e-9, "mu": 0.2, "nel": None}]
    res = SCResult(G=G, mu=0.2, iterations=2, converged=True, history=hist, Sigma=Sig)
    path = str(tmp_path / "sim.npz")
    save_results(path, res, sysm, params={"nel_cell": 4.0, "tol": 1e-9, "damping": 0.7})
    z = load_results(path)
    assert np.array_equal(z["G_tau"], G) and np.array_equal(z["Selfenergy"], Sig)
    assert z["params/mu"] == 0.2 and z["params/nk"] == sysm.nk and z["params/nts"] == sysm.nts
    assert z["params/ns"] == sysm.ns and z["params/beta"] == sysm.transforms.beta
    assert z["params/nao"] == sum(sysm.ao_sizes[int(sysm.irre_list[0])])
    assert z["params/NQ"] == sum(sysm.aux_sizes[int(sysm.irre_list[0])])
    assert z["params/nel_cell"] == 4.0 and z["params/damping"] == 0.7
    assert bool(z["converged"]) and z["iterations"] == 2
    assert np.array_equal(z["history/max_abs_dG"], [0.5, 1e-9])
    assert np.isnan(z["history/nel"]).all()
