"""Serving loops: admission policy on the CPU (stand-in runners and
backends), and on the GPU the same bits for a system alone and in a
concurrent mix, for both request models (resident systems, and requests that
bring their own system)."""
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


class _FakeBackend:
    """Backend that sleeps instead of computing; records loads and overlaps."""

    def __init__(self, runtime, fail_load=()):
        self.runtime = runtime
        self.fail_load = set(fail_load)
        self.lock = threading.Lock()
        self.loaded = set()
        self.load_calls = []
        self.running = {"n": 0, "max": 0}
        self.per_system = {}

    def load(self, i):
        if i in self.fail_load:
            raise RuntimeError(f"cannot load {i}")
        with self.lock:
            assert i not in self.loaded
            self.loaded.add(i)
            self.load_calls.append(i)
        return i

    def iterate(self, h, n, keep_G):
        with self.lock:
            assert h in self.loaded
            self.running["n"] += 1
            self.running["max"] = max(self.running["max"], self.running["n"])
            self.per_system[h] = self.per_system.get(h, 0) + 1
            assert self.per_system[h] == 1, "two requests of one system at once"
        time.sleep(self.runtime[h] * n)
        with self.lock:
            self.running["n"] -= 1
            self.per_system[h] -= 1
        return np.full(2, float(h)) if keep_G else None

    def unload(self, h):
        with self.lock:
            self.loaded.remove(h)


def test_server_resident_systems_budget_and_lru():
    fp = [400, 300, 300, 500, 2000]          # system 4 exceeds the budget
    be = _FakeBackend({i: 0.02 for i in range(5)})
    srv = SV.Server(fp, budget=1000, backend=be, max_active=3)
    assert srv.preload([0, 1, 2, 3]) >= 0.0
    assert be.load_calls == [0, 1, 2] and srv.used == 1000     # 3 does not fit
    reqs = [SV.IterRequest(0, 0, 0.0), SV.IterRequest(1, 1, 0.0), SV.IterRequest(2, 0, 0.0),
            SV.IterRequest(3, 3, 0.05), SV.IterRequest(4, 4, 0.05), SV.IterRequest(5, 2, 0.2, iterations=2)]
    res = srv.run(reqs, keep_G=True)
    assert [r.rid for r in res] == list(range(6))
    assert res[4].error and "exceeds" in res[4].error
    ok = [r for r in res if r.error is None]
    assert len(ok) == 5
    # system 3 was loaded cold, evicting least recently used idle systems
    assert res[3].cold and not res[0].cold and not res[1].cold
    assert srv.evictions >= 1 and srv.peak_used <= 1000
    # requests of one system never overlap; different systems do
    assert be.running["max"] >= 2
    assert res[2].started >= res[0].finished - 1e-6
    # system 2 was evicted for system 3, so its later request reloads it
    assert res[5].cold and res[5].iterations == 2 and np.array_equal(res[5].G, [2.0, 2.0])
    srv.close()
    assert not be.loaded and srv.used == 0
    s = SV.summarize_iter(res, 1.0, [type("Sys", (), {"nk": 8, "ao_sizes": [[i + 1]], "irre_list": [0]})()
                                     for i in range(5)])
    assert s["completed"] == 5 and s["failed"] == 1 and s["cold_starts"] == 2
    assert s["p99_latency_s"] >= s["p50_latency_s"] and s["unserved"] == 0


def test_server_oldest_cold_request_claims_memory_first():
    """A large cold system at the head of the queue is not overtaken by later
    small cold systems (no starvation); warm systems keep running meanwhile."""
    fp = [600, 300, 300, 300]
    be = _FakeBackend({0: 0.15, 1: 0.01, 2: 0.01, 3: 0.01})
    srv = SV.Server(fp, budget=1000, backend=be, max_active=4)
    srv.preload([1])
    reqs = [SV.IterRequest(0, 2, 0.0), SV.IterRequest(1, 3, 0.0), SV.IterRequest(2, 0, 0.001),
            SV.IterRequest(3, 2, 0.002)]
    res = srv.run(reqs)
    assert all(r.error is None for r in res)
    # 2 and 3 fill the budget (300 + 300 + 300 resident); 0 (600) needs evictions
    # and starts once they finish, before any later cold claim
    assert res[2].cold and res[2].started >= min(res[0].finished, res[1].finished) - 1e-6
    assert srv.peak_used <= 1000
    srv.close()


def test_server_deadline_leaves_long_requests_unserved():
    fp = [10, 10]
    be = _FakeBackend({0: 0.01, 1: 0.01})
    est = {0: 0.01, 1: 5.0}
    srv = SV.Server(fp, budget=100, backend=be, max_active=2, estimate=lambda i: est[i])
    srv.preload([0, 1])
    reqs = [SV.IterRequest(i, i % 2, 0.001 * i) for i in range(6)]
    res = srv.run(reqs, deadline=0.5)
    assert all(r.served for r in res if r.system == 0) and all(r.error is None for r in res if r.system == 0)
    assert all(not r.served and "deadline" in r.error for r in res if r.system == 1)
    s = SV.summarize_iter(res, 0.5, [type("Sys", (), {"nk": 8, "ao_sizes": [[4]], "irre_list": [0]})()] * 2,
                          deadline=0.5)
    assert s["unserved"] == 3 and s["completed"] == 3
    assert s["p99_latency_all_issued_lower_bound_s"] >= 0.49
    srv.close()


def test_server_long_requests_leave_slots_to_short_ones():
    fp = [10] * 6
    be = _FakeBackend({0: 0.2, 1: 0.2, 2: 0.2, 3: 0.01, 4: 0.01, 5: 0.01})
    est = {0: 5.0, 1: 5.0, 2: 5.0, 3: 0.01, 4: 0.01, 5: 0.01}
    srv = SV.Server(fp, budget=100, backend=be, max_active=4, estimate=lambda i: est[i], long_s=1.0, max_long=2)
    srv.preload(list(range(6)))
    reqs = [SV.IterRequest(i, i, 0.0) for i in range(3)] + [SV.IterRequest(3 + i, 3 + i, 0.01) for i in range(3)]
    res = srv.run(reqs)
    assert all(r.error is None for r in res)
    # two long requests ran first, the third waited for one of them; the short
    # ones arriving later took the open slots at once
    longs = sorted(res[:3], key=lambda r: r.started)
    assert longs[2].started >= min(longs[0].finished, longs[1].finished) - 1e-6
    for r in res[3:]:
        assert r.started - r.arrival < 0.1 and r.finished < longs[0].finished
    srv.close()


def test_server_failed_load_releases_its_memory():
    be = _FakeBackend({0: 0.01, 1: 0.01}, fail_load=[0])
    srv = SV.Server([600, 600], budget=1000, backend=be, max_active=2)
    res = srv.run([SV.IterRequest(0, 0, 0.0), SV.IterRequest(1, 1, 0.01)])
    assert res[0].error and "cannot load" in res[0].error
    assert res[1].error is None and res[1].cold
    assert srv.used == 600
    srv.close()


@pytest.mark.gpu
def test_resident_same_bits_alone_and_in_the_mix():
    """Resident serving: a system's G after its requests (interleaved with the
    other systems' on concurrent contexts) equals the G of the same number of
    iterations run alone, bit for bit.  One system is of serving-mix size:
    4x4x4 mesh (8 iBZ points), nao 156 in 8 blocks, naux 624, nts 50."""
    from mbsymmetry_amd import _lib
    from mbsymmetry_amd import system as S
    from mbsymmetry_amd.transforms import ir
    big = S.config5_system(2, 4, 156, ir(50.0, 1e3, 48))
    systems = [big, S.config1(),
               S.time_reversal_square(ns=2, ao=(5, 3, 2), aux=(12, 8, 6), ni=10, beta=12.0, seed=13),
               S.tiny(ns=2, n=2, ink=4, ao=(3, 2), aux=(4, 3), ni=8, beta=10.0, seed=11)]
    be = SV.DeviceBackend(systems, 0, _lib.PREC_FP64)
    alone = []
    for i in range(len(systems)):
        print(f"[case] resident alone: system {i} ({systems[i].name})", flush=True)
        h = be.load(i)
        alone.append(be.iterate(h, 2, True))
        be.unload(h)
    # two one-iteration requests per system, interleaved
    reqs = [SV.IterRequest(j, j % 4, 0.0 if j < 4 else 0.05) for j in range(8)]
    fp = [_lib.layout_footprint(s, _lib.PREC_FP64) for s in systems]
    srv = SV.Server(fp, budget=sum(fp) + (1 << 30), backend=be, max_active=4)
    print("[case] resident mix", flush=True)
    srv.preload(list(range(4)))
    res = srv.run(reqs, keep_G=True)
    srv.close()
    assert all(r.error is None for r in res), [r.error for r in res]
    assert srv.peak_active >= 2
    for r in res[4:]:
        assert np.array_equal(r.G, alone[r.system]), r.system


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
