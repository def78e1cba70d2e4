"""Serving loops: many GW systems on one device, continuously batched.

Two request models:
  * Server / serve_resident: systems stay resident (V or its generator, G,
    Sigma, P on the device) and a request is GW iterations of one of them;
    admission by HBM footprint with LRU eviction of idle systems.
  * Scheduler / serve: a request brings its own system, which is set up,
    iterated and released.

In both, device footprints come from the layout's footprint model
(symgw_layout_footprint), every running system has its own context (its own
HIP stream, default priority) driven from a worker thread, and the ctypes
calls release the GIL, so the systems' kernels overlap on the device.  A
finished request frees its slot at once and the next waiting request is
admitted: batching is continuous, not in waves.

Results are deterministic per system: every output element of every kernel is
written by one workgroup in a fixed order, so a system gives the same bits
alone and in the mix (tested).

There is no serving layer in the reference (SymmetryMBPT runs one system per
MPI job); the footprint check follows GW_check_devices_free_space
(SymmetryMBPT/src/solver/cugw_solver_block.cpp:471-493), applied per system.
"""
from __future__ import annotations

import threading
import time
from dataclasses import dataclass, field
from typing import Callable, List, Optional

import numpy as np

from . import _lib


@dataclass
class Request:
    rid: int
    system: object            # mbsymmetry_amd.system.System
    arrival: float            # seconds after the start of the run
    iterations: int = 1
    mu: float = 0.0


@dataclass
class Result:
    rid: int
    arrival: float
    admitted: float
    finished: float
    footprint: int
    iterations: int
    G: Optional[np.ndarray] = None
    error: Optional[str] = None

    @property
    def latency(self) -> float:
        return self.finished - self.arrival

    @property
    def queued(self) -> float:
        return self.admitted - self.arrival


def run_system(req: Request, device: int, precision: int, keep_G: bool) -> Optional[np.ndarray]:
    """Set up one system on its own context (V generated on the device, G0 from
    the Sigma = 0 Dyson solve) and run its GW iterations from per-phase
    hipGraphs; returns the final G (ts_kij) when asked."""
    s = req.system
    ctx = _lib.Context(s, device, precision)
    try:
        ctx.generate_V(s.v_seed, s.v_scale, s.decay_table())
        ctx.set_reps(s.orep, s.auxrep)
        ctx.set_transforms(s.transforms)
        ctx.set_fock(s.F, s.S)
        ctx.upload_G(s.zeros_ts_kij())
        ctx.dyson(req.mu, 0.0)
        for _ in range(req.iterations):
            ctx.gw_iteration(req.mu, use_graph=True)
        ctx.sync()
        return ctx.download_G((s.nts, s.ns, s.flat_ao)) if keep_G else None
    finally:
        ctx.close()


class Scheduler:
    """Arrival-ordered admission under an HBM budget and a concurrency cap.

    `runner(request) -> G or None` does the work of one request; it is called
    on a worker thread.  `footprint(request) -> bytes` sizes it.  Requests
    larger than the whole budget are rejected with an error result.  The head
    of the queue blocks the requests behind it (FIFO, no overtaking), so a
    large system cannot starve.
    """

    def __init__(self, runner: Callable[[Request], Optional[np.ndarray]], footprint: Callable[[Request], int],
                 hbm_budget: int, max_active: int = 8, clock: Callable[[], float] = time.perf_counter):
        if hbm_budget <= 0 or max_active < 1:
            raise ValueError("hbm_budget and max_active must be positive")
        self.runner = runner
        self.footprint = footprint
        self.budget = int(hbm_budget)
        self.max_active = int(max_active)
        self.clock = clock
        self._cv = threading.Condition()
        self._used = 0
        self._active = 0
        self.peak_used = 0
        self.peak_active = 0

    def run(self, requests: List[Request]) -> List[Result]:
        reqs = sorted(requests, key=lambda r: (r.arrival, r.rid))
        results: List[Result] = []
        threads = []
        t0 = self.clock()
        for req in reqs:
            now = self.clock() - t0
            if req.arrival > now:
                time.sleep(req.arrival - now)
            need = int(self.footprint(req))
            if need > self.budget:
                t = self.clock() - t0
                results.append(Result(req.rid, req.arrival, t, t, need, 0,
                                      error=f"footprint {need} B exceeds the HBM budget {self.budget} B"))
                continue
            with self._cv:
                while self._active >= self.max_active or self._used + need > self.budget:
                    self._cv.wait()
                self._used += need
                self._active += 1
                self.peak_used = max(self.peak_used, self._used)
                self.peak_active = max(self.peak_active, self._active)
            admitted = self.clock() - t0
            th = threading.Thread(target=self._work, args=(req, need, admitted, t0, results), daemon=True)
            th.start()
            threads.append(th)
        for th in threads:
            th.join()
        return sorted(results, key=lambda r: r.rid)

    def _work(self, req: Request, need: int, admitted: float, t0: float, results: List[Result]):
        G, err = None, None
        try:
            G = self.runner(req)
        except Exception as e:  # reported per request; the loop keeps serving
            err = f"{type(e).__name__}: {e}"
        finished = self.clock() - t0
        with self._cv:
            results.append(Result(req.rid, req.arrival, admitted, finished, need,
                                  0 if err else req.iterations, G=G, error=err))
            self._used -= need
            self._active -= 1
            self._cv.notify_all()


def serve(requests: List[Request], device: int = 0, precision: int = _lib.PREC_FP64,
          hbm_fraction: float = 0.8, max_active: int = 8, keep_G: bool = False) -> dict:
    """Run `requests` through the scheduler on `device`; the HBM budget is
    `hbm_fraction` of the device's free memory at the start.  Returns the
    per-request results and the latency summary."""
    free, _ = _lib.device_memory(device)
    foot = {}

    def footprint(r: Request) -> int:
        if r.rid not in foot:
            foot[r.rid] = _lib.layout_footprint(r.system, precision)
        return foot[r.rid]

    sch = Scheduler(lambda r: run_system(r, device, precision, keep_G), footprint, int(hbm_fraction * free),
                    max_active)
    t0 = time.perf_counter()
    res = sch.run(requests)
    wall = time.perf_counter() - t0
    return dict(results=res, wall_s=wall, summary=summarize(res, wall), peak_bytes=sch.peak_used,
                peak_active=sch.peak_active, budget=sch.budget)


# ---------------------------------------------------------------------------
# Resident serving: systems stay loaded, requests are GW iterations on them
# ---------------------------------------------------------------------------
@dataclass
class IterRequest:
    rid: int
    system: int               # index into the server's systems
    arrival: float            # seconds after the start of the run
    iterations: int = 1


@dataclass
class IterResult:
    rid: int
    system: int
    arrival: float
    started: float
    finished: float
    iterations: int
    cold: bool                # the system had to be loaded for this request
    G: Optional[np.ndarray] = None
    error: Optional[str] = None
    served: bool = True       # False: not started before the deadline

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


class Server:
    """Continuous batching of GW-iteration requests over resident systems.

    Every system is loaded at most once while it stays resident: its context
    keeps V (or its generator, for systems served with a streamed V), G,
    Sigma and P on the device between requests, so a request is one or more
    GW iterations from where the system's last request left it.  Admission:
    queued requests are taken shortest estimated work first, a request's
    estimate reduced by the seconds it has waited (so no request waits
    forever; with no estimates the order is the arrival order); a request
    whose system is loaded and idle starts at once (up to `max_active`
    requests run together, each system one request at a time, and at most
    `max_long` of them estimated longer than `long_s` seconds, so that long
    requests never hold every slot); a request
    whose system is not loaded loads it when its footprint fits the HBM
    budget, first evicting loaded idle systems, least recently used first;
    only the first such request in that order may claim memory, so a large
    system cannot be overtaken for ever by smaller ones.  `preload` loads
    systems before the clock starts, so that request latency excludes setup.
    With a deadline, a request starts only if it is estimated to end before
    it: from now, the lesser of its estimated seconds (`estimate(system) *
    iterations`) plus those of the running requests, and its own times
    `max_active` (the device shared by every slot); a request that would not
    end in time even alone is returned unserved.

    backend: load(i) -> handle (the set-up system), iterate(handle, n, keep_G)
    -> G or None, unload(handle).
    """

    def __init__(self, footprints: List[int], budget: int, backend, max_active: int = 8,
                 clock: Callable[[], float] = time.perf_counter,
                 estimate: Optional[Callable[[int], float]] = None, long_s: float = 1.0,
                 max_long: Optional[int] = None):
        if budget <= 0 or max_active < 1:
            raise ValueError("budget and max_active must be positive")
        self.long_s = float(long_s)
        self.max_long = max(1, max_active // 4) if max_long is None else int(max_long)
        self.active_long = 0
        self.fp = [int(f) for f in footprints]
        self.budget = int(budget)
        self.backend = backend
        self.max_active = int(max_active)
        self.clock = clock
        self.estimate = estimate or (lambda i: 0.0)
        n = len(self.fp)
        self.handle = [None] * n
        self.busy = [False] * n
        self.reserved = [False] * n   # memory claimed by a load in progress
        self.last_used = [0.0] * n
        self.used = 0
        self.active = 0
        self.outstanding = 0.0   # estimated seconds of the running requests
        self.peak_used = 0
        self.peak_active = 0
        self.loads = 0
        self.evictions = 0
        self.setup_s = 0.0
        self._cv = threading.Condition()

    def preload(self, order: List[int]) -> float:
        """Load systems in `order` while they fit the budget; returns the set-up seconds."""
        t0 = self.clock()
        for i in order:
            if self.handle[i] is None and self.used + self.fp[i] <= self.budget:
                self.handle[i] = self.backend.load(i)
                self.used += self.fp[i]
                self.loads += 1
        self.peak_used = max(self.peak_used, self.used)
        dt = self.clock() - t0
        self.setup_s += dt
        return dt

    def _evict_for(self, need: int, keep: int) -> bool:
        """Unload idle loaded systems (LRU first) until `need` bytes fit; False if they cannot."""
        idle = sorted((self.last_used[j], j) for j in range(len(self.fp))
                      if self.handle[j] is not None and not self.busy[j] and j != keep)
        free = self.budget - self.used + sum(self.fp[j] for _, j in idle)
        if free < need:
            return False
        for _, j in idle:
            if self.used + need <= self.budget:
                break
            self.backend.unload(self.handle[j])
            self.handle[j] = None
            self.used -= self.fp[j]
            self.evictions += 1
        return self.used + need <= self.budget

    def run(self, requests: List[IterRequest], time_limit: Optional[float] = None,
            keep_G: bool = False, deadline: Optional[float] = None) -> List[IterResult]:
        """Serve `requests` (arrivals after `time_limit` s are not issued; no
        request starts that is estimated to end after `deadline` s); returns
        one result per issued request."""
        reqs = sorted(requests, key=lambda r: (r.arrival, r.rid))
        if time_limit is not None:
            reqs = [r for r in reqs if r.arrival <= time_limit]
        results: List[IterResult] = []
        queue: List[IterRequest] = []
        threads = []
        t0 = self.clock()
        nxt = 0
        with self._cv:
            while nxt < len(reqs) or queue or self.active:
                now = self.clock() - t0
                while nxt < len(reqs) and reqs[nxt].arrival <= now:
                    r = reqs[nxt]
                    nxt += 1
                    if self.fp[r.system] > self.budget:
                        results.append(IterResult(r.rid, r.system, r.arrival, now, now, 0, False,
                                                  error=f"footprint {self.fp[r.system]} B exceeds the budget"))
                    else:
                        queue.append(r)
                started = []
                claimed = False
                # shortest estimated work first, aged by the time waited
                queue.sort(key=lambda r: (self.estimate(r.system) * r.iterations - (now - r.arrival), r.arrival))
                for r in queue:
                    if self.active >= self.max_active:
                        break
                    i = r.system
                    if self.busy[i]:
                        continue
                    est = self.estimate(i) * r.iterations
                    # end estimate: within the outstanding work plus its own
                    # (work-conserving device), or its own stretched by sharing
                    # the device with every slot
                    if deadline is not None and now + min(self.outstanding + est, self.max_active * est) > deadline:
                        continue
                    long = est > self.long_s
                    if long and self.active_long >= self.max_long:
                        continue           # the other slots stay open to short requests
                    cold = self.handle[i] is None
                    if cold:
                        if claimed:      # only the first cold request in this order may claim memory
                            continue
                        claimed = True
                        if not self._evict_for(self.fp[i], i):
                            continue
                        self.used += self.fp[i]
                        self.reserved[i] = True
                        self.peak_used = max(self.peak_used, self.used)
                    self.busy[i] = True
                    self.active += 1
                    self.active_long += long
                    self.outstanding += est
                    self.peak_active = max(self.peak_active, self.active)
                    th = threading.Thread(target=self._work, args=(r, cold, t0, keep_G, results, est, long),
                                          daemon=True)
                    threads.append(th)
                    started.append(r)
                    th.start()
                for r in started:
                    queue.remove(r)
                if deadline is not None:
                    now = self.clock() - t0
                    late = [r for r in queue if now + self.estimate(r.system) * r.iterations > deadline]
                    for r in late:
                        queue.remove(r)
                        results.append(IterResult(r.rid, r.system, r.arrival, now, now, 0, False,
                                                  error="not started: estimated to end after the deadline",
                                                  served=False))
                wait = None
                if nxt < len(reqs):
                    wait = max(0.0, reqs[nxt].arrival - (self.clock() - t0))
                if nxt < len(reqs) or queue or self.active:
                    self._cv.wait(timeout=wait if wait is not None else 0.5)
        for th in threads:
            th.join()
        return sorted(results, key=lambda r: r.rid)

    def _work(self, r: IterRequest, cold: bool, t0: float, keep_G: bool, results: List[IterResult], est: float,
              long: bool):
        i = r.system
        started = self.clock() - t0
        G, err = None, None
        try:
            if cold:
                ts = self.clock()
                h = self.backend.load(i)
                with self._cv:
                    self.handle[i] = h
                    self.reserved[i] = False
                    self.loads += 1
                    self.setup_s += self.clock() - ts
            G = self.backend.iterate(self.handle[i], r.iterations, keep_G)
        except Exception as e:  # reported per request; the server keeps serving
            err = f"{type(e).__name__}: {e}"
        finished = self.clock() - t0
        with self._cv:
            if err is not None and self.handle[i] is None and self.reserved[i]:
                self.used -= self.fp[i]       # a failed load gives its memory back
                self.reserved[i] = False
            results.append(IterResult(r.rid, i, r.arrival, started, finished, 0 if err else r.iterations, cold,
                                      G=G, error=err))
            self.busy[i] = False
            self.last_used[i] = finished
            self.active -= 1
            self.active_long -= long
            self.outstanding -= est
            self._cv.notify_all()

    def close(self):
        for i, h in enumerate(self.handle):
            if h is not None:
                self.backend.unload(h)
                self.handle[i] = None
        self.used = 0


class DeviceBackend:
    """Server backend on a device: load = context + V (or its generator) +
    reps, transforms, Fock and G0 from the Sigma = 0 Dyson solve; iterate =
    GW iterations replayed from the context's per-phase hipGraphs."""

    def __init__(self, systems, device: int = 0, precision: int = _lib.PREC_FP64, stream=None, mu: float = 0.0):
        self.systems = systems
        self.device = device
        self.precision = precision
        self.stream = stream or [False] * len(systems)
        self.mu = mu

    def precision_of(self, i: int) -> int:
        return self.precision | (_lib.FLAG_STREAM_V if self.stream[i] else 0)

    def load(self, i: int):
        s = self.systems[i]
        ctx = _lib.Context(s, self.device, self.precision_of(i))
        try:
            ctx.generate_V(s.v_seed, s.v_scale, s.decay_table())
            ctx.set_reps(s.orep, s.auxrep)
            ctx.set_transforms(s.transforms)
            ctx.set_fock(s.F, s.S)
            ctx.upload_G(s.zeros_ts_kij())
            ctx.dyson(self.mu, 0.0)
        except Exception:
            ctx.close()
            raise
        return (i, ctx)

    def iterate(self, handle, n: int, keep_G: bool):
        i, ctx = handle
        s = self.systems[i]
        for _ in range(n):
            ctx.gw_iteration(self.mu, use_graph=True)
        ctx.sync()
        return ctx.download_G((s.nts, s.ns, s.flat_ao)) if keep_G else None

    def unload(self, handle):
        handle[1].close()


def serve_resident(systems, requests: List[IterRequest], device: int = 0, precision: int = _lib.PREC_FP64,
                   hbm_fraction: float = 0.8, max_active: int = 8, time_limit: Optional[float] = None,
                   deadline: Optional[float] = None, rate_flops: float = 20e12, keep_G: bool = False) -> dict:
    """Serve GW-iteration requests over resident systems on `device`.  A
    system whose resident footprint exceeds half the budget is served
    with a streamed V.  Systems are preloaded in the order of their first
    request while they fit (set-up time reported apart); latency is arrival ->
    new G available.  With a deadline, a request's duration is estimated as
    its P0 + Sigma FLOPs over `rate_flops`."""
    free, _ = _lib.device_memory(device)
    budget = int(hbm_fraction * free)
    stream = []
    fp = []
    flops = []
    for s in systems:
        f = _lib.layout_footprint(s, precision)
        st = f > budget // 2
        if st:
            f = _lib.layout_footprint(s, precision | _lib.FLAG_STREAM_V)
        stream.append(st)
        fp.append(f)
        q = _lib.layout_query(s)
        flops.append(float(q.flop_P0 + q.flop_Sigma))
    be = DeviceBackend(systems, device, precision, stream)
    srv = Server(fp, budget, be, max_active, estimate=lambda i: flops[i] / rate_flops)
    first = []
    for r in sorted(requests, key=lambda r: r.arrival):
        if r.system not in first:
            first.append(r.system)
    setup = srv.preload(first)
    t0 = time.perf_counter()
    try:
        res = srv.run(requests, time_limit=time_limit, keep_G=keep_G, deadline=deadline)
    finally:
        wall = time.perf_counter() - t0
        srv.close()
    sm = summarize_iter(res, wall, systems, deadline)
    sm["P0_Sigma_TFLOP_served"] = sum(flops[r.system] * r.iterations for r in res if r.error is None) / 1e12
    sm["P0_Sigma_tflops"] = sm["P0_Sigma_TFLOP_served"] / wall if wall > 0 else 0.0
    return dict(results=res, wall_s=wall, setup_s=setup, summary=sm, peak_bytes=srv.peak_used,
                peak_active=srv.peak_active, budget=budget, loads=srv.loads, evictions=srv.evictions,
                streamed=sum(stream), flops=flops)


def _key(s) -> str:
    return f"{round(s.nk ** (1 / 3))}^3-nao{sum(s.ao_sizes[int(s.irre_list[0])])}"


def summarize_iter(results: List[IterResult], wall_s: float, systems, deadline: Optional[float] = None) -> dict:
    """Latency summary.  p50/p99 over the completed requests; with unserved
    requests (deadline) also over all issued requests, an unserved one counted
    at (deadline - arrival), which makes those quantiles lower bounds."""
    ok = [r for r in results if r.error is None]
    unserved = [r for r in results if not r.served]
    failed = [r for r in results if r.error is not None and r.served]
    lat = np.array([r.latency for r in ok]) if ok else np.full(1, np.nan)
    warm_l = [r.latency for r in ok if not r.cold]
    warm = np.array(warm_l) if warm_l else np.full(1, np.nan)
    its = sum(r.iterations for r in ok)
    by_size = {}
    for r in results:
        d = by_size.setdefault(_key(systems[r.system]), {"requests": 0, "completed": 0, "latencies": []})
        d["requests"] += 1
        if r.error is None:
            d["completed"] += 1
            d["latencies"].append(r.latency)
    by = {k: {"requests": v["requests"], "completed": v["completed"],
              "median_latency_s": float(np.median(v["latencies"])) if v["latencies"] else None}
          for k, v in sorted(by_size.items())}
    out = {"requests": len(results), "completed": len(ok), "failed": len(failed), "unserved": len(unserved),
           "cold_starts": sum(1 for r in ok if r.cold),
           "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
           "p50_warm_latency_s": float(np.percentile(warm, 50)), "p99_warm_latency_s": float(np.percentile(warm, 99)),
           "gw_iter_per_s": its / wall_s if wall_s > 0 else 0.0, "wall_s": wall_s, "by_system": by,
           "errors": sorted({r.error for r in failed})[:4]}
    if unserved and deadline is not None:
        allv = np.array([r.latency for r in ok] + [max(0.0, deadline - r.arrival) for r in unserved])
        out["p50_latency_all_issued_lower_bound_s"] = float(np.percentile(allv, 50))
        out["p99_latency_all_issued_lower_bound_s"] = float(np.percentile(allv, 99))
    return out


def summarize(results: List[Result], wall_s: float) -> dict:
    ok = [r for r in results if r.error is None]
    lat = np.array([r.latency for r in ok]) if ok else np.zeros(1)
    q = np.array([r.queued for r in ok]) if ok else np.zeros(1)
    its = sum(r.iterations for r in ok)
    return {"requests": len(results), "completed": len(ok), "failed": len(results) - len(ok),
            "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
            "mean_queue_s": float(q.mean()), "gw_iter_per_s": its / wall_s if wall_s > 0 else 0.0,
            "wall_s": wall_s, "errors": sorted({r.error for r in results if r.error})[:4]}


def poisson_arrivals(n: int, rate_per_s: float, seed: int) -> np.ndarray:
    """Arrival times (s) of a Poisson process: cumulative exponential gaps."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate_per_s, n))
