"""Serving loop: many GW systems on one device, continuously batched.

Requests (one system each, a fixed number of GW iterations) arrive over time.
The scheduler admits them in arrival order while the summed device footprint
of the running systems, from the layout's footprint model
(symgw_layout_footprint), stays within an HBM budget and fewer than
`max_active` systems run.  Every admitted system gets its own context (its own
HIP stream, default priority) and runs on a worker thread; the ctypes calls
release the GIL, so the systems' kernels overlap on the device.  A finished
system frees its context and its footprint at once, and the next waiting
request is admitted: batching is continuous, not in waves.

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
