"""Autotuner of the contraction work lists.

The planner (layout.cpp) splits every P0 and Sigma block contraction into
slabs of consecutive times and LDS chunks of aux rows and orbitals, choosing
among the shapes that fit by a cost model: executed MFMA FLOPs (16 x 16 x 4
tiles, padding included) plus a fixed cost per chunk for its barriers and
staging.  How large that fixed cost really is, how many aux rows a Sigma
chunk may take, how many times a Sigma slab may stack and how many output
rows a P0 slab may take depend on the block sizes, so they are the knobs
(symgw_plan) this module times on the device, per block-size histogram:

  * Sigma: sig_chunk_cost x sig_rows x sig_max_times;
  * P0:    p0_chunk_cost x p0_band_cap;

the two sweeps tuned independently (their knobs do not interact).  Each
candidate is applied with symgw_replan and timed with HIP events on a shard
of iBZ points sized to about `sample_s` seconds of work (the whole sweep when
that is shorter), after one untimed sweep of the default plan; the pick is
confirmed against the built-in rules on whole sweeps where those are short.
The fastest plan is applied and cached under the histogram's key (JSON; path from the
SYMGW_TUNE_CACHE environment variable, default ~/.cache/symgw/plans.json), so
a later context with the same block structure starts from it.

Tuning overwrites the device P0/P and Sigma buffers (it runs the sweeps):
tune a context after its V and G are set and before its first iteration.
There is no counterpart in the reference, which hands each block to cuBLAS
(SymmetryMBPT/src/cuda/cugw_qkpt_block.cu:233-260)."""
from __future__ import annotations

import collections
import hashlib
import json
import math
import os
import time
from typing import Callable, Dict, List, Optional

from . import _lib

SIG_CANDIDATES = [dict(sig_chunk_cost=c, sig_rows=r, sig_max_times=t)
                  for c in (6.25e4, 2.5e5, 1e6, 4e6) for r in (32, 64) for t in (1, 2, 4)]
P0_CANDIDATES = [dict(p0_chunk_cost=c, p0_band_cap=b) for c in (6.25e4, 2.5e5, 1e6, 4e6) for b in (8, 16)]


def histogram_key(sysm, precision: int) -> str:
    """Key of a system's block structure: the counts of its Sigma block shapes
    (k1, q, k2) and P0 block shapes (q, k1, k2), nts, ns and the precision."""
    nk = sysm.nk
    hist = collections.Counter()
    for k in sysm.irre_list:
        k = int(k)
        for q in range(nk):
            p = k * nk + int(sysm.kmq[k, q])
            for qs, a1, a2 in sysm.V_block_sizes(p):
                hist[("S", a1, qs, a2)] += 1
    for q in sysm.irre_list:
        q = int(q)
        for k in range(nk):
            p = k * nk + int(sysm.kpq[k, q])
            for qs, a1, a2 in sysm.V_block_sizes(p):
                hist[("P", qs, a1, a2)] += 1
    text = json.dumps([sorted(hist.items()), sysm.nts, sysm.ns, precision & 0xFF])
    return hashlib.sha256(text.encode()).hexdigest()[:24]


def _cache_path() -> str:
    return os.environ.get("SYMGW_TUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "symgw", "plans.json"))


def _load_cache(path: str) -> dict:
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(path: str, key: str, record: dict) -> None:
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        cache = _load_cache(path)
        cache[key] = record
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(cache, f, indent=1)
        os.replace(tmp, path)
    except OSError:
        pass  # a read-only home: the plan still applies to this context


def _timed(ctx, which: str) -> float:
    ctx.set_timing(True)
    if which == "P0":
        ctx.solve_P0()
    else:
        ctx.solve_Sigma()
    ctx.sync()
    ms = ctx.last_times()[0 if which == "P0" else 2]
    ctx.set_timing(False)
    return float(ms)


def tune(ctx, sysm, precision: int, sample_s: float = 0.4, use_cache: bool = True,
         log: Optional[Callable[[str], None]] = None, validate_s: float = 12.0) -> dict:
    """Tune the plan of `ctx` (a context of `sysm` in `precision` with V and G
    set) and leave the best plan applied.  A shard's pick is confirmed on
    whole sweeps when one takes at most `validate_s` seconds.  Returns a
    record of the search."""
    say = log or (lambda m: None)
    key = histogram_key(sysm, precision)
    path = _cache_path()
    if use_cache:
        hit = _load_cache(path).get(key)
        if hit:
            ctx.replan(**hit["plan"])
            say(f"autotune: cached plan {hit['plan']}")
            return dict(hit, cached=True)
    info = ctx.info()
    ink = sysm.ink
    t0 = time.perf_counter()
    out: Dict[str, object] = {"key": key, "cached": False, "candidates": {}}
    best_plan: Dict[str, float] = {}
    ctx.replan()
    ctx.solve_P0()
    ctx.solve_W()      # P from P0, so the Sigma sweeps (tuned first) read a physical P
    full_ms = {}
    for which, cands, flop in (("Sigma", SIG_CANDIDATES, info.flop_Sigma), ("P0", P0_CANDIDATES, info.flop_P0)):
        # a shard of about sample_s seconds (the whole sweep if shorter): first
        # sized at 30 TFLOP/s, then from the default plan's measured time
        count = max(1, min(ink, math.ceil(ink * sample_s / max(flop / 30e12, 1e-9))))
        ctx.replan()
        for _ in range(2):
            ctx.set_shard(0, count)
            _timed(ctx, which)                      # untimed warm-up of the default plan
            default_ms = _timed(ctx, which)
            if count == ink:
                full_ms[which] = default_ms
            want = max(1, min(ink, math.floor(count * 1e3 * sample_s / max(default_ms, 1e-3))))
            if want >= count or default_ms <= 2e3 * sample_s:
                break
            count = want
        results = []
        for c in cands:
            try:
                ctx.replan(**c)
            except _lib.SymGWError as e:             # a plan that does not fit this system
                results.append((float("inf"), c, str(e)))
                continue
            results.append((_timed(ctx, which), c, None))
        results.sort(key=lambda r: r[0])
        best_ms, best, _ = results[0]
        if best_ms >= default_ms:                   # keep the built-in rules unless beaten
            best_ms, best = default_ms, {}
        best_plan.update(best)
        out["candidates"][which] = [{"ms": ms, **c, **({"error": e} if e else {})} for ms, c, e in results]
        out[which] = {"shard_points": count, "default_ms": default_ms, "best_ms": best_ms,
                      "speedup": default_ms / best_ms if best_ms > 0 else 1.0, "plan": best}
        say(f"autotune {which}: shard of {count}/{ink} points, default {default_ms:.1f} ms, best {best_ms:.1f} ms "
            f"{best}")
    ctx.set_shard(0, ink)
    if best_plan and validate_s > 0:
        # the shard's choice against the built-in rules on whole sweeps (a
        # shard's grid is smaller: its tail can favour other shapes); the P0
        # sweeps above overwrote P, so the Sigma sweeps get it back first
        ctx.replan()
        ctx.solve_P0()
        ctx.solve_W()
        for which in ("Sigma", "P0"):
            picked = {k: v for k, v in best_plan.items() if (k.startswith("sig") == (which == "Sigma"))}
            if not picked:
                continue
            est = full_ms.get(which) or out[which]["default_ms"] * ink / out[which]["shard_points"]
            if est > 1e3 * validate_s:
                continue
            ctx.replan()
            d = full_ms.get(which) or _timed(ctx, which)
            ctx.replan(**picked)
            b = _timed(ctx, which)
            out[which]["whole_sweep_ms"] = {"default": d, "picked": b}
            if b >= d:
                for k in picked:
                    best_plan.pop(k)
                out[which]["plan"] = {}
                say(f"autotune {which}: whole sweep {b:.1f} ms vs {d:.1f} ms default: built-in rules kept")
    ctx.replan(**best_plan)
    out["plan"] = ctx.plan()
    out["seconds"] = time.perf_counter() - t0
    if use_cache:
        _store(path, key, {k: v for k, v in out.items() if k != "cached"})
    return out
