"""Benchmark: GW iterations/s of the symmetry-adapted GW hot path on one MI355X.

One step = one full GW iteration of one system, resident on the device:
P0 (block contraction) -> W ((1 - P0)^-1 P0 per bosonic frequency, with the
FP64 tau <-> i Omega transforms) -> Sigma (block contraction) -> Dyson (FP64
transforms + per-frequency inversions) -> rotation of the new G to the full
Brillouin zone.  All inputs (V, G, transforms, F, S) are in HBM before timing;
the step sequence is replayed as one hipGraph.

Default workload: the 6x6x6-mesh system reduced to 16 irreducible k points,
nao = 104 in 8 irrep blocks, naux = 400 in 8 blocks, nts = 106, ns = 1 (V is
52 GiB of complex128 generated on the device).

Prints one JSON line (see the keys below).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"fp64": 78.6, "fp32": 157.3, "int8": 5000.0, "hbm_gbps": 8000.0}  # MI355X datasheet: dense matrix
# TFLOP/s (INT8 TOPS), HBM3E GB/s


_T0 = time.perf_counter()


def log(*a):
    """Progress line on stderr, stamped with the seconds since the bench started."""
    print("[bench]", f"{time.perf_counter() - _T0:6.1f}s", *a, file=sys.stderr, flush=True)


def make_system(name):
    from mbsymmetry_amd import system as S
    if name == "config2":
        return S.config2()
    if name == "config1":
        return S.config1()
    raise SystemExit(f"unknown config {name}")


def workload_name(sysm):
    ao = sysm.ao_sizes[0]
    aux = sysm.aux_sizes[0]
    n = round(sysm.nk ** (1 / 3))
    return (f"{sysm.name}: {n}x{n}x{n} k-mesh ({sysm.ink} iBZ points), nao {sum(ao)} in blocks {ao}, "
            f"naux {sum(aux)} in blocks {aux}, nts {sysm.nts}, ns {sysm.ns}, {sysm.transforms.kind} grid; "
            f"one GW iteration = P0 + W + Sigma + Dyson")


def setup_ctx(sysm, precision, mu):
    from mbsymmetry_amd import _lib
    ctx = _lib.Context(sysm, 0, precision)
    ctx.generate_V(sysm.v_seed, sysm.v_scale, sysm.decay_table())
    ctx.set_reps(sysm.orep, sysm.auxrep)
    ctx.set_transforms(sysm.transforms)
    ctx.set_fock(sysm.F, sysm.S)
    return ctx


def run_path(sysm, precision, steps, warmup, mu, G0, keep_P=False, tune=True, plan=None):
    """Time `steps` GW iterations replayed from per-phase hipGraphs.

    The first iteration from G0 is untimed (it is warmup step 1) and gives the
    Sigma and G compared across precisions.  In the timed region HIP events sit
    between the phase graphs on the context stream, so the per-phase device
    times are averages over exactly the timed iterations.  keep_P: also return
    that iteration's P0 and P (recomputed from G0 after it, outside timing).
    tune: pick the work-list plan with the autotuner first (outside timing);
    plan: a fixed plan (dict of symgw_plan rules) instead."""
    ctx = setup_ctx(sysm, precision, mu)
    ctx.upload_G(G0)
    tuned = None
    if plan is not None:
        ctx.replan(**plan)
        tuned = {"plan": ctx.plan(), "given": True}
    elif tune:
        from mbsymmetry_amd import autotune
        tuned = autotune.tune(ctx, sysm, precision, use_cache=False, log=log)
        tuned = {k: tuned[k] for k in ("plan", "Sigma", "P0", "seconds")}
        ctx.upload_G(G0)
    ctx.gw_iteration(mu, use_graph=True)
    ctx.sync()
    sig = ctx.copy_Sigma_to_host(sysm.zeros_ts_kij()) / sysm.nk
    G1 = ctx.download_G(G0.shape)
    P0 = P = None
    if keep_P:
        ctx.upload_G(G0)
        ctx.solve_P0()
        P0 = ctx.copy_P0_to_host(sysm.zeros_q_t_QP()) / sysm.nk
        ctx.solve_W()
        P = ctx.copy_P0_to_host(sysm.zeros_q_t_QP())
        ctx.upload_G(G1)
    log("warmup 1 done (iteration from G0)")
    for i in range(1, warmup):
        ctx.gw_iteration(mu, use_graph=True)
        ctx.sync()
        log(f"warmup {i + 1}/{warmup} done")
    ctx.set_timing(True)
    ctx.sync()
    phase = np.zeros(4)
    step_s = []
    t0 = time.perf_counter()
    for i in range(steps):
        ts = time.perf_counter()
        ctx.gw_iteration(mu, use_graph=True)
        ctx.sync()
        step_s.append(time.perf_counter() - ts)
        phase += np.asarray(ctx.last_times()[:4])
    dt = time.perf_counter() - t0
    log(f"timed {steps} iterations: {dt:.3f} s")
    info = ctx.info()
    ctx.close()
    from mbsymmetry_amd import _lib
    from mbsymmetry_amd.transforms import leakage
    return dict(sec_per_iter=dt / steps, phase_ms=phase / steps, step_s=step_s, flop_p0=info.flop_P0,
                flop_sig=info.flop_Sigma, sigma=sig, G1=G1, P0=P0, P=P, leakage=leakage(sysm.transforms, G1),
                tuned=tuned,
                int8_ops=int8_ops(sysm) if precision == _lib.PREC_MIXED else 0.0)


def int8_ops(sysm):
    """INT8 MFMA operations of one mixed-mode Sigma sweep, a lower bound: per
    16 x 16 Z tile and 64-deep chunk of K (kp, the aux block rounded up to 16),
    the 60 instructions of 2 * 16 * 16 * 64 ops that the kernel issues (15
    slice pairs for each of P1, P2 twice and P3), for the blocks above the
    INT8 threshold, with the tiles counted over whole blocks (the padding of
    each LDS chunk's rows to 16 is left out)."""
    q_min = int(os.environ.get("SYMGW_Z8_MIN_Q", "32"))
    tot = 0.0
    nts = sysm.nts
    for k in sysm.irre_list:
        k = int(k)
        for q in range(sysm.nk):
            k2 = int(sysm.kmq[k, q])
            for (qs, a1, a2) in sysm.V_block_sizes(k * sysm.nk + k2):
                if qs <= q_min:
                    continue
                kp = (qs + 15) // 16 * 16
                inst = 60 * ((kp + 63) // 64)
                # Z rows: times x aux rows; columns: a1 x a2, one pass per slab band of a1
                tiles = ((nts * qs + 15) // 16) * ((a1 * a2 + 15) // 16)
                tot += sysm.ns * tiles * inst * 2 * 16 * 16 * 64
    return tot


def cpu_baseline(sysm, G0, P0, P, Sigma, max_workers=16):
    """The reference's CPU path restated (the oracle), timed on the host on a
    sample of the same iteration:
      * P0 and Sigma: one output slice each (iBZ point 0, tau 0), every block
        product a zgemm of numpy's OpenBLAS (oracle/blas.py); on 1 thread over
        the first nk/8 k (q) points of the slice's sum, and over the whole sum
        split across `workers` single-thread processes (the host's cores, at
        most max_workers; time = the slowest process's);
      * W ((1 - P0)^-1 P0 per bosonic frequency, with the tau <-> i Omega
        transforms), Dyson and the rotation of G to the full zone: whole, in
        the oracle's C loops, one thread.
    A GW iteration is then flop_P0 / rate_P0 + flop_Sigma / rate_Sigma + W +
    Dyson + rotation.  The whole-sum slices are also compared with the device's
    P0 and Sigma of the same iteration (the full-size parity check)."""
    from mbsymmetry_amd import _lib
    from oracle import blas as B
    from oracle.oracle import Oracle
    nproc = os.cpu_count()
    workers = max(1, min(len(os.sched_getaffinity(0)), max_workers))
    orc = Oracle(sysm)
    T = B.Tables(sysm)
    nk, nts = sysm.nk, sysm.nts
    q = int(sysm.irre_list[0])
    k = q
    p0_pairs = [kk * nk + int(sysm.kpq[kk, q]) for kk in range(nk)]
    sg_pairs = [k * nk + int(sysm.kmq[k, qq]) for qq in range(nk)]
    V, off = sysm.host_V_pairs(p0_pairs + sg_pairs)
    log(f"cpu baseline: host V of {len(set(p0_pairs + sg_pairs))} k-pairs ready ({V.nbytes / 2**30:.2f} GiB)")
    t0 = time.perf_counter()
    Gf = orc.rotate_G(G0)
    t_rot = time.perf_counter() - t0
    Pt = orc.rotate_P_t(P, 0)
    sub = list(range(max(1, nk // 8)))
    res = {}
    for nthr, sel in ((1, sub), (workers, None)):
        tm = {}
        a = B.p0_at(T, V, off, Gf, 0, 0, ks=sel, workers=nthr, timing=tm)
        tp = tm["seconds"]
        fp = B.flops_p0_at(T, 0, sel)
        log(f"cpu baseline: P0 slice ({len(sel) if sel else nk} k) on {nthr} threads: {tp:.2f} s, "
            f"{fp / tp / 1e9:.2f} GFLOP/s")
        b = B.sigma_at(T, V, off, Gf, Pt, 0, 0, 0, qs_sel=sel, workers=nthr, timing=tm)
        ts = tm["seconds"]
        fs = B.flops_sigma_at(T, 0, sel)
        log(f"cpu baseline: Sigma slice ({len(sel) if sel else nk} q) on {nthr} threads: {ts:.2f} s, "
            f"{fs / ts / 1e9:.2f} GFLOP/s")
        res[nthr] = dict(rate_p0=fp / tp, rate_sig=fs / ts, seconds=tp + ts, p0=a, sig=b)
    t0 = time.perf_counter()
    orc.W(P0)
    t_w = time.perf_counter() - t0
    t0 = time.perf_counter()
    orc.dyson(Sigma, 0.0)
    t_d = time.perf_counter() - t0
    log(f"cpu baseline: W {t_w:.2f} s, Dyson {t_d:.2f} s, G rotation {t_rot:.2f} s (1 thread)")
    info = _lib.layout_query(sysm)
    full = res[workers]
    qs = sysm.qslice(q)
    ks = sysm.kslice(k)
    d_p0 = float(np.abs(full["p0"] - P0[0:qs]).max())
    d_sig = float(np.abs(full["sig"] - Sigma[0, 0, 0:ks]).max())
    out = {}
    for nthr, r in res.items():
        it = info.flop_P0 / r["rate_p0"] + info.flop_Sigma / r["rate_sig"] + t_w + t_d + t_rot
        out[nthr] = dict(iter_s=it, gflops_p0=r["rate_p0"] / 1e9, gflops_sigma=r["rate_sig"] / 1e9,
                         timed_s=r["seconds"])
    return dict(by_threads=out, workers=workers, nproc=nproc, W_s=t_w, dyson_s=t_d, rot_s=t_rot,
                gpu_vs_cpu_max_abs_dP0=d_p0, gpu_vs_cpu_max_abs_dSigma=d_sig,
                max_abs_P0=float(np.abs(full["p0"]).max()), max_abs_Sigma=float(np.abs(full["sig"]).max()),
                sample_k=len(sub))


def config3_run(tune=True, with_fp32=False):
    """The uneven-batching system (64 irreducible points, each with its own
    2..10 irreps of sizes 1..256; the whole sweep of each contraction in one
    launch per slab depth): P0 and Sigma kernel time and rate from HIP events
    after one untimed pass, in FP64 and the mixed mode (and FP32 with
    with_fp32), with max|dSigma| of the others against FP64.  The autotuner
    picks the plan of the FP64 context; the other modes run it."""
    from mbsymmetry_amd import _lib, system as S
    sysm = S.config3()
    nb = [len(b) for b in sysm.ao_sizes]
    out = {"workload": (f"config3: 4x4x4 mesh, all 64 points irreducible, 2..10 irreps per point "
                        f"(ao blocks {min(min(b) for b in sysm.ao_sizes)}..{max(max(b) for b in sysm.ao_sizes)}, "
                        f"mean {np.mean(nb):.1f} per point), nts {sysm.nts}")}
    sig = {}
    modes = [(_lib.PREC_FP64, "fp64"), (_lib.PREC_MIXED, "mixed")] + ([(_lib.PREC_FP32, "fp32")] if with_fp32 else [])
    plan = None
    for prec, name in modes:
        ctx = setup_ctx(sysm, prec, 0.0)
        ctx.upload_G(sysm.zeros_ts_kij())
        ctx.dyson(0.0, 0.0)
        G0 = ctx.download_G((sysm.nts, sysm.ns, sysm.flat_ao))
        tuned = None
        if plan is not None:
            ctx.replan(**plan)
            tuned = {"plan": ctx.plan(), "from": "fp64"}
        elif tune:
            from mbsymmetry_amd import autotune
            ctx.upload_G(G0)
            rec = autotune.tune(ctx, sysm, prec, sample_s=0.3, use_cache=False, log=log)
            tuned = {k: rec[k] for k in ("plan", "Sigma", "P0", "seconds")}
            plan = rec["plan"]
        for timed in (False, True):
            ctx.upload_G(G0)
            ctx.set_timing(timed)
            ctx.solve_P0(); ctx.solve_W(); ctx.solve_Sigma()
            ctx.sync()
        lt, inf = ctx.last_times(), ctx.info()
        sig[name] = ctx.copy_Sigma_to_host(sysm.zeros_ts_kij()) / sysm.nk
        ctx.close()
        out[name] = {"P0_ms": float(lt[0]), "Sigma_ms": float(lt[2]), "P0_tflops": float(inf.flop_P0 / lt[0] / 1e9),
                     "Sigma_tflops": float(inf.flop_Sigma / lt[2] / 1e9), "autotune": tuned}
        if name != "fp64":
            out[name]["max_abs_dSigma_vs_fp64"] = float(np.abs(sig[name] - sig["fp64"]).max())
        log("config3", name, out[name])
    out["max_abs_Sigma"] = float(np.abs(sig["fp64"]).max())
    out["flop_P0"], out["flop_Sigma"] = float(inf.flop_P0), float(inf.flop_Sigma)
    return out


def config4_run(shard_points=1, nodes=2):
    """The precision-scaling system (8x8x8 mesh reduced to 29 irreducible
    points; nao 256 in 6 blocks, naux 1200 in 6 blocks; beta 200) on a
    sample, with the three GEMM paths: FP64, FP32 (3M complex products) and
    mixed (INT8 Z product).  Its V (~5 TB in FP64) is generated on the device
    per batch of k-pairs inside the sweeps (SYMGW_FLAG_STREAM_V), so the phase
    times include the generation.  The sample: P0 and Sigma over the first
    `shard_points` irreducible points (symgw_set_shard) on a `nodes`-node
    Chebyshev grid; per time the work is that of the 144-point grid, whose
    full-zone P (287 GB) does not fit one GPU.  P for Sigma comes from the FP64
    context's W solve (P0 zero outside the shard) and is uploaded to every
    path, so the three Sigma sweeps see the same inputs; max|dSigma| of FP32
    and mixed is against FP64 on the shard's rows."""
    from mbsymmetry_amd import _lib, system as S
    from mbsymmetry_amd.transforms import chebyshev
    t0 = time.perf_counter()
    sysm = S.config4(transforms=chebyshev(nodes, 200.0))
    log(f"config4: system built in {time.perf_counter() - t0:.1f} s")
    nk, ink = sysm.nk, sysm.ink
    out = {"workload": (f"config4: 8x8x8 k-mesh ({ink} iBZ points), nao 256 in blocks {sysm.ao_sizes[0]}, naux 1200 in "
                        f"blocks {sysm.aux_sizes[0]}, V streamed (generated on the device per batch of k-pairs)"),
           "sample": (f"P0 and Sigma over iBZ points 0..{shard_points - 1} of {ink} on a {nodes}-node Chebyshev grid "
                      f"(nts {sysm.nts}), beta 200; phase times include V generation")}
    ko = sysm.kao_off_irre() + [sysm.flat_ao]
    rows = slice(ko[0], ko[shard_points])
    sig = {}
    G0 = P = None
    info = None
    for name, prec in (("fp64", _lib.PREC_FP64), ("fp32", _lib.PREC_FP32), ("mixed", _lib.PREC_MIXED)):
        ctx = _lib.Context(sysm, 0, prec | _lib.FLAG_STREAM_V)
        ctx.generate_V(sysm.v_seed, sysm.v_scale, sysm.decay_table())
        ctx.set_reps(sysm.orep, sysm.auxrep)
        ctx.set_transforms(sysm.transforms)
        ctx.set_fock(sysm.F, sysm.S)
        if G0 is None:
            ctx.upload_G(sysm.zeros_ts_kij())
            ctx.dyson(0.0, 0.0)
            G0 = ctx.download_G((sysm.nts, sysm.ns, sysm.flat_ao))
        ctx.set_shard(0, shard_points)
        ctx.upload_G(G0)
        ctx.upload_P(sysm.zeros_q_t_QP())
        ctx.set_timing(True)
        log(f"config4 {name}: context ready, P0 sweep of the sample")
        ctx.solve_P0()
        ctx.sync()
        p0_ms = float(ctx.last_times()[0])
        log(f"config4 {name}: P0 {p0_ms:.1f} ms")
        if P is None:
            ctx.solve_W()
            P = ctx.copy_P0_to_host(sysm.zeros_q_t_QP())
        ctx.upload_P(P)
        ctx.solve_Sigma()
        ctx.sync()
        sig_ms = float(ctx.last_times()[2])
        log(f"config4 {name}: Sigma {sig_ms:.1f} ms")
        sig[name] = ctx.copy_Sigma_to_host(sysm.zeros_ts_kij())[:, :, rows] / nk
        info = ctx.info()
        ctx.close()
        f_p0 = info.flop_P0 * shard_points / ink
        f_sig = info.flop_Sigma * shard_points / ink
        out[name] = {"P0_ms": p0_ms, "Sigma_ms": sig_ms, "P0_tflops": f_p0 / (p0_ms * 1e-3) / 1e12,
                     "Sigma_tflops": f_sig / (sig_ms * 1e-3) / 1e12}
        if name != "fp64":
            out[name]["max_abs_dSigma_vs_fp64"] = float(np.abs(sig[name] - sig["fp64"]).max())
        log("config4", name, out[name])
    out["max_abs_Sigma"] = float(np.abs(sig["fp64"]).max())
    out["sample_flop_P0"] = info.flop_P0 * shard_points / ink
    out["sample_flop_Sigma"] = info.flop_Sigma * shard_points / ink
    # one whole iteration on the 144-point grid at the sample's rates (P0 and
    # Sigma scale with the number of times: nts/2 and nts)
    scale = 146.0 / sysm.nts
    out["full_iteration_estimate_s"] = {
        n: (info.flop_P0 * scale / (out[n]["P0_tflops"] * 1e12) + info.flop_Sigma * scale / (out[n]["Sigma_tflops"] * 1e12))
        for n in ("fp64", "fp32", "mixed")}
    return out


def serving_run(window_s, rate, deadline_s, precisions=("fp64", "mixed"), n_systems=32):
    """The serving loop on the 32-system mix (nao in {26, 52, 104, 156, 256},
    meshes 2^3 / 4^3 / 6^3): systems resident on the device (loaded before the
    clock; V streamed for the largest), GW-iteration requests for uniformly
    drawn systems arriving as a Poisson stream for `window_s` seconds, admitted
    by HBM footprint, continuously batched over up to 8 concurrent systems;
    no request starts that would end after `deadline_s` by the FLOP estimate.
    Latency = arrival -> new G on the device."""
    from mbsymmetry_amd import _lib, serving as SV
    from mbsymmetry_amd import system as S
    t0 = time.perf_counter()
    mix = S.config5(n_systems)
    log(f"serving: {len(mix)} systems built in {time.perf_counter() - t0:.1f} s")
    rng = np.random.default_rng(S.SEED_BASE + 55)
    arr = SV.poisson_arrivals(int(rate * window_s * 3) + 16, rate, seed=S.SEED_BASE + 5)
    arr = arr[arr <= window_s]
    reqs = [SV.IterRequest(i, int(rng.integers(0, len(mix))), float(a)) for i, a in enumerate(arr)]
    sizes = {}
    for s_ in mix:
        sizes[SV._key(s_)] = sizes.get(SV._key(s_), 0) + 1
    out = {"systems": len(mix), "system_sizes": dict(sorted(sizes.items())), "requests": len(reqs),
           "arrival_rate_per_s": rate, "arrival_window_s": window_s, "deadline_s": deadline_s,
           "mix": "config5: nao uniform in {26,52,104,156,256} (naux = 4 nao, <= 8 irreps), mesh uniform in "
                  "{2^3,4^3,6^3}, IR grid Lambda=1e3 (nts 50), beta 50; one GW iteration per request"}
    modes = {"fp64": _lib.PREC_FP64, "mixed": _lib.PREC_MIXED, "fp32": _lib.PREC_FP32}
    for name in precisions:
        r = SV.serve_resident(mix, reqs, device=0, precision=modes[name], max_active=8, time_limit=window_s,
                              deadline=deadline_s)
        sm = r["summary"]
        sm.update(peak_active=r["peak_active"], peak_resident_GiB=r["peak_bytes"] / 2 ** 30,
                  budget_GiB=r["budget"] / 2 ** 30, setup_s=r["setup_s"], loads=r["loads"],
                  evictions=r["evictions"], systems_with_streamed_V=r["streamed"])
        out[name] = sm
        log(f"serving {name}: {sm['completed']}/{sm['requests']} done, p50 {sm['p50_latency_s']:.3f} s, "
            f"p99 {sm['p99_latency_s']:.3f} s, {sm['gw_iter_per_s']:.2f} iter/s, "
            f"{sm['P0_Sigma_tflops']:.1f} TFLOP/s, unserved {sm['unserved']}, failed {sm['failed']}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="config2")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--fp32", action="store_true", help="also time the all-FP32 contraction path")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mu", type=float, default=0.0)
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-config3", action="store_true", help="skip the uneven-block (config 3) leg")
    ap.add_argument("--no-config4", action="store_true", help="skip the sampled precision-scaling (config 4) leg")
    ap.add_argument("--no-tune", action="store_true", help="keep the built-in work-list plan (no autotuner)")
    ap.add_argument("--plan", default=None,
                    help="fixed work-list plan for the config-2 legs instead of the autotuner, as JSON "
                         "(e.g. '{\"sig_max_times\": 2}'; the rules not named keep their defaults)")
    ap.add_argument("--serving-only", action="store_true", help="run only the serving mix and print its JSON")
    ap.add_argument("--serving-window", type=float, default=2.0, help="seconds of Poisson arrivals")
    ap.add_argument("--serving-deadline", type=float, default=20.0,
                    help="no serving request starts that is estimated to end later (s)")
    ap.add_argument("--serving-rate", type=float, default=50.0, help="Poisson arrival rate (requests/s)")
    ap.add_argument("--serving-modes", default="fp64,mixed", help="precision modes of the serving leg")
    ap.add_argument("--copy-calib", action="store_true",
                    help="run only a 4 GiB device-to-device copy kernel (scale reference for FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    # one hardware queue per serving slot (8), so that a short request's kernels
    # never queue behind a long request's on a shared queue; set before HIP starts
    os.environ["GPU_MAX_HW_QUEUES"] = "8"
    if args.copy_calib:
        import torch
        n = 4 << 30
        a = torch.empty(n // 8, dtype=torch.float64, device="cuda").fill_(1.0)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        torch.mul(a, 1.0, out=b)  # one elementwise kernel: reads 4 GiB, writes 4 GiB
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n, "copy_bytes_written": n}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 (the k-pair sharding "
                                   "over GPUs is a design, not part of this build)"}))
        sys.exit(2)
    from mbsymmetry_amd import _lib

    if args.serving_only:
        print(json.dumps({"serving": serving_run(args.serving_window, args.serving_rate, args.serving_deadline,
                                                 args.serving_modes.split(","))}),
              flush=True)
        return
    fixed_plan = json.loads(args.plan) if args.plan else None
    sysm = make_system(args.config)
    log("system", workload_name(sysm))
    peaks = _lib.measure_peaks(0)  # the box's own peaks, beside the datasheet values
    log("measured peaks", peaks)
    # G0 = Dyson with Sigma = 0 on the device (sc_loop init_g)
    ctx = setup_ctx(sysm, _lib.PREC_FP64, args.mu)
    ctx.upload_G(sysm.zeros_ts_kij())
    ctx.dyson(args.mu, 0.0)
    G0 = ctx.download_G((sysm.nts, sysm.ns, sysm.flat_ao))
    ctx.close()
    log("G0 ready")

    plain = run_path(sysm, _lib.PREC_FP64, args.steps, args.warmup, args.mu, G0, keep_P=not args.no_cpu,
                     tune=not args.no_tune, plan=fixed_plan)
    log("plain fp64", plain["sec_per_iter"], "s/iter; phases ms", plain["phase_ms"][:4])
    mixed = fp32 = None
    if not args.no_mixed:
        # the mixed mode runs the plan the autotuner picked for the plain path
        mixed = run_path(sysm, _lib.PREC_MIXED, args.steps, args.warmup, args.mu, G0, tune=False,
                         plan=fixed_plan if fixed_plan is not None else (plain["tuned"] or {}).get("plan"))
        log("mixed (int8 Z)", mixed["sec_per_iter"], "s/iter; phases ms", mixed["phase_ms"][:4])
    if args.fp32:
        fp32 = run_path(sysm, _lib.PREC_FP32, args.steps, args.warmup, args.mu, G0, tune=not args.no_tune,
                        plan=fixed_plan)
        log("fp32", fp32["sec_per_iter"], "s/iter; phases ms", fp32["phase_ms"][:4])

    p0_ms, sig_ms = float(plain["phase_ms"][0]), float(plain["phase_ms"][2])
    dom = "Sigma" if sig_ms >= p0_ms else "P0"
    dom_ms = max(sig_ms, p0_ms)
    dom_flop = plain["flop_sig"] if dom == "Sigma" else plain["flop_p0"]
    achieved = dom_flop / (dom_ms * 1e-3) / 1e12
    flops_iter = plain["flop_p0"] + plain["flop_sig"]
    if dom == "Sigma":
        # Sigma runs one launch per slab depth (4 times per slab for ao blocks up
        # to 32 orbitals, 2 or 1 for wider ones); the phase time covers them all
        # every block of at most 32 orbitals takes the deepest slab the plan allows
        widest = max(b for s in sysm.ao_sizes for b in s)
        plan = (plain["tuned"] or {}).get("plan") or _lib.plan_defaults().as_dict()
        tb = min(4, int(plan["sig_max_times"]))
        dom_kernel = f"sigma_mfma_kernel<double, {tb}, false>" if widest <= 32 else "sigma_mfma_kernel<double, 1|2|4, false>"
    else:
        dom_kernel = "p0_mfma_kernel<double>"
    traffic = None
    traffic_source = "no PMC record for this kernel and plan in profiles/"
    tpath = os.path.join(ROOT, "profiles", f"traffic_{dom.lower()}_{args.config}_fp64.json")
    if os.path.exists(tpath):  # HBM bytes per launch from the FETCH_SIZE / WRITE_SIZE runs
        with open(tpath) as f:
            tr = json.load(f)
        if tr.get("kernel") == dom_kernel and tr.get("workload") == args.config:
            traffic = tr["bytes_per_launch"]
            traffic_source = (f"profiles/{os.path.basename(tpath)}: rocprofv3 FETCH_SIZE / WRITE_SIZE passes of "
                              f"{tr.get('command', 'bench.py')} at commit {tr.get('commit', '?')}")
    out = {
        "metric": "GW iter/s & P0+Sigma TFLOP/s vs plain FP64 GPU at max|dSigma|<=1e-8 Ha; p99 latency",
        "value": 1.0 / plain["sec_per_iter"],
        "unit": "GW iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": plain["sec_per_iter"] * 1e3,
        "ms_per_step_median": float(np.median(plain["step_s"])) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex128 (fp64)",
        "data": "synthetic (seeded device-generated V, Haar-random rotations and Fock blocks)",
        "config": {"workload": workload_name(sysm), "flop_P0": plain["flop_p0"], "flop_Sigma": plain["flop_sig"]},
        "p0_sigma_tflops": flops_iter / (1e-3 * (p0_ms + sig_ms)) / 1e12,
        "phase_ms": {"P0_kernel": p0_ms, "W+P_rotation": float(plain["phase_ms"][1]), "Sigma_kernel": sig_ms,
                     "Dyson+G_rotation": float(plain["phase_ms"][3])},
        "roofline": {"bound": "mfma", "kernel": dom_kernel, "achieved": achieved, "peak": PEAK["fp64"],
                     "unit": "TFLOP/s", "frac": achieved / PEAK["fp64"], "traffic": traffic, "traffic_source": traffic_source,
                     "measured_peak": peaks["fp64_mfma_tflops"],
                     "frac_of_measured_peak": achieved / peaks["fp64_mfma_tflops"]},
        "peaks": {"datasheet": {"fp64_mfma_tflops": PEAK["fp64"], "fp32_mfma_tflops": PEAK["fp32"],
                                "hbm_gbps": PEAK["hbm_gbps"]},
                  "measured": peaks},
        "kernel_tflops": {"P0": plain["flop_p0"] / (p0_ms * 1e-3) / 1e12,
                          "Sigma": plain["flop_sig"] / (sig_ms * 1e-3) / 1e12},
    }
    out["autotune"] = {"fp64": plain["tuned"], "mixed": mixed["tuned"] if mixed else None}
    out["physics"] = {"max_abs_Sigma": float(np.abs(plain["sigma"]).max()),
                      "max_abs_G_after_iteration": float(np.abs(plain["G1"]).max()),
                      "causal_bound_|G|<=1": bool(np.abs(plain["G1"]).max() <= 1.0 + 1e-9),
                      "dyson_leakage": plain["leakage"]}

    def compare(res, name):
        dsig = float(np.abs(res["sigma"] - plain["sigma"]).max())
        dg = float(np.abs(res["G1"] - plain["G1"]).max())
        rp0, rsig = float(res["phase_ms"][0]), float(res["phase_ms"][2])
        kt = {"P0": res["flop_p0"] / (rp0 * 1e-3) / 1e12, "Sigma": res["flop_sig"] / (rsig * 1e-3) / 1e12}
        return {"mode": name, "iter_per_s": 1.0 / res["sec_per_iter"],
                "speedup_vs_plain": plain["sec_per_iter"] / res["sec_per_iter"],
                "max_abs_dSigma_vs_plain": dsig, "max_abs_dG_vs_plain": dg, "meets_1e-8": dsig <= 1e-8 and dg <= 1e-8,
                "phase_ms": list(map(float, res["phase_ms"][:4])), "kernel_tflops": kt,
                "p0_sigma_tflops": flops_iter / (1e-3 * (rp0 + rsig)) / 1e12}

    if mixed is not None:
        out["mixed"] = compare(mixed, "Sigma's q x q x k1k2 product (67% of its FLOPs) on INT8 matrix cores: 5 "
                                      "slices per operand, exact int32 sums, FP64 reconstruction; P0, the other "
                                      "products, W, Dyson and the transforms FP64")
        out["mixed"]["int8_ops_Sigma"] = mixed["int8_ops"]
        out["mixed"]["int8_rate_lower_bound_TOPS"] = mixed["int8_ops"] / (mixed["phase_ms"][2] * 1e-3) / 1e12
        out["mixed"]["frac_int8_peak_lower_bound"] = out["mixed"]["int8_rate_lower_bound_TOPS"] / PEAK["int8"]
    if fp32 is not None:
        out["fp32"] = compare(fp32, "every contraction product in FP32 (3M), FP32 accumulation")
        out["fp32"]["frac_fp32_peak"] = {k: v / PEAK["fp32"] for k, v in out["fp32"]["kernel_tflops"].items()}
    if not args.no_config3:
        out["config3"] = config3_run(tune=not args.no_tune)
    if not args.no_config4:
        out["config4"] = config4_run()
    if not args.no_serving:
        out["serving"] = serving_run(args.serving_window, args.serving_rate, args.serving_deadline,
                                     args.serving_modes.split(","))
    if not args.no_cpu:
        cb = cpu_baseline(sysm, G0, plain["P0"], plain["P"], plain["sigma"])
        w, one = cb["workers"], cb["by_threads"][1]
        many = cb["by_threads"][w]
        out["cpu_baseline"] = {
            "value": 1.0 / many["iter_s"], "unit": "GW iterations/s", "cores": w, "kind": "port",
            "nproc": cb["nproc"], "value_1_thread": 1.0 / one["iter_s"],
            "sample": (f"the oracle (gw_block.cpp loop nest restated, block products on numpy's OpenBLAS zgemm) on "
                       f"one P0 slice (q_ir 0, tau 0) and one Sigma slice (k_ir 0, tau 0): over the whole "
                       f"k / q sum on {w} threads ({many['timed_s']:.1f} s; {many['gflops_p0']:.1f} / "
                       f"{many['gflops_sigma']:.1f} GFLOP/s), over {cb['sample_k']} of {sysm.nk} points on 1 thread "
                       f"({one['timed_s']:.1f} s; {one['gflops_p0']:.2f} / {one['gflops_sigma']:.2f} GFLOP/s); "
                       f"W {cb['W_s']:.2f} s, Dyson {cb['dyson_s']:.2f} s and G rotation {cb['rot_s']:.2f} s whole "
                       f"(C loops, 1 thread); an iteration = flop_P0 / rate_P0 + flop_Sigma / rate_Sigma + "
                       f"W + Dyson + rotation"),
            "iteration_s": {"1_thread": one["iter_s"], f"{w}_threads": many["iter_s"]},
            "gpu_vs_cpu_full_size_slices": {"max_abs_dP0": cb["gpu_vs_cpu_max_abs_dP0"],
                                            "max_abs_P0": cb["max_abs_P0"],
                                            "max_abs_dSigma": cb["gpu_vs_cpu_max_abs_dSigma"],
                                            "max_abs_Sigma": cb["max_abs_Sigma"]}}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
