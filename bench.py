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


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


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


def run_path(sysm, precision, steps, warmup, mu, G0):
    """Time `steps` GW iterations replayed from per-phase hipGraphs.

    The first iteration from G0 is untimed (it is warmup step 1) and gives the
    Sigma and G compared across precisions.  In the timed region HIP events sit
    between the phase graphs on the context stream, so the per-phase device
    times are averages over exactly the timed iterations."""
    ctx = setup_ctx(sysm, precision, mu)
    ctx.upload_G(G0)
    ctx.gw_iteration(mu, use_graph=True)
    ctx.sync()
    sig = ctx.copy_Sigma_to_host(sysm.zeros_ts_kij()) / sysm.nk
    G1 = ctx.download_G(G0.shape)
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
                flop_sig=info.flop_Sigma, sigma=sig, G1=G1, leakage=leakage(sysm.transforms, G1),
                int8_ops=int8_ops(sysm) if precision == _lib.PREC_MIXED else 0.0)


def int8_ops(sysm):
    """INT8 MFMA operations one mixed-mode Sigma sweep executes (executed, with
    the padding of 16-wide tiles and 64-deep slice-pair chunks): per Z job
    tile, 3 real products x sum_{d<5} ceil((d + 1) kp / 64) instructions of
    2 * 16 * 16 * 64 ops, for the blocks above the INT8 threshold."""
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
                inst = 3 * sum((d * kp + 63) // 64 for d in range(1, 6))
                # Z rows: times x aux rows; columns: a1 x a2, one pass per slab band of a1
                tiles = ((nts * qs + 15) // 16) * ((a1 * a2 + 15) // 16)
                tot += sysm.ns * tiles * inst * 2 * 16 * 16 * 64
    return tot


def cpu_baseline(sysm, G0, budget_s=15.0):
    """The FP64 CPU oracle (the reference's CPU loop nest restated) timed on a
    bounded sample of the same workload: P0 for min(cores, ink) iBZ q points
    (one per thread), the first nkk k points and the first ntau tau points,
    sized to about `budget_s` seconds; scaled to GW iterations/s with the FLOP
    model."""
    from oracle.oracle import Oracle
    cores = len(os.sched_getaffinity(0))
    orc = Oracle(sysm)
    nk = sysm.nk

    def flops(q, nkk):
        f = 0.0
        for k in range(nkk):
            p = k * nk + int(sysm.kpq[k, q])
            for (qs, a1, a2) in sysm.V_block_sizes(p):
                f += sysm.ns * 8.0 * (a2 * qs * a1 * a1 + qs * a1 * a2 * a2 + qs * qs * a1 * a2)
        return f

    nq = min(cores, sysm.ink)
    qs_ = [int(sysm.irre_list[i]) for i in range(nq)]
    # at most ~256 k-pairs of V are generated on the host (V of config 2 is 8 MiB per pair)
    nkk = min(nk, max(1, 256 // nq))
    pairs = {k * nk + int(sysm.kpq[k, q]) for q in qs_ for k in range(nkk)}
    V = sysm.host_V(pairs=pairs)
    Gf = orc.rotate_G(G0)
    log(f"cpu baseline: host V of {len(pairs)} k-pairs ready")
    # calibrate on one k of the first q (one thread), then size the sample
    t0 = time.perf_counter()
    orc.P0_sample(V, Gf, 0, 1, 1, 1, nthreads=1)
    rate1 = flops(qs_[0], 1) / max(time.perf_counter() - t0, 1e-6)
    per_tau = max(flops(q, nkk) for q in qs_) / rate1      # seconds for nkk k, one tau, per thread
    if per_tau >= budget_s:
        nkk, ntau = max(1, int(nkk * budget_s / per_tau)), 1
    else:
        ntau = int(min(sysm.nts // 2, max(1, budget_s // per_tau)))
    log(f"cpu baseline: timing {nq} q x {nkk} k x {ntau} tau on {nq} threads")
    t0 = time.perf_counter()
    orc.P0_sample(V, Gf, 0, ntau, nq, nkk, nthreads=nq)
    dt = time.perf_counter() - t0
    fl = ntau * sum(flops(q, nkk) for q in qs_)
    return dict(rate=fl / dt, seconds=dt, cores=nq, nkk=nkk, nq=nq, ntau=ntau, flops=fl)


def uneven_run():
    """Irrep blocks of every size from 1 to 256 in one system (the uneven-batching
    stress case at a size that fits one test): P0 and Sigma kernel rates from
    HIP events, after one untimed pass, in both precisions."""
    from mbsymmetry_amd import _lib, system as S
    from mbsymmetry_amd.transforms import chebyshev
    ao, aux = [1, 3, 9, 27, 81, 256], [1, 4, 16, 64, 256]
    sysm = S.synthetic("uneven-1-256", 2, 4, ao, aux, 1, chebyshev(20, 20.0), S.SEED_BASE + 3)
    out = {"workload": f"2x2x2 k-mesh (4 iBZ points), ao blocks {ao}, aux blocks {aux}, nts {sysm.nts}"}
    for prec, name in ((_lib.PREC_FP64, "fp64"), (_lib.PREC_FP32, "fp32")):
        ctx = setup_ctx(sysm, prec, 0.0)
        ctx.upload_G(sysm.zeros_ts_kij())
        ctx.dyson(0.0, 0.0)
        for timed in (False, True):
            ctx.set_timing(timed)
            ctx.solve_P0(); ctx.solve_W(); ctx.solve_Sigma()
            ctx.sync()
        lt, inf = ctx.last_times(), ctx.info()
        ctx.close()
        out[name] = {"P0_ms": float(lt[0]), "Sigma_ms": float(lt[2]), "P0_tflops": float(inf.flop_P0 / lt[0] / 1e9),
                     "Sigma_tflops": float(inf.flop_Sigma / lt[2] / 1e9)}
        log("uneven", name, out[name])
    return out


def _ints(text):
    return [int(x) for x in text.split(",") if x]


def serving_run(n_requests, rate, with_fp32, naos=(26, 52, 104), meshes=(2, 4)):
    """The serving loop on a seeded request mix: Poisson arrivals, HBM-footprint
    admission, continuous batching over per-system contexts; latency from
    arrival to the final G on the host."""
    from mbsymmetry_amd import _lib, serving as SV
    from mbsymmetry_amd import system as S
    mix = S.serving_mix(n_requests, meshes=tuple(meshes), naos=tuple(naos))
    arr = SV.poisson_arrivals(len(mix), rate, seed=S.SEED_BASE + 5)
    reqs = [SV.Request(i, s, float(arr[i]), iterations=it) for i, (s, it) in enumerate(mix)]
    flops = sum(it * (_lib.layout_query(s).flop_P0 + _lib.layout_query(s).flop_Sigma) for s, it in mix)
    out = {"requests": len(reqs), "arrival_rate_per_s": rate,
           "mix": f"synthetic systems on {' / '.join(f'{m}^3' for m in meshes)} meshes, nao in "
                  f"{{{', '.join(map(str, naos))}}} (naux = 4 nao), nts 32, 2 GW iterations each",
           "P0_Sigma_TFLOP": flops / 1e12}
    for name, prec in (("fp64", _lib.PREC_FP64), ("fp32", _lib.PREC_FP32)):
        if name == "fp32" and not with_fp32:
            continue
        r = SV.serve(reqs, device=0, precision=prec, max_active=8)
        sm = r["summary"]
        sm["peak_active"] = r["peak_active"]
        sm["peak_footprint_GiB"] = r["peak_bytes"] / 2 ** 30
        out[name] = sm
        log(f"serving {name}: p50 {sm['p50_latency_s']:.3f} s  p99 {sm['p99_latency_s']:.3f} s  "
            f"{sm['gw_iter_per_s']:.2f} iter/s")
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
    ap.add_argument("--no-uneven", action="store_true", help="skip the blocks-1-to-256 stress case")
    ap.add_argument("--serving-naos", default="26,52,104", help="nao values of the serving mix")
    ap.add_argument("--serving-meshes", default="2,4", help="k-mesh edges of the serving mix")
    ap.add_argument("--serving-only", action="store_true", help="run only the serving mix and print its JSON")
    ap.add_argument("--serving-requests", type=int, default=32)
    ap.add_argument("--serving-rate", type=float, default=50.0, help="Poisson arrival rate (requests/s)")
    ap.add_argument("--copy-calib", action="store_true",
                    help="run only a 4 GiB device-to-device copy kernel (scale reference for FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
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
        print(json.dumps({"serving": serving_run(args.serving_requests, args.serving_rate, not args.no_mixed,
                                                 _ints(args.serving_naos), _ints(args.serving_meshes))}),
              flush=True)
        return
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

    plain = run_path(sysm, _lib.PREC_FP64, args.steps, args.warmup, args.mu, G0)
    log("plain fp64", plain["sec_per_iter"], "s/iter; phases ms", plain["phase_ms"][:4])
    mixed = fp32 = None
    if not args.no_mixed:
        mixed = run_path(sysm, _lib.PREC_MIXED, args.steps, args.warmup, args.mu, G0)
        log("mixed (int8 Z)", mixed["sec_per_iter"], "s/iter; phases ms", mixed["phase_ms"][:4])
        if args.fp32:
            fp32 = run_path(sysm, _lib.PREC_FP32, args.steps, args.warmup, args.mu, G0)
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
        widest = max(b for s in sysm.ao_sizes for b in s)
        dom_kernel = "sigma_mfma_kernel<double, 4>" if widest <= 32 else "sigma_mfma_kernel<double, 1|2|4>"
    else:
        dom_kernel = "p0_mfma_kernel<double>"
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{dom.lower()}_{args.config}_fp64.json")
    if os.path.exists(tpath):  # HBM bytes per launch from the FETCH_SIZE / WRITE_SIZE runs
        with open(tpath) as f:
            tr = json.load(f)
        if tr.get("kernel") == dom_kernel and tr.get("workload") == args.config:
            traffic = tr["bytes_per_launch"]
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
                     "unit": "TFLOP/s", "frac": achieved / PEAK["fp64"], "traffic": traffic,
                     "measured_peak": peaks["fp64_mfma_tflops"],
                     "frac_of_measured_peak": achieved / peaks["fp64_mfma_tflops"]},
        "peaks": {"datasheet": {"fp64_mfma_tflops": PEAK["fp64"], "fp32_mfma_tflops": PEAK["fp32"],
                                "hbm_gbps": PEAK["hbm_gbps"]},
                  "measured": peaks},
        "kernel_tflops": {"P0": plain["flop_p0"] / (p0_ms * 1e-3) / 1e12,
                          "Sigma": plain["flop_sig"] / (sig_ms * 1e-3) / 1e12},
    }
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
    if not args.no_uneven:
        out["uneven_blocks"] = uneven_run()
    if not args.no_serving:
        out["serving"] = serving_run(args.serving_requests, args.serving_rate, not args.no_mixed,
                                     _ints(args.serving_naos), _ints(args.serving_meshes))
    if not args.no_cpu:
        cb = cpu_baseline(sysm, G0)
        out["cpu_baseline"] = {"value": cb["rate"] / flops_iter, "unit": "GW iterations/s", "cores": cb["cores"],
                               "kind": "port",
                               "sample": f"FP64 oracle P0 loop nest (gw_block.cpp:6-84 restated), {cb['nq']} iBZ q x "
                                         f"{cb['nkk']} k x {cb['ntau']} tau, {cb['flops'] / 1e9:.1f} GFLOP in {cb['seconds']:.1f} s "
                                         f"({cb['rate'] / 1e9:.2f} GFLOP/s), scaled by the FLOP model"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
