"""Self-consistency driver over the device GW iteration.

The reference stops at one solver pass plus a Dyson step per call
(SymmetryMBPT/src/scmbpt/sc_loop_gf_block.cpp:5-21, init_g with the density
matrix dmr = -(2/ns) G(beta)); its loop driver and convergence check are not
in the tree (SymmetryMBPT/CMakeLists.txt:59-63).  This driver iterates
P0 -> W -> Sigma -> Dyson on one context until the Green's function stops
changing:

  G_{n+1} = damp * GW(G_n) + (1 - damp) * G_n

with the reference's damping convention (1: no damping, 0: full damping,
sc_loop_block_t.h:70-72).  With `nel` given, mu is re-found after every
iteration by bisection over device Dyson solves so that the electron count
-(2/ns) sum_k w_k Tr[G_k(beta) S_k] / nk stays at nel.

`save_results` writes the outcome to one .npz archive: the run parameters
under the names the reference stores in its results file
(SymmetryMBPT/src/common/params_t.cpp:27-48, default "sim.h5"), the final G
and Sigma (ts_kij layout) and the per-iteration history.  HDF5 is not used
because no HDF5 library is available on the device hosts.
"""
from __future__ import annotations

import argparse
import json
import time
from dataclasses import dataclass, field
from typing import Callable, List, Optional

import numpy as np


@dataclass
class SCResult:
    G: np.ndarray                     # final G, ts_kij layout
    mu: float
    iterations: int
    converged: bool
    history: List[dict] = field(default_factory=list)   # per iteration: dG, mu, nel
    Sigma: Optional[np.ndarray] = None  # Sigma of the last iteration (/nk), ts_kij layout, if kept


def self_consistent(ctx, system, G0: np.ndarray, mu: float, max_iter: int = 20, tol: float = 1e-8,
                    damp: float = 1.0, nel: Optional[float] = None, mu_bracket=(-2.0, 2.0),
                    use_graph: bool = True, keep_sigma: bool = False,
                    stall_ratio: Optional[float] = None,
                    on_iteration: Optional[Callable[[dict], None]] = None) -> SCResult:
    """Iterate GW on `ctx` (V, rotations, transforms and Fock already set)
    from G0 until max|G_{n+1} - G_n| < tol or max_iter iterations, or, with
    stall_ratio, until an update shrinks by less than that factor (the loop
    has reached its precision's noise floor; converged stays False).  With
    keep_sigma the last iteration's Sigma is copied back, divided by nk as the
    reference's caller does (cugw_solver_block.cpp:302-304).  on_iteration
    receives each iteration's history entry as it is made."""
    if not 0.0 < damp <= 1.0:
        raise ValueError("damp must be in (0, 1]")
    shape = (system.nts, system.ns, system.flat_ao)
    G = np.ascontiguousarray(G0, dtype=np.complex128).reshape(shape)
    ctx.upload_G(G)
    hist = []
    converged = False
    it = 0
    for it in range(1, max_iter + 1):
        ctx.gw_iteration(mu, use_graph=use_graph)
        ctx.sync()
        n_now = None
        if nel is not None:
            mu, n_now = ctx.mu_search(nel, system.kweights, 1.0 / system.nk, mu_bracket[0], mu_bracket[1], 1e-12)
        G_new = ctx.download_G(shape)
        if damp < 1.0:
            G_new = damp * G_new + (1.0 - damp) * G
            ctx.upload_G(G_new)
        dG = float(np.abs(G_new - G).max())
        hist.append({"iteration": it, "max_abs_dG": dG, "mu": mu, "nel": n_now})
        if on_iteration is not None:
            on_iteration(hist[-1])
        G = G_new
        if dG < tol:
            converged = True
            break
        if stall_ratio is not None and len(hist) > 1 and dG > stall_ratio * hist[-2]["max_abs_dG"]:
            break
    Sigma = None
    if keep_sigma:
        Sigma = ctx.copy_Sigma_to_host(system.zeros_ts_kij()) / system.nk
    return SCResult(G=G, mu=mu, iterations=it, converged=converged, history=hist, Sigma=Sigma)


def self_consistent_cleanup(ctx_sp, ctx_dp, system, G0: np.ndarray, mu: float, max_iter: int = 20,
                            tol: float = 1e-8, sp_tol: float = 1e-6, damp: float = 1.0,
                            nel: Optional[float] = None, mu_bracket=(-2.0, 2.0), use_graph: bool = True,
                            keep_sigma: bool = False,
                            on_iteration: Optional[Callable[[dict], None]] = None) -> SCResult:
    """Single precision with a cleanup in double precision (the reference's
    single_prec = 2, SymmetryMBPT/src/common/params_t.cpp:76, switched per
    solver by solver_block_t::single_prec_run, cugw_solver_block.h:70-77):
    iterate on `ctx_sp` (complex64 contractions) until max|dG| < sp_tol or
    the updates stop shrinking by half (the FP32 noise floor), then continue
    from that G on `ctx_dp` (complex128) until max|dG| < tol.  Both contexts
    hold the same system.  History entries carry the precision of their
    iteration.  It saves time only where the loop needs many iterations: the
    FP64 phase still takes the iterations from the FP32 floor down to tol."""
    def tagged(precision, offset):
        def f(h):
            h["precision"] = precision
            h["iteration"] += offset
            if on_iteration is not None:
                on_iteration(h)
        return f

    r1 = self_consistent(ctx_sp, system, G0, mu, max_iter, sp_tol, damp, nel, mu_bracket, use_graph,
                         stall_ratio=0.5, on_iteration=tagged("fp32", 0))
    r2 = self_consistent(ctx_dp, system, r1.G, r1.mu, max_iter - r1.iterations, tol, damp, nel, mu_bracket,
                         use_graph, keep_sigma, on_iteration=tagged("fp64", r1.iterations))
    return SCResult(G=r2.G, mu=r2.mu, iterations=r1.iterations + r2.iterations, converged=r2.converged,
                    history=r1.history + r2.history, Sigma=r2.Sigma)


def save_results(path: str, res: SCResult, system, params: Optional[dict] = None) -> None:
    """Write `res` to `path` (.npz).  Keys: "params/<name>" for the system
    sizes and run options (nel_cell, nao, nk, ink, NQ, ns, nts, beta, mu, tol,
    damping, ... plus anything in `params`), "G_tau", "Selfenergy" (when kept),
    "iterations", "converged" and "history/<field>" arrays."""
    out = {
        "params/nao": np.int64(sum(system.ao_sizes[int(system.irre_list[0])])),
        "params/NQ": np.int64(sum(system.aux_sizes[int(system.irre_list[0])])),
        "params/nk": np.int64(system.nk),
        "params/ink": np.int64(len(system.irre_list)),
        "params/ns": np.int64(system.ns),
        "params/nts": np.int64(system.nts),
        "params/beta": np.float64(system.transforms.beta),
        "params/mu": np.float64(res.mu),
        "G_tau": res.G,
        "iterations": np.int64(res.iterations),
        "converged": np.bool_(res.converged),
    }
    for k, v in (params or {}).items():
        out["params/" + k] = np.asarray(v)
    if res.Sigma is not None:
        out["Selfenergy"] = res.Sigma
    for fld in ("iteration", "max_abs_dG", "mu", "nel"):
        out["history/" + fld] = np.array([np.nan if h[fld] is None else h[fld] for h in res.history], dtype=np.float64)
    np.savez(path, **out)


def load_results(path: str) -> dict:
    """Read an archive written by save_results (no pickled objects)."""
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def main(argv=None) -> int:
    """Self-consistent GW on a synthetic system from the Sigma = 0 Green's
    function at mu = 0, printing each iteration; with --compare, the same loop
    in a second precision from the same G0 and the largest difference between
    the two final G."""
    from . import _lib, system as S
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--system", default="config2", choices=["config1", "config2"])
    precisions = {"fp64": _lib.PREC_FP64, "mixed": _lib.PREC_MIXED, "fp32": _lib.PREC_FP32}
    ap.add_argument("--precision", default="fp64", choices=sorted(precisions))
    ap.add_argument("--compare", default=None, choices=sorted(precisions))
    ap.add_argument("--iters", type=int, default=10)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--damp", type=float, default=1.0)
    ap.add_argument("--out", default=None, help="save the first run's results (.npz)")
    ap.add_argument("--cleanup", action="store_true",
                    help="run --precision (fp32 or mixed) to its noise floor, then FP64 to --tol "
                         "(self_consistent_cleanup)")
    args = ap.parse_args(argv)
    sysm = S.config2() if args.system == "config2" else S.config1()

    def context(prec):
        ctx = _lib.Context(sysm, 0, prec)
        ctx.generate_V(sysm.v_seed, sysm.v_scale, sysm.decay_table())
        ctx.set_reps(sysm.orep, sysm.auxrep)
        ctx.set_transforms(sysm.transforms)
        ctx.set_fock(sysm.F, sysm.S)
        return ctx

    summary = {"system": args.system, "runs": {}}
    G0 = None
    finals = {}
    for name in [args.precision] + ([args.compare] if args.compare else []):
        ctx = context(precisions[name])
        if G0 is None:
            ctx.upload_G(sysm.zeros_ts_kij())
            ctx.dyson(0.0, 0.0)
            G0 = ctx.download_G((sysm.nts, sysm.ns, sysm.flat_ao))
        t0 = time.perf_counter()
        last = [t0]

        def report(h, name=name, last=last):
            now = time.perf_counter()
            print(f"[scloop] {name} iteration {h['iteration']} ({h.get('precision', name)}): "
                  f"max|dG| {h['max_abs_dG']:.3e}, {now - last[0]:.2f} s", flush=True)
            last[0] = now

        if args.cleanup and name == args.precision:
            ctx_dp = context(_lib.PREC_FP64)
            res = self_consistent_cleanup(ctx, ctx_dp, sysm, G0, 0.0, args.iters, args.tol, damp=args.damp,
                                          on_iteration=report)
            ctx_dp.close()
        else:
            res = self_consistent(ctx, sysm, G0, 0.0, args.iters, args.tol, args.damp, on_iteration=report)
        finals[name] = res.G
        summary["runs"][name] = {"iterations": res.iterations, "converged": res.converged,
                                 "seconds": time.perf_counter() - t0, "max_abs_G": float(np.abs(res.G).max()),
                                 "max_abs_dG": [h["max_abs_dG"] for h in res.history],
                                 "precisions": [h.get("precision", name) for h in res.history]}
        if args.out and name == args.precision:
            save_results(args.out, res, sysm, {"damping": args.damp, "tol": args.tol})
        ctx.close()
    if args.compare:
        summary["max_abs_G_difference"] = float(np.abs(finals[args.precision] - finals[args.compare]).max())
    print(json.dumps(summary), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())

