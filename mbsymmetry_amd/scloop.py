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
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np


@dataclass
class SCResult:
    G: np.ndarray                     # final G, ts_kij layout
    mu: float
    iterations: int
    converged: bool
    history: List[dict] = field(default_factory=list)   # per iteration: dG, mu, nel


def self_consistent(ctx, system, G0: np.ndarray, mu: float, max_iter: int = 20, tol: float = 1e-8,
                    damp: float = 1.0, nel: Optional[float] = None, mu_bracket=(-2.0, 2.0),
                    use_graph: bool = True) -> SCResult:
    """Iterate GW on `ctx` (V, rotations, transforms and Fock already set)
    from G0 until max|G_{n+1} - G_n| < tol or max_iter iterations."""
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
        G = G_new
        if dG < tol:
            converged = True
            break
    return SCResult(G=G, mu=mu, iterations=it, converged=converged, history=hist)
