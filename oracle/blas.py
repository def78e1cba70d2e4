"""CPU timing twin of the oracle's sampled P0 / Sigma slices, on BLAS.

Test infrastructure, like the rest of oracle/: only the tests and the CPU
baseline leg of the benchmark import it.  Same loop nest and the same three
products per block as orc_P0_at / orc_Sigma_at in oracle.c (which restate
SymmetryMBPT/src/cpu_ref/gw_block.cpp:6-84 for P0 and :86-158 for Sigma),
with every block product a numpy matmul, i.e. a zgemm of numpy's bundled
OpenBLAS (the host has no other BLAS), so that the CPU baseline is timed on
tuned GEMMs as the reference's Eigen products are, not on the oracle's
reference loops.

The loop nest is first flattened into task rows, one per (k point, spin,
block) for P0 and per (q point, block) for Sigma, holding the offsets of the
block's operands.  The rows run in this process (OpenBLAS held to one thread)
or are split over worker processes (`python -m oracle.blas`), each on one
thread, like the reference's ranks; the parent sums the partial slices.
Equality with the C oracle is tested (tests/test_oracle_blas.py)."""
from __future__ import annotations

import json
import os
import shutil
import subprocess
import sys
import tempfile
import time
from typing import Dict, List, Optional, Sequence

import numpy as np


def _offsets(sizes_sq):
    return np.concatenate([[0], np.cumsum(sizes_sq)[:-1]]).astype(np.int64)


class Tables:
    """Block offsets of a system in the full Brillouin zone (the layouts of
    G_full, ts_kij over every k, and P_full_t, Q x P over every q)."""

    def __init__(self, s):
        self.s = s
        nk = s.nk
        self.ao_off = [_offsets([b * b for b in s.ao_sizes[k]]) for k in range(nk)]
        self.aux_off = [_offsets([b * b for b in s.aux_sizes[q]]) for q in range(nk)]
        self.kao_full = _offsets([s.kslice(k) for k in range(nk)])
        self.qaux_full = _offsets([s.qslice(q) for q in range(nk)])
        self.flat_ao_full = s.flat_ao_full
        self.blocks: Dict[int, list] = {}

    def pair_blocks(self, p):
        """[(iq, i1, i2, qs, a1, a2, offset in the pair's V)] of pair p."""
        if p not in self.blocks:
            s = self.s
            k1, k2 = divmod(p, s.nk)
            q = s.pair_q(k1, k2)
            out, o = [], 0
            for iq, i1, i2 in s.V_irreps.get(p, []):
                qs, a1, a2 = s.aux_sizes[q][iq], s.ao_sizes[k1][i1], s.ao_sizes[k2][i2]
                out.append((iq, i1, i2, qs, a1, a2, o))
                o += qs * a1 * a2
            self.blocks[p] = out
        return self.blocks[p]


# ---------------------------------------------------------------------------
# task rows and their execution
# ---------------------------------------------------------------------------
# P0 row:    v, qs, a1, a2, g0 (G(beta - t) block), g1 (G(t) block), o (output)
# Sigma row: v, qs, a1, a2, g1 (G(t) block), po (P(t) block), o (output)
def _exec_p0(V, Gc, rows, n_out):
    acc = np.zeros(n_out, dtype=np.complex128)
    for v, qs, a1, a2, g0, g1, o in rows:
        K = a1 * a2
        Vb = V[v:v + qs * K].reshape(qs, a1, a2)
        Gtp = Gc[g0:g0 + a1 * a1].reshape(a1, a1)
        G1 = Gc[g1:g1 + a2 * a2].reshape(a2, a2)
        X1 = np.matmul(Gtp, Vb).reshape(qs, K)                         # G(beta - t) V
        Y = (Vb.reshape(qs * a1, a2).conj() @ G1.T).reshape(qs, K)     # conj(V) G(t)^T
        acc[o:o + qs * qs] += (X1 @ Y.T).ravel()                       # [Q][P] = sum_x X1[Q][x] Y[P][x]
    return acc


def _exec_sigma(V, Gc, Pt, rows, n_out):
    acc = np.zeros(n_out, dtype=np.complex128)
    for v, qs, a1, a2, g1, po, o in rows:
        K = a1 * a2
        Vb = V[v:v + qs * K].reshape(qs, a1, a2)
        G1 = Gc[g1:g1 + a2 * a2].reshape(a2, a2)
        Pq = Pt[po:po + qs * qs].reshape(qs, qs)
        Y1 = (Vb.reshape(qs * a1, a2) @ G1).reshape(qs, K)            # V G(t)
        Y2 = Y1.T @ Pq                                                  # (k1 k2) x q
        Wt = Vb.conj().transpose(2, 0, 1).reshape(a2 * qs, a1)          # [(n, P)][j] = conj(V[P][j][n])
        acc[o:o + a1 * a1] += (Y2.reshape(a1, a2 * qs) @ Wt).ravel()
    return acc


def p0_rows(T: Tables, off: Dict[int, int], iqr: int, t: int, ks: Optional[Sequence[int]] = None):
    """Task rows of the P0 slice (iqr, t), grouped per k point, and the
    slice length.  G offsets index [G(beta - t) over spins, G(t) over spins]."""
    s = T.s
    nk, ns, nts, F = s.nk, s.ns, s.nts, T.flat_ao_full
    q = int(s.irre_list[iqr])
    groups = []
    for k1 in (range(nk) if ks is None else ks):
        k2 = int(s.kpq[k1, q])
        p = int(k1) * nk + k2
        rows = []
        for sp in range(ns):
            for iq, i1, i2, qs, a1, a2, bo in T.pair_blocks(p):
                rows.append((off[p] + bo, qs, a1, a2, sp * F + T.kao_full[k1] + T.ao_off[k1][i1],
                             (ns + sp) * F + T.kao_full[k2] + T.ao_off[k2][i2], T.aux_off[q][iq]))
        groups.append(rows)
    return groups, s.qslice(q)


def sigma_rows(T: Tables, off: Dict[int, int], ikr: int, qs_sel: Optional[Sequence[int]] = None):
    """Task rows of the Sigma slice of iBZ k index ikr, grouped per q point,
    and the slice length.  G offsets index G(t) of one spin."""
    s = T.s
    nk = s.nk
    k1 = int(s.irre_list[ikr])
    groups = []
    for q in (range(nk) if qs_sel is None else qs_sel):
        p = k1 * nk + int(s.kmq[k1, q])
        rows = []
        for iq, i1, i2, qs, a1, a2, bo in T.pair_blocks(p):
            k2 = p % nk
            rows.append((off[p] + bo, qs, a1, a2, T.kao_full[k2] + T.ao_off[k2][i2],
                         T.qaux_full[q] + T.aux_off[q][iq], T.ao_off[k1][i1]))
        groups.append(rows)
    return groups, s.kslice(k1)


def _split(groups: List[list], n: int) -> List[np.ndarray]:
    n = max(1, min(n, len(groups)))
    parts = []
    for i in range(n):
        rows = [r for g in groups[len(groups) * i // n: len(groups) * (i + 1) // n] for r in g]
        parts.append(np.asarray(rows, dtype=np.int64).reshape(-1, 7))
    return parts


def _run(kind, arrays, parts, n_out):
    """Execute the row parts: in process for one part, else one worker
    process per part on one BLAS thread.  Returns (sum of the partial
    slices, compute seconds: the slowest part's)."""
    if len(parts) == 1:
        from threadpoolctl import threadpool_limits
        with threadpool_limits(1, user_api="blas"):
            t0 = time.perf_counter()
            if kind == "p0":
                acc = _exec_p0(arrays["V"], arrays["G"], parts[0], n_out)
            else:
                acc = _exec_sigma(arrays["V"], arrays["G"], arrays["P"], parts[0], n_out)
            return acc, time.perf_counter() - t0
    root = "/dev/shm" if os.path.isdir("/dev/shm") else None
    d = tempfile.mkdtemp(prefix="orc_blas_", dir=root)
    try:
        for name, a in arrays.items():
            np.save(os.path.join(d, name + ".npy"), a)
        for i, rows in enumerate(parts):
            np.save(os.path.join(d, f"rows{i}.npy"), rows)
        env = dict(os.environ, OPENBLAS_NUM_THREADS="1", OMP_NUM_THREADS="1", MKL_NUM_THREADS="1")
        top = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
        procs = [subprocess.Popen([sys.executable, "-m", "oracle.blas", d, kind, str(i), str(n_out)], env=env,
                                  cwd=top) for i in range(len(parts))]
        codes = [p.wait() for p in procs]
        if any(codes):
            raise RuntimeError(f"oracle BLAS workers failed: exit codes {codes}")
        acc = np.zeros(n_out, dtype=np.complex128)
        secs = []
        for i in range(len(parts)):
            acc += np.load(os.path.join(d, f"out{i}.npy"))
            with open(os.path.join(d, f"time{i}.json")) as f:
                secs.append(json.load(f)["seconds"])
        return acc, max(secs)
    finally:
        shutil.rmtree(d, ignore_errors=True)


def p0_at(T: Tables, V: np.ndarray, off: Dict[int, int], G_full: np.ndarray, iqr: int, t: int,
          ks: Optional[Sequence[int]] = None, workers: int = 1, timing: Optional[dict] = None) -> np.ndarray:
    """P0 of iBZ q index iqr at time t (the output of orc_P0_at), summed over
    the k points `ks` (default: all).  timing, if given, receives the compute
    seconds."""
    s = T.s
    G = G_full.reshape(s.nts, s.ns, T.flat_ao_full)
    Gc = np.concatenate([G[s.nts - 1 - t].ravel(), G[t].ravel()])
    groups, n_out = p0_rows(T, off, iqr, t, ks)
    acc, sec = _run("p0", {"V": V, "G": Gc}, _split(groups, workers), n_out)
    if timing is not None:
        timing["seconds"] = sec
    return acc * (-1.0 / s.nk * (2.0 if s.ns == 1 else 1.0))


def sigma_at(T: Tables, V: np.ndarray, off: Dict[int, int], G_full: np.ndarray, P_full_t: np.ndarray, ikr: int,
             t: int, spin: int = 0, qs_sel: Optional[Sequence[int]] = None, workers: int = 1,
             timing: Optional[dict] = None) -> np.ndarray:
    """Sigma of iBZ k index ikr at (t, spin) (the output of orc_Sigma_at),
    summed over the q points `qs_sel` (default: all)."""
    s = T.s
    Gc = np.ascontiguousarray(G_full.reshape(s.nts, s.ns, T.flat_ao_full)[t, spin])
    groups, n_out = sigma_rows(T, off, ikr, qs_sel)
    acc, sec = _run("sigma", {"V": V, "G": Gc, "P": np.ascontiguousarray(P_full_t)}, _split(groups, workers), n_out)
    if timing is not None:
        timing["seconds"] = sec
    return -acc / s.nk


def flops_p0_at(T: Tables, iqr: int, ks: Optional[Sequence[int]] = None) -> float:
    """Real FLOPs of p0_at (8 per complex multiply-add, three products per block)."""
    s = T.s
    q = int(s.irre_list[iqr])
    f = 0
    for k1 in (range(s.nk) if ks is None else ks):
        for _, _, _, qs, a1, a2, _ in T.pair_blocks(int(k1) * s.nk + int(s.kpq[k1, q])):
            f += qs * a1 * a1 * a2 + qs * a1 * a2 * a2 + qs * qs * a1 * a2
    return 8.0 * s.ns * f


def flops_sigma_at(T: Tables, ikr: int, qs_sel: Optional[Sequence[int]] = None) -> float:
    s = T.s
    k1 = int(s.irre_list[ikr])
    f = 0
    for q in (range(s.nk) if qs_sel is None else qs_sel):
        for _, _, _, qs, a1, a2, _ in T.pair_blocks(k1 * s.nk + int(s.kmq[k1, q])):
            f += qs * a1 * a2 * a2 + a1 * a2 * qs * qs + a1 * a2 * qs * a1
    return 8.0 * f


def _worker(d: str, kind: str, i: int, n_out: int) -> None:
    V = np.load(os.path.join(d, "V.npy"), mmap_mode="r")
    G = np.load(os.path.join(d, "G.npy"))
    rows = np.load(os.path.join(d, f"rows{i}.npy"))
    for v, qs, a1, a2 in rows[:, :4]:   # map this part's V pages before the clock starts
        float(np.abs(V[v:v + qs * a1 * a2]).max())
    t0 = time.perf_counter()
    if kind == "p0":
        acc = _exec_p0(V, G, rows, n_out)
    else:
        acc = _exec_sigma(V, G, np.load(os.path.join(d, "P.npy"), mmap_mode="r"), rows, n_out)
    sec = time.perf_counter() - t0
    np.save(os.path.join(d, f"out{i}.npy"), acc)
    with open(os.path.join(d, f"time{i}.json"), "w") as f:
        json.dump({"seconds": sec}, f)


if __name__ == "__main__":
    _worker(sys.argv[1], sys.argv[2], int(sys.argv[3]), int(sys.argv[4]))
