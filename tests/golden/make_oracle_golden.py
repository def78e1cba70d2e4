"""Regression vectors of the FP64 oracle: one GW iteration from the Sigma = 0
Green's function, on the LiH fixture and on the seeded 2x2x2 synthetic system
(nao 26, naux 128), written as fixed random projections of P0, P, Sigma and
the new G.

    python tests/golden/make_oracle_golden.py

The oracle is pinned independently (brute-force loops, closed forms, the
reference's block == dense relations); these numbers freeze its outputs so
that later changes to the oracle or to the systems cannot drift unnoticed.
Output: tests/golden/oracle_golden.json.
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

NPROJ = 6


def projections(x, seed):
    """Norm, max |x| and NPROJ complex projections on seeded Gaussian vectors."""
    x = np.asarray(x).ravel()
    rng = np.random.default_rng(seed)
    out = {"norm": float(np.linalg.norm(x)), "maxabs": float(np.abs(x).max()), "proj": []}
    for _ in range(NPROJ):
        w = rng.standard_normal(x.size) + 1j * rng.standard_normal(x.size)
        v = np.vdot(w, x)
        out["proj"].append([float(v.real), float(v.imag)])
    return out


def systems():
    from mbsymmetry_amd import system as S
    from mbsymmetry_amd.transforms import ir_from_fixture
    g = os.path.join(ROOT, "tests", "golden")
    lih = S.lih_fixture(np.load(os.path.join(g, "lih_nk2.npz")), ir_from_fixture(np.load(os.path.join(g, "ir_1e3_72.npz")), 5.0))
    return {"lih": lih, "config1": S.config1()}


def iterate(sysm, mu=0.0):
    from oracle.oracle import Oracle
    orc = Oracle(sysm)
    G0 = orc.dyson(sysm.zeros_ts_kij(), mu)
    P0, P, Sig, G1 = orc.gw_iteration(G0, mu, nthreads=os.cpu_count() or 1)
    return {"P0": P0, "P": P, "Sigma": Sig, "G1": G1}


def main():
    out = {"mu": 0.0, "nproj": NPROJ, "systems": {}}
    for name, sysm in systems().items():
        res = iterate(sysm)
        out["systems"][name] = {k: projections(v, seed=1000 + i) for i, (k, v) in enumerate(sorted(res.items()))}
        print(name, {k: v["norm"] for k, v in out["systems"][name].items()})
    with open(os.path.join(ROOT, "tests", "golden", "oracle_golden.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
