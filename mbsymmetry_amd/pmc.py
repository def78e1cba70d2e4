"""Summaries of rocprofv3 output for the contraction kernels.

The numbers behind the roofline figures come from three kinds of runs of the
same command (see `profiles/`):

  rocprofv3 --kernel-trace --stats --output-format csv -d DIR -o NAME -- CMD
  rocprofv3 --pmc FETCH_SIZE --output-format csv -d DIR -o NAME -- CMD
  rocprofv3 --pmc WRITE_SIZE ... (one pass per counter group)

plus the two PMC passes over `bench.py --copy-calib` (a 4 GiB elementwise copy
kernel of known size), which give the factor that turns FETCH_SIZE and
WRITE_SIZE into bytes.  This module reads those CSV files:

  python -m mbsymmetry_amd.pmc stats DIR
      per-kernel calls, average and total duration (the *_kernel_stats.csv)
  python -m mbsymmetry_amd.pmc counters DIR [DIR ...] [--kernel REGEX]
      per-kernel mean value per dispatch of every counter collected
  python -m mbsymmetry_amd.pmc traffic --fetch DIR --write DIR --calib-fetch DIR
      --calib-write DIR --kernel REGEX [--calib-bytes N] [--out FILE]
      HBM bytes per launch of the matching kernel: FETCH_SIZE and WRITE_SIZE
      (KiB per dispatch) scaled by what they read for the copy of N bytes
      read and N bytes written, as JSON (the bench line's roofline.traffic)
"""
from __future__ import annotations

import argparse
import collections
import csv
import glob
import json
import os
import re
import sys
from typing import Dict, List


def _rows(pattern_dir: str, suffix: str) -> List[dict]:
    files = sorted(glob.glob(os.path.join(pattern_dir, "**", f"*{suffix}"), recursive=True))
    if not files:
        raise SystemExit(f"no *{suffix} under {pattern_dir}")
    out = []
    for f in files:
        with open(f, newline="") as fh:
            out.extend(csv.DictReader(fh))
    return out


def _short(name: str) -> str:
    """Kernel name without its argument list (the parenthesised group that
    ends the name; "(anonymous namespace)" inside the name stays)."""
    name = name.strip()
    if not name.endswith(")"):
        return name
    depth = 0
    for i in range(len(name) - 1, -1, -1):
        if name[i] == ")":
            depth += 1
        elif name[i] == "(":
            depth -= 1
            if depth == 0:
                return name[:i].rstrip()
    return name


def kernel_stats(d: str) -> List[dict]:
    """Rows of the kernel-trace statistics: name, calls, average and total ms."""
    out = []
    for r in _rows(d, "kernel_stats.csv"):
        calls = int(r["Calls"])
        total = float(r["TotalDurationNs"])
        out.append({"kernel": _short(r["Name"]), "calls": calls, "avg_ms": total / calls / 1e6,
                    "total_ms": total / 1e6, "percent": float(r.get("Percentage", "nan"))})
    out.sort(key=lambda x: -x["total_ms"])
    return out


def counters(dirs: List[str], kernel: str = "") -> Dict[str, Dict[str, float]]:
    """Mean value per dispatch of each counter, per kernel (a dispatch's
    value is the sum over the rows rocprofv3 writes for it)."""
    rx = re.compile(kernel) if kernel else None
    per: Dict[str, Dict[str, Dict[str, float]]] = collections.defaultdict(lambda: collections.defaultdict(dict))
    for d in dirs:
        for r in _rows(d, "counter_collection.csv"):
            name = _short(r["Kernel_Name"])
            if rx and not rx.search(name):
                continue
            key = (r.get("Dispatch_Id") or r.get("Correlation_Id"), d)
            cval = per[name][r["Counter_Name"]]
            cval[key] = cval.get(key, 0.0) + float(r["Counter_Value"])
    out = {}
    for name, cs in per.items():
        out[name] = {c: sum(v.values()) / len(v) for c, v in cs.items()}
        out[name]["dispatches"] = max(len(v) for v in cs.values())
    return out


def _one(dirs: List[str], counter: str, kernel: str) -> float:
    c = counters(dirs, kernel)
    vals = [v[counter] for v in c.values() if counter in v]
    if len(vals) != 1:
        raise SystemExit(f"{counter}: expected one kernel matching {kernel!r}, found {sorted(c)}")
    return vals[0]


def traffic(fetch: str, write: str, calib_fetch: str, calib_write: str, kernel: str,
            calib_bytes: int = 4 << 30, calib_kernel: str = "") -> dict:
    """HBM bytes per launch of `kernel` from FETCH_SIZE / WRITE_SIZE, each
    scaled by the factor that makes the copy kernel's reading equal its
    known bytes (calib_bytes read and written in total over its dispatches)."""
    cf = counters([calib_fetch], calib_kernel)
    cw = counters([calib_write], calib_kernel)
    # the copy may be issued as several dispatches: total over them
    f_cal = sum(v["FETCH_SIZE"] * v["dispatches"] for v in cf.values() if "FETCH_SIZE" in v) * 1024.0
    w_cal = sum(v["WRITE_SIZE"] * v["dispatches"] for v in cw.values() if "WRITE_SIZE" in v) * 1024.0
    sf, sw = calib_bytes / f_cal, calib_bytes / w_cal
    f = _one([fetch], "FETCH_SIZE", kernel) * 1024.0
    w = _one([write], "WRITE_SIZE", kernel) * 1024.0
    return {"kernel_regex": kernel, "fetch_KiB": f / 1024.0, "write_KiB": w / 1024.0, "fetch_scale": sf,
            "write_scale": sw, "bytes_per_launch": f * sf + w * sw}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m mbsymmetry_amd.pmc")
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("stats")
    a.add_argument("dir")
    a = sub.add_parser("counters")
    a.add_argument("dirs", nargs="+")
    a.add_argument("--kernel", default="")
    a = sub.add_parser("traffic")
    for k in ("fetch", "write", "calib_fetch", "calib_write"):
        a.add_argument("--" + k.replace("_", "-"), dest=k, required=True)
    a.add_argument("--kernel", required=True)
    a.add_argument("--calib-kernel", default="")
    a.add_argument("--calib-bytes", type=int, default=4 << 30)
    a.add_argument("--out")
    args = ap.parse_args(argv)
    if args.cmd == "stats":
        rows = kernel_stats(args.dir)
        print("| kernel | calls | average ms | total % |\n|---|---|---|---|")
        for r in rows:
            print(f"| `{r['kernel']}` | {r['calls']} | {r['avg_ms']:.3f} | {r['percent']:.2f} |")
    elif args.cmd == "counters":
        res = counters(args.dirs, args.kernel)
        names = sorted({c for v in res.values() for c in v})
        print("| counter | " + " | ".join(f"`{k}`" for k in res) + " |")
        print("|---|" + "---|" * len(res))
        for c in names:
            print(f"| {c} | " + " | ".join(f"{res[k].get(c, float('nan')):.4e}" for k in res) + " |")
    else:
        t = traffic(args.fetch, args.write, args.calib_fetch, args.calib_write, args.kernel, args.calib_bytes,
                    args.calib_kernel)
        text = json.dumps(t, indent=1)
        if args.out:
            with open(args.out, "w") as f:
                f.write(text + "\n")
        print(text)
    return 0


if __name__ == "__main__":
    sys.exit(main())
