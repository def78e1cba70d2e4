"""Build libsymgw.so (gfx950 HIP kernels + C ABI) in-tree with hipcc.

Usage: python -m mbsymmetry_amd.build   (or build() from __graft_entry__)
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libsymgw.so")
SOURCES = ["kernels.hip", "contract.hip", "peaks.hip", "capi.cpp", "layout.cpp"]
# host-only C (binary128 arithmetic and OpenMP, built with gcc)
C_SOURCES = ["irbasis.c"]
LINK_LIBS = ["-lquadmath", "-lgomp"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def compile_flags():
    return [
        "--offload-arch=gfx950",
        "-O3",
        "-std=c++17",
        "-fPIC",
        # packed FP32 instructions stay disabled in every HIP translation unit
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wno-unused-command-line-argument",
    ]


def build(verbose: bool = False) -> str:
    objs = []
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    cc = hipcc()
    procs = []
    for src in SOURCES:
        obj = os.path.join(bdir, os.path.splitext(src)[0] + ".o")
        cmd = [cc, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj] + compile_flags()
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
    for src in C_SOURCES:
        obj = os.path.join(bdir, os.path.splitext(src)[0] + ".o")
        cmd = ["gcc", "-O2", "-fPIC", "-fopenmp", "-std=c11", "-Wall", "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
    for src, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        # hipcc's host pass does not know the packed-fp32 feature name; that warning is harmless
        lines = [l for l in text.splitlines() if "packed-fp32-ops" not in l and l.strip()]
        if p.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src}:\n{text}")
        if verbose and lines:
            print("\n".join(lines))
    tmp = LIB + ".tmp"
    cmd = [cc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + LINK_LIBS
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout.decode(errors="replace"))
    os.replace(tmp, LIB)
    return LIB


def build_abi_consumer(verbose: bool = False) -> str:
    """tests/abi_consumer: a C11 program built with gcc against include/symgw.h
    and the library alone (no HIP headers), as a C or C++ host of the
    reference would bind it."""
    root = os.path.dirname(HERE)
    out = os.path.join(root, "tests", "abi_consumer")
    cmd = ["gcc", "-std=c11", "-O1", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(root, "include"),
           os.path.join(root, "tests", "abi_consumer.c"), "-o", out + ".tmp", "-L", HERE,
           "-l:" + os.path.basename(LIB), "-Wl,-rpath,$ORIGIN/../mbsymmetry_amd"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError("abi_consumer build failed:\n" + r.stdout.decode(errors="replace"))
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
