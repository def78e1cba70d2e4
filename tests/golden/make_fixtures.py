"""Convert the reference's HDF5 input fixtures into numpy .npz fixtures.

Run in the development container only (it reads the reference tree and uses
the HDF5 command-line tools under /opt/conda/bin; the GPU box has neither):

    python tests/golden/make_fixtures.py

Sources (CQMP/MBSymmetry, data files of its own tests, MIT licence):
  SymmetryMBPT/data/LiH/integral/nk2/block/symmetry_info.h5   block tables
  SymmetryMBPT/data/LiH/integral/nk2/block/df_int/meta.h5     k mesh, sizes
  SymmetryMBPT/data/LiH/integral/nk2/block/df_int/VQ_flat_0.h5 flat V
  SymmetryMBPT/data/LiH/integral/nk2/block/input_flat.h5      F, S, H (flat)
  SymmetryMBPT/data/LiH/integral/nk2/block/input_blk.h5       F, S, H (dense)
  SymmetryMBPT/data/ir_grid/1e3_72.h5                         IR basis, Lambda=1e3
Outputs: tests/golden/lih_nk2.npz, tests/golden/ir_1e3_72.npz (data only).
"""
import os
import re
import subprocess
import tempfile

import numpy as np

REF = "/root/reference/SymmetryMBPT/data"
H5DUMP = "/opt/conda/bin/h5dump"
H5LS = "/opt/conda/bin/h5ls"
OUT = os.path.dirname(os.path.abspath(__file__))

_TYPES = {
    "H5T_STD_I32LE": np.int32,
    "H5T_STD_I64LE": np.int64,
    "H5T_IEEE_F64LE": np.float64,
}


def read(path, dset):
    head = subprocess.run([H5DUMP, "-H", "-d", dset, path], capture_output=True, text=True, check=True).stdout
    dtype = _TYPES[re.search(r"DATATYPE\s+(\S+)", head).group(1)]
    m = re.search(r"DATASPACE\s+SIMPLE\s*\{\s*\(([^)]*)\)", head)
    shape = tuple(int(x) for x in m.group(1).split(",")) if m else ()
    with tempfile.TemporaryDirectory() as td:
        fn = os.path.join(td, "d.bin")
        subprocess.run([H5DUMP, "-d", dset, "-b", "LE", "-o", fn, path], capture_output=True, check=True)
        arr = np.fromfile(fn, dtype=dtype)
    return arr.reshape(shape) if shape else arr.reshape(())


def datasets(path):
    out = subprocess.run([H5LS, "-r", path], capture_output=True, text=True, check=True).stdout
    return [l.split()[0] for l in out.splitlines() if "Dataset" in l]


def lih():
    base = os.path.join(REF, "LiH/integral/nk2")
    sym = os.path.join(base, "block/symmetry_info.h5")
    meta = os.path.join(base, "block/df_int/meta.h5")
    out = {}
    for d in datasets(sym):
        out["sym" + d.replace("/", "__")] = read(sym, d)
    for d in datasets(meta):
        out["meta" + d.replace("/", "__")] = read(meta, d)
    out["V_flat"] = read(os.path.join(base, "block/df_int/VQ_flat_0.h5"), "/0")
    for name, fn in (("flat", "block/input_flat.h5"), ("blk", "block/input_blk.h5")):
        p = os.path.join(base, fn)
        for key in ("Fock-k", "S-k", "H-k"):
            out[f"{name}__{key}"] = read(p, "/HF/" + key)
    dense_meta = os.path.join(base, "df_int/meta.h5")
    for d in datasets(dense_meta):
        out["dense_meta" + d.replace("/", "__")] = read(dense_meta, d)
    np.savez_compressed(os.path.join(OUT, "lih_nk2.npz"), **out)


def ir():
    p = os.path.join(REF, "ir_grid/1e3_72.h5")
    out = {}
    for d in datasets(p):
        if d.endswith("/uwl"):  # compound complex copy of uwl_re / uwl_im
            continue
        out[d.strip("/").replace("/", "__")] = read(p, d)
    np.savez_compressed(os.path.join(OUT, "ir_1e3_72.npz"), **out)


if __name__ == "__main__":
    lih()
    ir()
    for f in ("lih_nk2.npz", "ir_1e3_72.npz"):
        print(f, os.path.getsize(os.path.join(OUT, f)))
