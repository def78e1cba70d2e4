"""The C ABI as a compiled C host sees it: tests/abi_consumer.c, built with
gcc against include/symgw.h and the library only.  On the CPU: struct sizes
and the layout query agree with the library's Python binding.  On the GPU:
P0 and Sigma of LiH run through the C program and are accumulated into host
buffers that already hold data (the += of cugw_routines_block.cu:236-274)."""
import ctypes as C
import json
import os
import subprocess

import numpy as np
import pytest

from mbsymmetry_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "abi_consumer")


def _run(*args):
    if not os.path.exists(EXE):
        pytest.fail("tests/abi_consumer is not built (run __graft_entry__.build())")
    r = subprocess.run([EXE, *args], stdout=subprocess.PIPE, stderr=subprocess.PIPE, timeout=600)
    assert r.returncode == 0, r.stderr.decode()
    return r.stdout.decode()


def write_desc(sysm, d):
    """desc.bin: the header and the 16 index arrays of symgw_layout_desc."""
    nk = sysm.nk
    arrays = [
        [len(s) for s in sysm.ao_sizes], [b for s in sysm.ao_sizes for b in s],
        [len(s) for s in sysm.aux_sizes], [b for s in sysm.aux_sizes for b in s],
        sysm.irre_list, sysm.index, sysm.irre_index, sysm.conj_index, sysm.kpq, sysm.kmq,
        [len(x) for x in sysm.orep_irreps], [v for x in sysm.orep_irreps for pr in x for v in pr],
        [len(x) for x in sysm.auxrep_irreps], [v for x in sysm.auxrep_irreps for pr in x for v in pr],
        [len(sysm.V_irreps.get(p, [])) for p in range(nk * nk)],
        [v for p in range(nk * nk) for t in sysm.V_irreps.get(p, []) for v in t],
    ]
    parts = [np.array([nk, sysm.ink, sysm.ns, sysm.nts, sysm.transforms.nw, sysm.transforms.nw_b], np.int32)]
    for a in arrays:
        a = np.asarray(a, np.int32).ravel()
        parts += [np.array([a.size], np.int32), a]
    np.concatenate(parts).tofile(os.path.join(d, "desc.bin"))


def test_struct_sizes_match_the_binding():
    got = json.loads(_run("sizes"))
    assert got["symgw_layout_desc"] == C.sizeof(_lib.LayoutDesc)
    assert got["symgw_layout_info"] == C.sizeof(_lib.LayoutInfo)


@pytest.mark.parametrize("name", ["lih", "config1"])
def test_query_matches_the_binding(name, lih, config1, tmp_path):
    sysm = {"lih": lih, "config1": config1}[name]
    write_desc(sysm, str(tmp_path))
    got = json.loads(_run("query", str(tmp_path)))
    info = _lib.layout_query(sysm)
    for f in ("flat_ao_size", "flat_ao_size_full", "flat_aux_size", "flat_aux_size_full", "flat_V_size",
              "flat_orep_size", "flat_auxrep_size", "n_P0_blocks", "n_Sigma_blocks"):
        assert got[f] == getattr(info, f), f
    assert got["flop_P0"] == info.flop_P0 and got["flop_Sigma"] == info.flop_Sigma
    assert got["footprint_fp64"] == _lib.layout_footprint(sysm, _lib.PREC_FP64)


@pytest.mark.gpu
def test_c_host_accumulates_P0_and_Sigma(lih, tmp_path):
    from oracle.oracle import Oracle
    sysm = lih
    d = str(tmp_path)
    write_desc(sysm, d)
    orc = Oracle(sysm)
    G0 = orc.dyson(sysm.zeros_ts_kij(), 0.0)
    V = sysm.host_V()
    Gf = orc.rotate_G(G0)
    P0_ref = orc.P0(V, Gf)
    P = orc.W(P0_ref)
    S_ref = orc.Sigma(V, Gf, orc.rotate_P(P))
    rng = np.random.default_rng(5)
    P0_init = rng.standard_normal(P0_ref.shape) + 1j * rng.standard_normal(P0_ref.shape)
    S_init = rng.standard_normal(S_ref.shape) + 1j * rng.standard_normal(S_ref.shape)
    for name, a in (("V", V), ("G", G0), ("P0_init", P0_init), ("P", P), ("Sig_init", S_init),
                    ("orep", sysm.orep if sysm.orep.size else np.zeros(1, np.complex128)),
                    ("auxrep", sysm.auxrep if sysm.auxrep.size else np.zeros(1, np.complex128))):
        np.ascontiguousarray(a, np.complex128).tofile(os.path.join(d, name + ".bin"))
    print("[case] abi_consumer run on LiH", flush=True)
    assert "ok" in _run("run", d)
    P0_out = np.fromfile(os.path.join(d, "P0_out.bin"), np.complex128)
    S_out = np.fromfile(os.path.join(d, "Sig_out.bin"), np.complex128).reshape(S_ref.shape)
    nk = sysm.nk
    # the device results are not divided by nk; the host buffers kept their data
    dp = np.abs((P0_out - P0_init) / nk - P0_ref).max()
    ds = np.abs((S_out - S_init) / nk - S_ref).max()
    assert dp <= 1e-10 * max(1.0, np.abs(P0_ref).max()), dp
    assert ds <= 1e-10 * max(1.0, np.abs(S_ref).max()), ds
