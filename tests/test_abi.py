"""C ABI checks that need no GPU: the library loads, exports every entry point
declared in include/symgw.h, validates layouts and runs its host-side parts."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from mbsymmetry_amd import _lib
from mbsymmetry_amd import system as S
from oracle.oracle import Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "symgw.h")).read()
    return sorted(set(re.findall(r"\b(symgw_[a-z_A-Z0-9]+)\s*\(", text)))


def test_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_symbols()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)


def test_layout_query_matches_oracle_and_flop_model(config1, lih):
    for sysm, p0, sg, nb in ((config1, 6.37927424e9, 12.75854848e9, 640), (lih, 2.918034304e9, 5.836068608e9, 1553)):
        inf = _lib.layout_query(sysm)
        sz = Oracle(sysm).flat_sizes()
        assert inf.flat_ao_size == sz["flat_ao"] and inf.flat_aux_size == sz["flat_aux"]
        assert inf.flat_ao_size_full == sz["flat_ao_full"] and inf.flat_aux_size_full == sz["flat_aux_full"]
        assert inf.flat_V_size == sz["flat_V"] == sysm.flat_V()
        # FLOP model of analysis.cpp:229-298 (8MNK per complex GEMM)
        assert abs(inf.flop_P0 - p0) < 1 and abs(inf.flop_Sigma - sg) < 1
        assert inf.n_P0_blocks == nb


def test_layout_errors_are_reported():
    sysm = S.tiny()
    lib = _lib.load()
    d = sysm.desc()
    d.nts = 9  # odd nts is rejected (cugw_solver_block.h:39-40)
    inf = _lib.LayoutInfo()
    assert lib.symgw_layout_query(C.byref(d), C.byref(inf)) == 1
    assert b"nts" in lib.symgw_create_error()
    d = sysm.desc()
    d.ns = 3
    assert lib.symgw_layout_query(C.byref(d), C.byref(inf)) == 1


def test_host_generator_deterministic_and_gaussian_like():
    sysm = S.tiny(n=2, ink=4, ao=(5, 4), aux=(7, 6))
    V1 = sysm.host_V()
    V2 = sysm.host_V()
    assert np.array_equal(V1, V2)
    sysm.v_seed += 1
    V3 = sysm.host_V()
    assert not np.array_equal(V1, V3)
    # undo the exp(-|i-j|/4) factor on one block and check unit variance per part
    p = next(iter(sysm.kpair_offsets()))
    o, _ = sysm.kpair_offsets()[p]
    qs, a1, a2 = sysm.V_block_sizes(p)[0]
    iq, i1, i2 = sysm.V_irreps[p][0]
    k1, k2 = divmod(p, sysm.nk)
    ab = np.cumsum([0] + sysm.ao_sizes[k1])[i1]
    bb = np.cumsum([0] + sysm.ao_sizes[k2])[i2]
    d = np.abs((ab + np.arange(a1))[:, None] - (bb + np.arange(a2))[None, :])
    blk = V1[o:o + qs * a1 * a2].reshape(qs, a1, a2) / (np.exp(-d / 4.0)[None] / np.sqrt(2.0))
    allv = np.concatenate([blk.real.ravel(), blk.imag.ravel()])
    assert abs(allv.mean()) < 0.3 and 0.6 < allv.var() < 1.4


def test_create_without_device_fails_cleanly():
    """On a machine without a gfx950 device creation returns an error, never a fallback."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("device present")
    sysm = S.tiny()
    with pytest.raises(_lib.SymGWError):
        _lib.Context(sysm)


def test_planner_fits_every_block_width():
    """The host planner finds work lists within the LDS and accumulator
    capacity of a workgroup for ao blocks of every width up to 256 and aux
    blocks up to 256, in both precisions: the widths sit on the edges of the
    planner's choices (16-wide tiles, whole-block Sigma slabs up to 32 / 48,
    64- / 80- / 96-wide bands, P0 slabs stacked up to 192 or 256 aux rows).
    symgw_layout_footprint runs the same planner as symgw_create and reports
    its errors."""
    from mbsymmetry_amd.transforms import chebyshev
    widths = [1, 15, 16, 17, 31, 32, 33, 47, 48, 49, 63, 64, 65, 79, 80, 81, 96, 97, 128, 129, 160, 161, 200,
              255, 256]
    for k1 in widths:
        for qs in (1, 17, 191, 192, 193, 256):
            sysm = S.synthetic("widths", 2, 4, [k1, 5], [qs, 7], 1, chebyshev(4, 6.0), 3)
            for prec in (_lib.PREC_FP64, _lib.PREC_FP32):
                assert _lib.layout_footprint(sysm, prec) > 0, (k1, qs, prec)
