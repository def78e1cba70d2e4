"""MI355X-native symmetry-adapted GW hot path (P0, W, Sigma, Dyson) behind a C ABI.

libsymgw.so (built in-tree from csrc/ by mbsymmetry_amd.build) holds the gfx950
kernels and the extern "C" entry points of include/symgw.h; _lib binds them.
"""
from . import _lib  # noqa: F401
from ._lib import Context, PREC_FP32, PREC_FP64, SymGWError  # noqa: F401

__all__ = ["Context", "PREC_FP64", "PREC_FP32", "SymGWError"]
