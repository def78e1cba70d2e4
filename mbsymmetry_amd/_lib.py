"""ctypes binding of libsymgw (include/symgw.h).

The library is loaded from this package directory only; a missing library is
an error (there is no CPU fallback for the device path).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsymgw.so")

SYMGW_OK = 0
PREC_FP64 = 0
PREC_FP32 = 1
PREC_MIXED = 2
FLAG_STREAM_V = 0x100

_i32p = C.POINTER(C.c_int32)
_dp = C.POINTER(C.c_double)


class LayoutDesc(C.Structure):
    _fields_ = [
        ("nk", C.c_int32), ("ink", C.c_int32), ("ns", C.c_int32), ("nts", C.c_int32),
        ("nw", C.c_int32), ("nw_b", C.c_int32),
        ("kao_nblocks", _i32p), ("ao_sizes", _i32p), ("qaux_nblocks", _i32p), ("aux_sizes", _i32p),
        ("irre_list", _i32p), ("index", _i32p), ("irre_index", _i32p), ("conj_index", _i32p),
        ("kpq", _i32p), ("kmq", _i32p),
        ("orep_nblocks", _i32p), ("orep_irreps", _i32p), ("auxrep_nblocks", _i32p), ("auxrep_irreps", _i32p),
        ("kpair_nblocks", _i32p), ("V_irreps", _i32p),
    ]


class LayoutInfo(C.Structure):
    _fields_ = [
        ("flat_ao_size", C.c_int64), ("flat_ao_size_full", C.c_int64), ("flat_aux_size", C.c_int64),
        ("flat_aux_size_full", C.c_int64), ("flat_V_size", C.c_int64), ("flat_orep_size", C.c_int64),
        ("flat_auxrep_size", C.c_int64), ("n_P0_blocks", C.c_int64), ("n_Sigma_blocks", C.c_int64),
        ("flop_P0", C.c_double), ("flop_Sigma", C.c_double), ("device_bytes", C.c_int64),
    ]


class Plan(C.Structure):
    """symgw_plan: rules of the work-list planner (the autotuner's knobs)."""
    _fields_ = [
        ("sig_chunk_cost", C.c_double), ("sig_rows", C.c_int32), ("sig_max_times", C.c_int32),
        ("p0_chunk_cost", C.c_double), ("p0_band_cap", C.c_int32), ("reserved", C.c_int32),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_ if f != "reserved"}


# every entry point declared in include/symgw.h: name -> (restype, argtypes)
SIGNATURES = {
    "symgw_plan_defaults": (C.c_int, [C.POINTER(Plan)]),
    "symgw_replan": (C.c_int, [C.c_void_p, C.POINTER(Plan)]),
    "symgw_get_plan": (C.c_int, [C.c_void_p, C.POINTER(Plan)]),
    "symgw_layout_query": (C.c_int, [C.POINTER(LayoutDesc), C.POINTER(LayoutInfo)]),
    "symgw_kpair_offset": (C.c_int64, [C.c_void_p, C.c_int, C.c_int]),
    "symgw_layout_footprint": (C.c_int, [C.POINTER(LayoutDesc), C.c_int, C.POINTER(C.c_int64)]),
    "symgw_device_memory": (C.c_int, [C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "symgw_measure_peaks": (C.c_int, [C.c_int, C.POINTER(C.c_double)]),
    "symgw_create": (C.c_int, [C.POINTER(LayoutDesc), C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    "symgw_destroy": (C.c_int, [C.c_void_p]),
    "symgw_last_error": (C.c_char_p, [C.c_void_p]),
    "symgw_create_error": (C.c_char_p, []),
    "symgw_layout_info_ctx": (C.c_int, [C.c_void_p, C.POINTER(LayoutInfo)]),
    "symgw_set_V": (C.c_int, [C.c_void_p, _dp, C.c_int64]),
    "symgw_generate_V": (C.c_int, [C.c_void_p, C.c_uint64, C.c_double, _dp, C.c_int]),
    "symgw_generate_V_pair_host": (C.c_int, [C.POINTER(LayoutDesc), C.c_uint64, C.c_double, _dp, C.c_int,
                                             C.c_int, C.c_int, _dp]),
    "symgw_generate_V_pairs_host": (C.c_int, [C.POINTER(LayoutDesc), C.c_uint64, C.c_double, _dp, C.c_int, C.c_int,
                                              _i32p, C.POINTER(C.c_int64), _dp]),
    "symgw_set_reps": (C.c_int, [C.c_void_p, _dp, _dp]),
    "symgw_upload_G": (C.c_int, [C.c_void_p, _dp]),
    "symgw_download_G": (C.c_int, [C.c_void_p, _dp]),
    "symgw_solve_P0": (C.c_int, [C.c_void_p]),
    "symgw_copy_P0_to_host": (C.c_int, [C.c_void_p, _dp]),
    "symgw_upload_P": (C.c_int, [C.c_void_p, _dp]),
    "symgw_set_transforms": (C.c_int, [C.c_void_p, _dp, _dp, _dp, _dp, _dp]),
    "symgw_solve_W": (C.c_int, [C.c_void_p]),
    "symgw_solve_W_noalloc": (C.c_int, [C.c_void_p]),
    "symgw_solve_Sigma": (C.c_int, [C.c_void_p]),
    "symgw_copy_Sigma_to_host": (C.c_int, [C.c_void_p, _dp]),
    "symgw_upload_Sigma": (C.c_int, [C.c_void_p, _dp]),
    "symgw_set_fock": (C.c_int, [C.c_void_p, _dp, _dp]),
    "symgw_dyson": (C.c_int, [C.c_void_p, C.c_double, C.c_double]),
    "symgw_density": (C.c_int, [C.c_void_p, _dp, _dp]),
    "symgw_mu_search": (C.c_int, [C.c_void_p, C.c_double, _dp, C.c_double, C.c_double, C.c_double, C.c_double,
                                  _dp, _dp]),
    "symgw_gw_iteration": (C.c_int, [C.c_void_p, C.c_double, C.c_int]),
    "symgw_set_shard": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "symgw_sync": (C.c_int, [C.c_void_p]),
    "symgw_last_times": (C.c_int, [C.c_void_p, _dp]),
    "symgw_set_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "symgw_ir_basis": (C.c_int, [C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), _dp, _dp, _dp,
                                 _dp, C.POINTER(C.c_int64), _dp]),
    "symgw_ir_uhat": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp, C.c_int,
                                C.POINTER(C.c_int64), _dp]),
}

_lib = None


def load():
    """Load libsymgw.so from the package directory (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run `python -m mbsymmetry_amd.build`")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def dptr(a: np.ndarray):
    if not isinstance(a, np.ndarray) or a.dtype not in (np.float64, np.complex128) or not a.flags["C_CONTIGUOUS"]:
        raise TypeError("expected a C-contiguous float64 or complex128 numpy array")
    return a.ctypes.data_as(_dp)


def checked(a, dtype, n, what):
    """`a` as a C-contiguous array of `dtype` holding exactly `n` elements
    (the length the library reads or writes for it); ValueError otherwise."""
    a = np.ascontiguousarray(a, dtype=dtype)
    if a.size != n:
        raise ValueError(f"{what}: {a.size} elements given, the layout needs {n}")
    return a


def checked_out(a, n, what):
    """An output buffer the library writes in place: must already be a
    C-contiguous complex128 array of exactly `n` elements."""
    if not isinstance(a, np.ndarray) or a.dtype != np.complex128 or not a.flags["C_CONTIGUOUS"]:
        raise TypeError(f"{what}: expected a C-contiguous complex128 numpy array")
    if a.size != n:
        raise ValueError(f"{what}: {a.size} elements given, the layout needs {n}")
    return a


class SymGWError(RuntimeError):
    pass


class Context:
    """One device-resident system (symgw_ctx)."""

    def __init__(self, system, device: int = 0, precision: int = PREC_FP64):
        self.lib = load()
        self.system = system
        self.desc = system.desc()
        h = C.c_void_p()
        rc = self.lib.symgw_create(C.byref(self.desc), device, precision, C.byref(h))
        if rc != SYMGW_OK:
            raise SymGWError(f"symgw_create failed ({rc}): {self.lib.symgw_create_error().decode()}")
        self.h = h
        self.precision = precision
        inf = self.info()
        d = self.desc
        # element counts of every host array the entry points read or write
        self.n_ts_kij = d.nts * d.ns * inf.flat_ao_size
        self.n_q_t_QP = d.nts * inf.flat_aux_size
        self.n_fock = d.ns * inf.flat_ao_size
        self.n_V = inf.flat_V_size
        self.n_orep, self.n_auxrep = inf.flat_orep_size, inf.flat_auxrep_size
        self.nts, self.nw, self.nw_b, self.ink = d.nts, d.nw, d.nw_b, d.ink

    def _chk(self, rc, what):
        if rc != SYMGW_OK:
            raise SymGWError(f"{what} failed ({rc}): {self.lib.symgw_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.symgw_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> LayoutInfo:
        inf = LayoutInfo()
        self._chk(self.lib.symgw_layout_info_ctx(self.h, C.byref(inf)), "layout_info")
        return inf

    def set_V(self, V: np.ndarray):
        V = checked(V, np.complex128, self.n_V, "set_V")
        self._chk(self.lib.symgw_set_V(self.h, dptr(V), V.size), "set_V")

    def generate_V(self, seed: int, scale: float, decay: np.ndarray):
        decay = np.ascontiguousarray(decay, dtype=np.float64)
        self._chk(self.lib.symgw_generate_V(self.h, seed, scale, dptr(decay), decay.size), "generate_V")

    def set_reps(self, orep: np.ndarray, auxrep: np.ndarray):
        o = checked(orep, np.complex128, self.n_orep, "set_reps orep")
        a = checked(auxrep, np.complex128, self.n_auxrep, "set_reps auxrep")
        self._chk(self.lib.symgw_set_reps(self.h, dptr(o) if o.size else None, dptr(a) if a.size else None),
                  "set_reps")

    def upload_G(self, G):
        G = checked(G, np.complex128, self.n_ts_kij, "upload_G")
        self._chk(self.lib.symgw_upload_G(self.h, dptr(G)), "upload_G")

    def download_G(self, shape):
        out = checked_out(np.zeros(shape, dtype=np.complex128), self.n_ts_kij, "download_G")
        self._chk(self.lib.symgw_download_G(self.h, dptr(out)), "download_G")
        return out

    def solve_P0(self):
        self._chk(self.lib.symgw_solve_P0(self.h), "solve_P0")

    def copy_P0_to_host(self, out: np.ndarray):
        checked_out(out, self.n_q_t_QP, "copy_P0_to_host")
        self._chk(self.lib.symgw_copy_P0_to_host(self.h, dptr(out)), "copy_P0_to_host")
        return out

    def upload_P(self, P):
        P = checked(P, np.complex128, self.n_q_t_QP, "upload_P")
        self._chk(self.lib.symgw_upload_P(self.h, dptr(P)), "upload_P")

    def set_transforms(self, tr):
        ni = self.nts - 2
        want = ((tr.Tnt, self.nw * ni, "Tnt"), (tr.Ttn, self.nts * self.nw, "Ttn"),
                (tr.Tnt_BF, self.nw_b * ni, "Tnt_BF"), (tr.Ttn_FB, self.nts * self.nw_b, "Ttn_FB"),
                (tr.iwn, self.nw, "iwn"))
        arrs = [checked(a, np.complex128, n, "set_transforms " + name) for a, n, name in want]
        self._keep = arrs
        self._chk(self.lib.symgw_set_transforms(self.h, *[dptr(a) for a in arrs]), "set_transforms")

    def solve_W(self):
        self._chk(self.lib.symgw_solve_W(self.h), "solve_W")

    def solve_W_noalloc(self):
        self._chk(self.lib.symgw_solve_W_noalloc(self.h), "solve_W_noalloc")

    def solve_Sigma(self):
        self._chk(self.lib.symgw_solve_Sigma(self.h), "solve_Sigma")

    def copy_Sigma_to_host(self, out: np.ndarray):
        checked_out(out, self.n_ts_kij, "copy_Sigma_to_host")
        self._chk(self.lib.symgw_copy_Sigma_to_host(self.h, dptr(out)), "copy_Sigma_to_host")
        return out

    def upload_Sigma(self, S):
        S = checked(S, np.complex128, self.n_ts_kij, "upload_Sigma")
        self._chk(self.lib.symgw_upload_Sigma(self.h, dptr(S)), "upload_Sigma")

    def set_fock(self, F, S):
        F = checked(F, np.complex128, self.n_fock, "set_fock F")
        S = checked(S, np.complex128, self.n_fock, "set_fock S")
        self._chk(self.lib.symgw_set_fock(self.h, dptr(F), dptr(S)), "set_fock")

    def dyson(self, mu: float, sigma_scale: float):
        self._chk(self.lib.symgw_dyson(self.h, mu, sigma_scale), "dyson")

    def density(self, kweights):
        kw = checked(kweights, np.float64, self.ink, "density kweights")
        v = C.c_double()
        self._chk(self.lib.symgw_density(self.h, dptr(kw), C.byref(v)), "density")
        return v.value

    def mu_search(self, nel, kweights, sigma_scale, lo, hi, tol):
        kw = checked(kweights, np.float64, self.ink, "mu_search kweights")
        mu, n = C.c_double(), C.c_double()
        self._chk(self.lib.symgw_mu_search(self.h, nel, dptr(kw), sigma_scale, lo, hi, tol, C.byref(mu), C.byref(n)),
                  "mu_search")
        return mu.value, n.value

    def gw_iteration(self, mu: float, use_graph: bool = True):
        self._chk(self.lib.symgw_gw_iteration(self.h, mu, 1 if use_graph else 0), "gw_iteration")

    def set_shard(self, first: int, count: int):
        self._chk(self.lib.symgw_set_shard(self.h, first, count), "set_shard")

    def replan(self, **rules):
        """Rebuild the work lists under the default plan with `rules` changed
        (sig_chunk_cost, sig_rows, sig_max_times, p0_chunk_cost, p0_band_cap)."""
        p = plan_defaults()
        for k, v in rules.items():
            if k not in p.as_dict():
                raise KeyError(f"unknown plan rule {k}")
            setattr(p, k, v)
        self._chk(self.lib.symgw_replan(self.h, C.byref(p)), "replan")

    def plan(self) -> dict:
        p = Plan()
        self._chk(self.lib.symgw_get_plan(self.h, C.byref(p)), "get_plan")
        return p.as_dict()

    def sync(self):
        self._chk(self.lib.symgw_sync(self.h), "sync")

    def set_timing(self, on: bool):
        self._chk(self.lib.symgw_set_timing(self.h, 1 if on else 0), "set_timing")

    def last_times(self):
        out = np.zeros(5)
        self._chk(self.lib.symgw_last_times(self.h, dptr(out)), "last_times")
        return out


def plan_defaults() -> Plan:
    lib = load()
    p = Plan()
    if lib.symgw_plan_defaults(C.byref(p)) != 0:
        raise SymGWError("symgw_plan_defaults failed")
    return p


def layout_query(system) -> LayoutInfo:
    lib = load()
    d = system.desc()
    inf = LayoutInfo()
    rc = lib.symgw_layout_query(C.byref(d), C.byref(inf))
    if rc != SYMGW_OK:
        raise SymGWError(f"symgw_layout_query failed: {lib.symgw_create_error().decode()}")
    return inf


def layout_footprint(system, precision=PREC_FP64) -> int:
    """Device bytes a context of `system` allocates in `precision`."""
    lib = load()
    d = system.desc()
    out = C.c_int64(0)
    rc = lib.symgw_layout_footprint(C.byref(d), precision, C.byref(out))
    if rc != SYMGW_OK:
        raise SymGWError(f"symgw_layout_footprint failed: {lib.symgw_create_error().decode()}")
    return int(out.value)


def device_memory(device=0):
    """(free, total) bytes of `device`."""
    lib = load()
    fr, tot = C.c_int64(0), C.c_int64(0)
    rc = lib.symgw_device_memory(device, C.byref(fr), C.byref(tot))
    if rc != SYMGW_OK:
        raise SymGWError(f"symgw_device_memory failed: {lib.symgw_create_error().decode()}")
    return int(fr.value), int(tot.value)


def measure_peaks(device=0):
    """Measured peaks of `device`: FP64 and FP32 matrix-core TFLOP/s of the
    contraction kernels' complex MFMA step, and HBM copy GB/s."""
    lib = load()
    out = (C.c_double * 3)()
    rc = lib.symgw_measure_peaks(device, out)
    if rc != SYMGW_OK:
        raise SymGWError(f"symgw_measure_peaks failed: {lib.symgw_create_error().decode()}")
    return {"fp64_mfma_tflops": out[0], "fp32_mfma_tflops": out[1], "hbm_copy_gbps": out[2]}


def generate_V_pair_host(system, seed, scale, decay, k1, k2, n_complex):
    lib = load()
    p = k1 * system.nk + k2
    offs = system.kpair_offsets()
    if p not in offs:
        raise ValueError(f"generate_V_pair_host: pair ({k1}, {k2}) is not stored")
    if n_complex != offs[p][1]:
        raise ValueError(f"generate_V_pair_host: {n_complex} elements requested, the pair holds {offs[p][1]}")
    d = system.desc()
    out = np.zeros(n_complex, dtype=np.complex128)
    decay = np.ascontiguousarray(decay, dtype=np.float64)
    rc = lib.symgw_generate_V_pair_host(C.byref(d), seed, scale, dptr(decay), decay.size, k1, k2, dptr(out))
    if rc != SYMGW_OK:
        raise SymGWError("symgw_generate_V_pair_host failed")
    return out


def ir_basis(statistics: int, lam: float, nl: int, deg: int = 24, max_seg: int = 256):
    """IR basis of the logistic kernel (symgw_ir_basis): dict with the
    segment edges, per-segment Legendre coefficients coef[l][seg][k], the
    singular values and the tau / Matsubara sampling points with u^ there."""
    lib = load()
    nseg = C.c_int(0)
    edges = np.zeros(max_seg + 1)
    coef = np.zeros((nl, max_seg, deg))
    s = np.zeros(nl)
    xs = np.zeros(nl)
    ws = np.zeros(nl, dtype=np.int64)
    uwl = np.zeros((nl, nl, 2))
    rc = lib.symgw_ir_basis(statistics, lam, nl, max_seg, deg, C.byref(nseg), dptr(edges), dptr(coef), dptr(s),
                            dptr(xs), ws.ctypes.data_as(C.POINTER(C.c_int64)), dptr(uwl))
    if rc != SYMGW_OK:
        raise SymGWError(f"symgw_ir_basis failed ({rc}) for statistics {statistics}, Lambda {lam}, nl {nl}")
    ns = nseg.value
    return {"statistics": statistics, "lambda": lam, "nl": nl, "deg": deg, "max_seg": max_seg, "nseg": ns,
            "edges": edges, "coef": coef, "s": s, "xsample": xs, "wsample": ws,
            "uwl": uwl[..., 0] + 1j * uwl[..., 1]}


def ir_uhat(basis, n):
    """u^_l(i w_n) of a basis from ir_basis at the Matsubara indices n: (len(n), nl)."""
    lib = load()
    n = np.ascontiguousarray(n, dtype=np.int64)
    out = np.zeros((n.size, basis["nl"], 2))
    rc = lib.symgw_ir_uhat(basis["statistics"], basis["nl"], basis["nseg"], basis["deg"], basis["max_seg"],
                           dptr(basis["edges"]), dptr(basis["coef"]), n.size, n.ctypes.data_as(C.POINTER(C.c_int64)),
                           dptr(out))
    if rc != SYMGW_OK:
        raise SymGWError(f"symgw_ir_uhat failed ({rc})")
    return out[..., 0] + 1j * out[..., 1]


def generate_V_pairs_host(system, seed, scale, decay, pairs):
    """Host twin of the device generator for a list of stored pairs (k1, k2):
    returns (V, offsets) with pair i's flat slice at V[offsets[i]:]."""
    lib = load()
    offs = system.kpair_offsets()
    ps = np.ascontiguousarray(np.asarray(pairs, dtype=np.int32).reshape(-1, 2))
    sizes = []
    for k1, k2 in ps:
        p = int(k1) * system.nk + int(k2)
        if p not in offs:
            raise ValueError(f"generate_V_pairs_host: pair ({k1}, {k2}) is not stored")
        sizes.append(offs[p][1])
    start = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64) if sizes else np.zeros(0, np.int64)
    out = np.zeros(int(sum(sizes)), dtype=np.complex128)
    decay = np.ascontiguousarray(decay, dtype=np.float64)
    d = system.desc()
    rc = lib.symgw_generate_V_pairs_host(C.byref(d), seed, scale, dptr(decay), decay.size, len(ps),
                                         ps.ctypes.data_as(_i32p), start.ctypes.data_as(C.POINTER(C.c_int64)),
                                         dptr(out) if out.size else None)
    if rc != SYMGW_OK:
        raise SymGWError(f"symgw_generate_V_pairs_host failed: {lib.symgw_create_error().decode()}")
    return out, start
