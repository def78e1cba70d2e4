"""ctypes binding of the FP64 CPU oracle (oracle.c) — TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.  build() compiles liboracle.so with the Makefile next to it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
_dp = C.POINTER(C.c_double)
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = C.CDLL(LIB)
        vp = C.c_void_p
        sig = {
            "orc_flat_ao": (C.c_int64, [vp]), "orc_flat_ao_full": (C.c_int64, [vp]),
            "orc_flat_aux": (C.c_int64, [vp]), "orc_flat_aux_full": (C.c_int64, [vp]),
            "orc_flat_V": (C.c_int64, [vp]), "orc_kpair_offset": (C.c_int64, [vp, C.c_int, C.c_int]),
            "orc_rotate_G": (C.c_int, [vp, _dp, _dp, _dp]),
            "orc_rotate_P": (C.c_int, [vp, _dp, _dp, _dp]),
            "orc_P0": (C.c_int, [vp, _dp, _dp, C.c_int, C.c_int, _dp, C.c_int]),
            "orc_P0_sample": (C.c_int, [vp, _dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int, _dp, C.c_int]),
            "orc_Sigma": (C.c_int, [vp, _dp, _dp, _dp, C.c_int, C.c_int, _dp, C.c_int]),
            "orc_W": (C.c_int, [vp, _dp, _dp, _dp, _dp]),
            "orc_dyson": (C.c_int, [vp, _dp, _dp, _dp, C.c_double, _dp, _dp, _dp, _dp]),
            "orc_density": (C.c_double, [vp, _dp, _dp, _dp]),
            "orc_P0_at": (C.c_int, [vp, _dp, C.POINTER(C.c_int64), _dp, C.c_int, C.c_int, _dp, C.c_int]),
            "orc_Sigma_at": (C.c_int, [vp, _dp, C.POINTER(C.c_int64), _dp, _dp, C.c_int, C.c_int, C.c_int, _dp,
                                       C.c_int]),
            "orc_rotate_P_t": (C.c_int, [vp, _dp, _dp, C.c_int, _dp]),
            "orc_inverse_lu": (C.c_int, [C.c_int, _dp, _dp]),
            "orc_inverse_chol": (C.c_int, [C.c_int, _dp, _dp]),
        }
        for k, (r, a) in sig.items():
            f = getattr(lib, k)
            f.restype, f.argtypes = r, a
        _lib = lib
    return _lib


def _p(a):
    assert a.dtype in (np.complex128, np.float64) and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags)
    return a.ctypes.data_as(_dp)


class Oracle:
    """Oracle bound to one System (the same layout descriptor as the product)."""

    def __init__(self, system):
        self.lib = load()
        self.s = system
        self.d = system.desc()
        self.dp = C.cast(C.pointer(self.d), C.c_void_p)

    def flat_sizes(self):
        L = self.lib
        return dict(flat_ao=L.orc_flat_ao(self.dp), flat_ao_full=L.orc_flat_ao_full(self.dp),
                    flat_aux=L.orc_flat_aux(self.dp), flat_aux_full=L.orc_flat_aux_full(self.dp),
                    flat_V=L.orc_flat_V(self.dp))

    def kpair_offset(self, k1, k2):
        return self.lib.orc_kpair_offset(self.dp, k1, k2)

    def rotate_G(self, G_ibz):
        s = self.s
        out = np.zeros((s.nts, s.ns, self.flat_sizes()["flat_ao_full"]), dtype=np.complex128)
        orep = np.ascontiguousarray(s.orep if s.orep.size else np.zeros(1, np.complex128))
        self.lib.orc_rotate_G(self.dp, _p(np.ascontiguousarray(G_ibz)), _p(orep), _p(out))
        return out

    def rotate_P(self, P_ibz):
        s = self.s
        out = np.zeros(s.nts * self.flat_sizes()["flat_aux_full"], dtype=np.complex128)
        aux = np.ascontiguousarray(s.auxrep if s.auxrep.size else np.zeros(1, np.complex128))
        self.lib.orc_rotate_P(self.dp, _p(np.ascontiguousarray(P_ibz)), _p(aux), _p(out))
        return out

    def P0(self, V, G_full, t0=0, t1=None, nthreads=1, out=None):
        s = self.s
        t1 = s.nts // 2 if t1 is None else t1
        out = s.zeros_q_t_QP() if out is None else out
        self.lib.orc_P0(self.dp, _p(V), _p(np.ascontiguousarray(G_full)), t0, t1, _p(out), nthreads)
        return out

    def P0_sample(self, V, G_full, t0, t1, nq, nkk, nthreads=1):
        out = self.s.zeros_q_t_QP()
        self.lib.orc_P0_sample(self.dp, _p(V), _p(np.ascontiguousarray(G_full)), t0, t1, nq, nkk, _p(out), nthreads)
        return out

    def Sigma(self, V, G_full, P_full, t0=0, t1=None, nthreads=1, out=None):
        s = self.s
        t1 = s.nts if t1 is None else t1
        out = s.zeros_ts_kij() if out is None else out
        self.lib.orc_Sigma(self.dp, _p(V), _p(np.ascontiguousarray(G_full)), _p(np.ascontiguousarray(P_full)), t0, t1,
                           _p(out), nthreads)
        return out

    def _voff(self, pairs_off):
        """pair -> complex offset map as the int64[nk*nk] table orc_*_at read."""
        voff = np.full(self.s.nk * self.s.nk, -1, dtype=np.int64)
        for p, o in pairs_off.items():
            voff[p] = o
        return voff

    def P0_at(self, V, pairs_off, G_full, iqr, t, nthreads=1):
        """P0 of iBZ q index iqr at time t summed over every k (qslice
        complex, scaled as P0); V holds pair p's slice at pairs_off[p]."""
        s = self.s
        out = np.zeros(s.qslice(int(s.irre_list[iqr])), dtype=np.complex128)
        voff = self._voff(pairs_off)
        rc = self.lib.orc_P0_at(self.dp, _p(V), voff.ctypes.data_as(C.POINTER(C.c_int64)),
                                _p(np.ascontiguousarray(G_full)), iqr, t, _p(out), nthreads)
        if rc:
            raise RuntimeError(f"orc_P0_at failed ({rc})")
        return out

    def rotate_P_t(self, P_ibz, t):
        s = self.s
        out = np.zeros(self.flat_sizes()["flat_aux_full"], dtype=np.complex128)
        aux = np.ascontiguousarray(s.auxrep if s.auxrep.size else np.zeros(1, np.complex128))
        self.lib.orc_rotate_P_t(self.dp, _p(np.ascontiguousarray(P_ibz)), _p(aux), t, _p(out))
        return out

    def Sigma_at(self, V, pairs_off, G_full, P_full_t, ikr, t, spin=0, nthreads=1):
        """Sigma of iBZ k index ikr at (t, spin) summed over every q (kslice
        complex, scaled as Sigma); P_full_t from rotate_P_t."""
        s = self.s
        out = np.zeros(s.kslice(int(s.irre_list[ikr])), dtype=np.complex128)
        voff = self._voff(pairs_off)
        rc = self.lib.orc_Sigma_at(self.dp, _p(V), voff.ctypes.data_as(C.POINTER(C.c_int64)),
                                   _p(np.ascontiguousarray(G_full)), _p(np.ascontiguousarray(P_full_t)), ikr, t, spin,
                                   _p(out), nthreads)
        if rc:
            raise RuntimeError(f"orc_Sigma_at failed ({rc})")
        return out

    def W(self, P0):
        s, tr = self.s, self.s.transforms
        out = s.zeros_q_t_QP()
        rc = self.lib.orc_W(self.dp, _p(np.ascontiguousarray(P0)), _p(np.ascontiguousarray(tr.Tnt_BF)),
                            _p(np.ascontiguousarray(tr.Ttn_FB)), _p(out))
        if rc:
            raise RuntimeError("oracle W: matrix not positive definite")
        return out

    def dyson(self, Sigma, mu):
        s, tr = self.s, self.s.transforms
        out = s.zeros_ts_kij()
        rc = self.lib.orc_dyson(self.dp, _p(np.ascontiguousarray(Sigma)), _p(np.ascontiguousarray(s.F)),
                                _p(np.ascontiguousarray(s.S)), mu, _p(np.ascontiguousarray(tr.Tnt)),
                                _p(np.ascontiguousarray(tr.Ttn)), _p(np.ascontiguousarray(tr.iwn)), _p(out))
        if rc:
            raise RuntimeError("oracle Dyson: singular matrix")
        return out

    def density(self, G):
        s = self.s
        return self.lib.orc_density(self.dp, _p(np.ascontiguousarray(G)), _p(np.ascontiguousarray(s.S)),
                                    _p(np.ascontiguousarray(s.kweights, dtype=np.float64)))

    def inverse_lu(self, A):
        A = np.ascontiguousarray(A, dtype=np.complex128)
        out = np.zeros_like(A)
        if self.lib.orc_inverse_lu(A.shape[0], _p(A), _p(out)):
            raise np.linalg.LinAlgError("singular")
        return out

    def inverse_chol(self, A):
        A = np.ascontiguousarray(A, dtype=np.complex128)
        out = np.zeros_like(A)
        if self.lib.orc_inverse_chol(A.shape[0], _p(A), _p(out)):
            raise np.linalg.LinAlgError("not positive definite")
        return out

    def gw_iteration(self, G_ibz, mu, nthreads=1, V=None):
        """One plain-FP64 GW iteration on the host: P0 -> W -> Sigma -> Dyson.
        Returns (P0 (scaled), P, Sigma, G_new)."""
        s = self.s
        V = s.host_V() if V is None else V
        Gf = self.rotate_G(G_ibz)
        P0 = self.P0(V, Gf, nthreads=nthreads)
        P = self.W(P0)
        Pf = self.rotate_P(P)
        Sig = self.Sigma(V, Gf, Pf, nthreads=nthreads)
        Gn = self.dyson(Sig, mu)
        return P0, P, Sig, Gn
