"""FP64 tau <-> Matsubara transform matrices (host-side setup).

Three constructions behind one record:
  * IR basis generated by the library (symgw_ir_basis: singular-value
    expansion of the logistic kernel for any cutoff Lambda), assembled into
    the same matrices as below;
  * IR basis read from a fixture of an IR grid file (transformer_t.cpp:7-82):
    Tnc = uwl*sqrt(beta), Tct = pinv(uxl)*sqrt(beta/2), Ttc = [ux1l_minus; uxl;
    ux1l]*sqrt(2/beta), Tcn = pinv(Tnc); Tnt = Tnc Tct, Ttn = Ttc Tcn,
    Tnt_BF = Tnc_B Tct_B Ttc_other[1:-1] Tct, Ttn_FB = Ttc_B_other Tcn_B.
  * Chebyshev basis on the Chebyshev tau mesh (imtime_mesh_t.h:55-64,
    transformer_t.cpp:128-179 for Ttc/Tct).  The reference throws "Chebyshev is
    not implemented" for every Matsubara transform (transformer_t.h:68-69);
    here Tnc is the exact Fourier integral of the Chebyshev polynomials,
        Tnc[n][c] = int_0^beta exp(i w_n tau) T_c(2 tau/beta - 1) dtau,
    evaluated with composite Gauss-Legendre quadrature, on the lowest
    frequencies and their negatives; fermionic functions use ni polynomials,
    bosonic ones ni-1 (the ncheb -> ncheb-1 convention of
    fermi_boson_trans_cheby, transformer_t.cpp:208-238).
"""
import functools
from dataclasses import dataclass
from typing import Optional

import numpy as np
from numpy.polynomial import legendre as _leg


@dataclass
class Transforms:
    beta: float
    ni: int            # interior tau points; nts = ni + 2
    nw: int
    nw_b: int
    tau: np.ndarray    # (nts,) fermionic tau mesh
    wsample: np.ndarray   # (nw,) fermionic Matsubara indices n: w = (2n+1) pi / beta
    wsample_b: np.ndarray  # (nw_b,) bosonic indices m: W = 2 m pi / beta
    Tnt: np.ndarray    # (nw, ni)
    Ttn: np.ndarray    # (nts, nw)
    Tnt_BF: np.ndarray  # (nw_b, ni)
    Ttn_FB: np.ndarray  # (nts, nw_b)
    kind: str
    Tct: Optional[np.ndarray] = None  # (nl, ni) fermionic basis coefficients from the interior tau values

    @property
    def nts(self):
        return self.ni + 2

    @property
    def iwn(self):
        return 1j * (2 * self.wsample + 1) * np.pi / self.beta

    @property
    def iwb(self):
        return 1j * 2 * self.wsample_b * np.pi / self.beta


def _pinv(a):
    return np.linalg.pinv(a, rcond=1e-15)


def ir_from_fixture(z, beta: float) -> Transforms:
    """IR transforms from an npz made of an IR grid file (keys fermi__*, bose__*)."""
    return _ir_transforms(lambda key: z[key], beta)


def _ir_transforms(z, beta: float, kind: str = "ir") -> Transforms:
    """The matrices of transformer_t::read_trans_ir (transformer_t.cpp:48-82)
    from the grid arrays z(key), keys as in the grid files (fermi__*, bose__*)."""
    def stat(p):
        uwl = z(p + "__uwl_re") + 1j * z(p + "__uwl_im")
        uxl = z(p + "__uxl")
        nl = uxl.shape[1]
        Ttc = np.zeros((uxl.shape[0] + 2, nl))
        Ttc[0] = z(p + "__ux1l_minus")
        Ttc[-1] = z(p + "__ux1l")
        Ttc[1:-1] = uxl
        Tcn = _pinv(uwl)
        Tct = _pinv(uxl)
        return uwl, Tcn, Ttc, Tct
    Tnc_f, Tcn_f, Ttc_f, Tct_f = stat("fermi")
    Tnc_b, Tcn_b, Ttc_b, Tct_b = stat("bose")
    other_f = z("fermi__other_uxl")   # (nx_b, nl_f)
    other_b = z("bose__other_uxl")    # (nx_f, nl_b)
    Ttc_other = Ttc_b @ Tct_b @ other_f * np.sqrt(2.0 / beta)
    Ttc_b_other = Ttc_f @ Tct_f @ other_b * np.sqrt(2.0 / beta)
    Ttc_f = Ttc_f * np.sqrt(2.0 / beta)
    Ttc_b = Ttc_b * np.sqrt(2.0 / beta)
    Tct_f = Tct_f * np.sqrt(beta / 2.0)
    Tct_b = Tct_b * np.sqrt(beta / 2.0)
    Tnc_f = Tnc_f * np.sqrt(beta)
    Tnc_b = Tnc_b * np.sqrt(beta)
    Tcn_f = Tcn_f * np.sqrt(1.0 / beta)
    Tcn_b = Tcn_b * np.sqrt(1.0 / beta)
    ncheb_b = Tct_b.shape[0]
    Tnt = Tnc_f @ Tct_f
    Ttn = Ttc_f @ Tcn_f
    Tnt_BF = Tnc_b @ Tct_b @ Ttc_other[1:1 + ncheb_b, :] @ Tct_f
    Ttn_FB = Ttc_b_other @ Tcn_b
    x = z("fermi__xsample")
    tau = np.concatenate([[0.0], (x + 1) * beta / 2, [beta]])
    return Transforms(beta, len(x), Tnt.shape[0], Tnt_BF.shape[0], tau, z("fermi__wsample").astype(np.int64),
                      z("bose__wsample").astype(np.int64), Tnt, Ttn, Tnt_BF, Ttn_FB, kind, Tct=Tct_f)


def ir_eval(basis, x) -> np.ndarray:
    """u_l(x) of a generated basis at the points x: (len(x), nl)."""
    x = np.atleast_1d(np.asarray(x, dtype=np.float64))
    e = basis["edges"][:basis["nseg"] + 1]
    seg = np.clip(np.searchsorted(e, x, side="right") - 1, 0, basis["nseg"] - 1)
    out = np.zeros((x.size, basis["nl"]))
    for sg in np.unique(seg):
        sel = seg == sg
        c, h = 0.5 * (e[sg] + e[sg + 1]), 0.5 * (e[sg + 1] - e[sg])
        out[sel] = _leg.legval((x[sel] - c) / h, basis["coef"][:, sg, :].T).T
    return out


@functools.lru_cache(maxsize=8)
def ir_grid(lam: float, nl_f: int, nl_b: Optional[int] = None) -> dict:
    """The arrays of an IR grid file (the keys transformer_t.cpp:48-82 reads:
    xsample, wsample, uxl, uwl, ux1l, ux1l_minus, other_uxl per statistics)
    from bases generated by the library.  nl_b defaults to nl_f - 1 (the
    pattern of the reference's 1e3_72 grid: 72 fermionic, 71 bosonic)."""
    from . import _lib
    nl_b = nl_f - 1 if nl_b is None else nl_b
    bf = _lib.ir_basis(1, lam, nl_f)
    bb = _lib.ir_basis(0, lam, nl_b)
    out = {}
    for p, b, other in (("fermi", bf, bb), ("bose", bb, bf)):
        out[p + "__xsample"] = b["xsample"]
        out[p + "__wsample"] = b["wsample"]
        out[p + "__uxl"] = ir_eval(b, b["xsample"])
        out[p + "__uwl_re"] = b["uwl"].real.copy()
        out[p + "__uwl_im"] = b["uwl"].imag.copy()
        out[p + "__ux1l"] = ir_eval(b, [1.0])[0]
        out[p + "__ux1l_minus"] = ir_eval(b, [-1.0])[0]
        out[p + "__other_uxl"] = ir_eval(b, other["xsample"])
        out[p + "__s"] = b["s"]
    return out


def ir(beta: float, lam: float, nl_f: int, nl_b: Optional[int] = None) -> Transforms:
    """IR transforms at inverse temperature beta from a basis generated for
    the cutoff Lambda (w_max = Lambda / beta) with nl_f fermionic and nl_b
    bosonic functions."""
    g = ir_grid(float(lam), int(nl_f), None if nl_b is None else int(nl_b))
    return _ir_transforms(lambda key: g[key], beta, kind=f"ir(Lambda={lam:g})")


def leakage(tr: Transforms, G_ts: np.ndarray) -> float:
    """Dyson leakage of sc_type_block.cpp:66-86: the largest magnitude of the
    last basis coefficient of G(tau) over the largest of the first,
    max |c_{L-1}| / max |c_0| with c = Tct G(interior tau), over every spin,
    k point and matrix element.  Large values mean G is not resolved by the
    grid."""
    if tr.Tct is None:
        raise ValueError("these transforms carry no coefficient matrix")
    G = np.asarray(G_ts).reshape(G_ts.shape[0], -1)[1:-1]
    c_last = tr.Tct[-1] @ G
    c_first = tr.Tct[0] @ G
    return float(np.abs(c_last).max() / np.abs(c_first).max())


def chebyshev_tau_mesh(ni: int, beta: float) -> np.ndarray:
    pts = np.zeros(ni + 2)
    pts[-1] = beta
    for it in range(1, ni + 1):
        z = np.cos(np.pi * (it - 0.5) / ni)
        pts[ni + 1 - it] = (z + 1) * beta / 2.0
    return pts


def cheb_T(nc: int, x: np.ndarray) -> np.ndarray:
    """T_c(x) for c < nc, shape (len(x), nc), by the three-term recurrence."""
    x = np.asarray(x, dtype=np.float64)
    T = np.zeros((x.size, nc))
    T[:, 0] = 1.0
    if nc > 1:
        T[:, 1] = x
    for c in range(2, nc):
        T[:, c] = 2 * x * T[:, c - 1] - T[:, c - 2]
    return T


def cheb_fourier(nc: int, omegas: np.ndarray, beta: float) -> np.ndarray:
    """F[n][c] = int_0^beta exp(i w_n tau) T_c(2 tau/beta - 1) dtau (composite Gauss-Legendre)."""
    xg, wg = np.polynomial.legendre.leggauss(24)
    out = np.zeros((omegas.size, nc), dtype=np.complex128)
    for i, w in enumerate(omegas):
        a = w * beta / 2.0
        panels = max(8, int(np.ceil(abs(a) / 2.0)) + nc // 2)
        edges = np.linspace(-1.0, 1.0, panels + 1)
        mid = 0.5 * (edges[1:] + edges[:-1])
        half = 0.5 * (edges[1:] - edges[:-1])
        xs = (mid[:, None] + half[:, None] * xg[None, :]).ravel()
        ws = (half[:, None] * wg[None, :]).ravel()
        T = cheb_T(nc, xs)
        ph = np.exp(1j * a * (xs + 1.0)) * ws
        out[i] = (beta / 2.0) * (ph @ T)
    return out


def chebyshev(ni: int, beta: float, nw_pos: int = None) -> Transforms:
    """Chebyshev transforms with the nw_pos lowest non-negative Matsubara
    frequencies and their negatives (nw = 2*nw_pos fermionic samples,
    nw_b = 2*nw_pos + 1 bosonic ones).  The back transforms are least-squares
    pseudo-inverses: with nw_pos = ni the single-pole closed form is
    reproduced to ~1e-11 (a square sampling with nw = ni is ill-conditioned
    for Chebyshev polynomials, condition number > 1e10)."""
    if ni % 2:
        raise ValueError("number of fermionic Chebyshev nodes must be even")
    nw_pos = ni if nw_pos is None else nw_pos
    tau = chebyshev_tau_mesh(ni, beta)
    x = 2.0 * tau / beta - 1.0
    pos = np.arange(nw_pos, dtype=np.int64)
    wsample = np.concatenate([-pos[::-1] - 1, pos])
    wsample_b = np.arange(-nw_pos, nw_pos + 1, dtype=np.int64)
    nb = ni - 1
    Ttc = cheb_T(ni, x)                     # (nts, ni)
    fac = np.full(ni, 2.0 / ni)
    fac[0] = 1.0 / ni
    Tct = (Ttc[1:-1, :] * fac[None, :]).T   # (ni, ni) on the interior nodes
    wf = (2 * wsample + 1) * np.pi / beta
    wb = 2 * wsample_b * np.pi / beta
    Tnc = cheb_fourier(ni, wf, beta)        # (nw, ni)
    Tnc_b = cheb_fourier(nb, wb, beta)      # (nw_b, nb)
    Tnt = Tnc @ Tct
    Ttn = Ttc @ np.linalg.pinv(Tnc, rcond=1e-13)
    Tnt_BF = Tnc_b @ Tct[:nb, :]
    Ttn_FB = cheb_T(nb, x) @ np.linalg.pinv(Tnc_b, rcond=1e-13)
    return Transforms(beta, ni, wsample.size, wsample_b.size, tau, wsample, wsample_b, Tnt, Ttn, Tnt_BF, Ttn_FB,
                      "chebyshev", Tct=Tct)
