"""The library's IR basis generator (symgw_ir_basis) against the reference's
IR grid file (SymmetryMBPT/data/ir_grid/1e3_72.h5, converted to
tests/golden/ir_1e3_72.npz), and the generated Lambda = 1e4 grid of the
6x6x6 benchmark system against closed forms.  CPU only."""
import numpy as np
import pytest

from tests import helpers as H


@pytest.fixture(scope="module")
def grid_1e3():
    from mbsymmetry_amd.transforms import ir_grid
    return ir_grid(1e3, 72, 71)


def test_generator_reproduces_reference_grid(grid_1e3, ir_npz):
    """Sampling points, u_l at the file's tau points, u^_l at the file's
    Matsubara points, u_l(+-1) and the cross-statistics tables of the
    reference's Lambda = 1e3 grid (72 fermionic, 71 bosonic functions)."""
    from mbsymmetry_amd import _lib
    from mbsymmetry_amd.transforms import ir_eval
    g, z = grid_1e3, ir_npz
    for p, nl in (("fermi", 72), ("bose", 71)):
        assert np.abs(g[p + "__xsample"] - z[p + "__xsample"]).max() < 1e-12, p
        # every Matsubara point but the outermost pair: there the extremum of
        # u^_{L-1} is so flat (|f(n)| changes by ~1e-5 over +-5 indices) that
        # the file's generator, which evaluates the tail differently, lands a
        # few indices away (1738 vs 1743 fermionic, 1779 vs 1780 bosonic)
        w, wz = g[p + "__wsample"], z[p + "__wsample"]
        assert len(w) == nl and np.array_equal(w[1:-1], wz[1:-1]), p
        assert abs(int(w[-1]) - int(wz[-1])) <= 8 and abs(int(w[0]) - int(wz[0])) <= 8, p
    bf = _lib.ir_basis(1, 1e3, 72)
    bb = _lib.ir_basis(0, 1e3, 71)
    for b, p, other in ((bf, "fermi", "bose"), (bb, "bose", "fermi")):
        assert np.abs(ir_eval(b, z[p + "__xsample"]) - z[p + "__uxl"]).max() < 1e-10, p
        uw = z[p + "__uwl_re"] + 1j * z[p + "__uwl_im"]
        assert np.abs(_lib.ir_uhat(b, z[p + "__wsample"]) - uw).max() < 1e-12, p
        assert np.abs(ir_eval(b, [1.0])[0] - z[p + "__ux1l"]).max() < 1e-10, p
        assert np.abs(ir_eval(b, [-1.0])[0] - z[p + "__ux1l_minus"]).max() < 1e-10, p
        assert np.abs(ir_eval(b, z[other + "__xsample"]) - z[p + "__other_uxl"]).max() < 1e-10, p


def test_generated_functions_are_orthonormal_with_parity():
    """int u_l u_m dx = delta_lm (Gauss-Legendre on the basis's own
    segments) and u_l(-x) = (-1)^l u_l(x): properties of the singular
    functions that do not depend on the file."""
    from mbsymmetry_amd import _lib
    from mbsymmetry_amd.transforms import ir_eval
    b = _lib.ir_basis(1, 1e2, 24)
    e = b["edges"][:b["nseg"] + 1]
    xg, wg = np.polynomial.legendre.leggauss(40)
    xs = np.concatenate([0.5 * (e[i] + e[i + 1]) + 0.5 * (e[i + 1] - e[i]) * xg for i in range(b["nseg"])])
    ws = np.concatenate([0.5 * (e[i + 1] - e[i]) * wg for i in range(b["nseg"])])
    U = ir_eval(b, xs)
    assert np.abs(U.T @ (ws[:, None] * U) - np.eye(24)).max() < 1e-12
    par = (-1.0) ** np.arange(24)
    assert np.abs(ir_eval(b, -xs[:50]) - ir_eval(b, xs[:50]) * par[None, :]).max() < 1e-12
    assert np.all(np.diff(b["s"]) < 0)


@pytest.fixture(scope="module")
def tr_config2():
    from mbsymmetry_amd.transforms import ir
    return ir(100.0, 1e4, 104)


def test_config2_grid_size_from_the_cutoff():
    """104 fermionic singular values of the Lambda = 1e4 kernel lie above
    1e-15 s_0: the 104-point grid the 6x6x6 benchmark names is the full basis
    at double precision."""
    from mbsymmetry_amd.transforms import ir_grid
    s = ir_grid(1e4, 104)["fermi__s"]
    assert s[-1] / s[0] > 1e-15
    from mbsymmetry_amd import _lib
    s105 = _lib.ir_basis(1, 1e4, 106)["s"]
    assert s105[104] / s105[0] < 1e-15


def test_config2_grid_single_pole(tr_config2):
    """Fermionic and bosonic single poles across the energy window the grid
    resolves (w_max = Lambda / beta = 100 Ha): tau <-> i w_n round trips, the
    cross-statistics matrices of the W step and the sum rule."""
    tr = tr_config2
    assert tr.nts == 106 and tr.nw == 104 and tr.nw_b == 103
    for eps in (-30.0, -2.0, -0.3, 0.0, 0.7, 3.0, 30.0):
        gt = H.fermi_g([eps], tr.tau, tr.beta)[:, 0]
        gw = 1.0 / (tr.iwn - eps)
        assert np.abs(tr.Tnt @ gt[1:-1] - gw).max() < 1e-10, eps
        assert np.abs(tr.Ttn @ gw - gt).max() < 1e-10, eps
        assert abs(gt[0] + gt[-1] + 1.0) < 1e-14
    for eps in (0.05, 0.5, 4.0, 40.0):
        # bosonic propagator on the fermionic tau points: -e^{-eps t} / (1 - e^{-beta eps})
        bt = -np.exp(-eps * tr.tau) / (1.0 - np.exp(-tr.beta * eps))
        bw = 1.0 / (tr.iwb - eps)
        assert np.abs(tr.Tnt_BF @ bt[1:-1] - bw).max() < 1e-9 * np.abs(bw).max(), eps
        assert np.abs(tr.Ttn_FB @ bw - bt).max() < 1e-9 * np.abs(bt).max(), eps


def test_leakage_flags_unresolved_functions(tr_config2):
    """The Dyson leakage of sc_type_block.cpp:66-86: tiny for poles inside the
    grid's window, large for a pole far outside it (a grid with
    w_max = 10 Ha and a pole at 400 Ha)."""
    from mbsymmetry_amd.transforms import ir, leakage
    tr = tr_config2
    g = H.fermi_g([-1.0, 0.2, 0.9], tr.tau, tr.beta)
    assert leakage(tr, g) < 1e-10
    small = ir(100.0, 1e3, 72)
    g = H.fermi_g([400.0], small.tau, small.beta)
    assert leakage(small, g) > 1e-6
