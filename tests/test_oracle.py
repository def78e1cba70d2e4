"""Pins of the FP64 CPU oracle (no GPU): brute-force index loops, closed
forms, the reference's block == dense relations and its fixture tables."""
import numpy as np
import pytest

from mbsymmetry_amd import system as S
from mbsymmetry_amd.transforms import chebyshev
from oracle.oracle import Oracle
from tests import helpers as H


@pytest.fixture(scope="module")
def tiny_rot():
    # 2x2x2 mesh with rotated stars, ns = 2
    return S.tiny(ns=2, n=2, ink=4, ao=(3, 2), aux=(4, 3), ni=8, beta=10.0, seed=11)


@pytest.fixture(scope="module")
def tiny_big():
    # 3x3x3 mesh (non-TRIM k points: k+q != k-q), ns = 1
    return S.tiny(ns=1, n=3, ink=4, ao=(2, 3, 1), aux=(3, 2), ni=6, beta=8.0, seed=5)


@pytest.fixture(scope="module")
def tiny_tr():
    # 3x3 mesh with mirror + time reversal: conjugate partners, ns = 2
    return S.time_reversal_square(ns=2, ao=(3, 2), aux=(4, 3), ni=8, beta=10.0, seed=13)


def _G_full(orc, sysm, seed=3):
    rng = np.random.default_rng(seed)
    G = rng.standard_normal((sysm.nts, sysm.ns, sysm.flat_ao)) + 1j * rng.standard_normal((sysm.nts, sysm.ns, sysm.flat_ao))
    return G, orc.rotate_G(G)


@pytest.mark.parametrize("name", ["tiny_rot", "tiny_big", "tiny_tr"])
def test_P0_matches_bruteforce(name, request):
    sysm = request.getfixturevalue(name)
    orc = Oracle(sysm)
    V = sysm.host_V()
    _, Gf = _G_full(orc, sysm)
    ts = list(range(sysm.nts // 2))
    ref = H.brute_P0(sysm, V, Gf, ts)
    got = orc.P0(V, Gf)
    assert np.abs(got - ref).max() < 1e-11 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("name", ["tiny_rot", "tiny_big", "tiny_tr"])
def test_Sigma_matches_bruteforce(name, request):
    sysm = request.getfixturevalue(name)
    orc = Oracle(sysm)
    V = sysm.host_V()
    _, Gf = _G_full(orc, sysm)
    rng = np.random.default_rng(9)
    Pi = rng.standard_normal(sysm.nts * sysm.flat_aux) + 1j * rng.standard_normal(sysm.nts * sysm.flat_aux)
    Pf = orc.rotate_P(Pi)
    ts = list(range(sysm.nts))
    ref = H.brute_Sigma(sysm, V, Gf, Pf, ts)
    got = orc.Sigma(V, Gf, Pf)
    assert np.abs(got - ref).max() < 1e-11 * max(1.0, np.abs(ref).max())


def test_rotation_is_O_G_OH(tiny_rot):
    sysm = tiny_rot
    orc = Oracle(sysm)
    G, Gf = _G_full(orc, sysm)
    koff = H.kao_off_full(sysm)
    koffi = sysm.kao_off_irre()
    ro = 0
    for k in range(sysm.nk):
        ik = int(sysm.irre_index[k])
        for (a1, a2) in sysm.orep_irreps[k]:
            m, n = sysm.ao_sizes[k][a1], sysm.ao_sizes[int(sysm.irre_list[ik])][a2]
            O = sysm.orep[ro:ro + m * n].reshape(m, n)
            ro += m * n
            src = koffi[ik] + sysm.ao_block_offsets(int(sysm.irre_list[ik]))[a2]
            dst = koff[k] + sysm.ao_block_offsets(k)[a1]
            for t in (0, 3):
                X = G[t, 1, src:src + n * n].reshape(n, n)
                Y = Gf[t, 1, dst:dst + m * m].reshape(m, m)
                assert np.abs(Y - O @ X @ O.conj().T).max() < 1e-13
        if sysm.index[k] == k:
            sl = slice(koffi[ik], koffi[ik] + sysm.kslice(k))
            assert np.array_equal(Gf[:, :, koff[k]:koff[k] + sysm.kslice(k)], G[:, :, sl])


def _expected_full(sysm, X_ibz, lead, sizes, block_offs, slice_len, off_irre, reps, rep_irreps, off_full):
    """Full-BZ slices from the definition (cugw_routines_block.cu:124-186,
    cugw_qkpt_block.cu:288-340): the TR partner ck of k is copied from its
    iBZ slice when it is a representative, else rotated block by block
    (O X O^H); the slice of k is the complex conjugate when k != ck."""
    out = np.zeros(lead + (off_full[-1] + slice_len(sysm.nk - 1),), dtype=np.complex128)
    rep_base = np.cumsum([0] + [sum(sizes(k)[a1] * sizes(int(sysm.index[k]))[a2] for a1, a2 in rep_irreps[k])
                                for k in range(sysm.nk)])
    for k in range(sysm.nk):
        ck = int(sysm.conj_index[k])
        ik = int(sysm.irre_index[ck])
        kr = int(sysm.irre_list[ik])
        src = X_ibz[..., off_irre[ik]:off_irre[ik] + slice_len(kr)]
        if int(sysm.index[ck]) == ck:
            Y = src.copy()
        else:
            Y = np.zeros(lead + (slice_len(ck),), dtype=np.complex128)
            ro = rep_base[ck]
            for a1, a2 in rep_irreps[ck]:
                m, n = sizes(ck)[a1], sizes(kr)[a2]
                O = reps[ro:ro + m * n].reshape(m, n)
                ro += m * n
                so, do = block_offs(kr)[a2], block_offs(ck)[a1]
                Xb = src[..., so:so + n * n].reshape(lead + (n, n))
                Y[..., do:do + m * m] = (O @ Xb @ O.conj().T).reshape(lead + (m * m,))
        if ck != k:
            Y = Y.conj()
        out[..., off_full[k]:off_full[k] + slice_len(k)] = Y
    return out


def test_rotation_time_reversal(tiny_tr):
    """Full-BZ G and P of a system with rotations and time-reversal partners
    against the definition, slice by slice; the system covers all four ways
    a slice is derived."""
    sysm = tiny_tr
    cases = set()
    for k in range(sysm.nk):
        ck = int(sysm.conj_index[k])
        cases.add((ck != k, int(sysm.index[ck]) != ck))
    assert cases == {(False, False), (False, True), (True, False), (True, True)}
    orc = Oracle(sysm)
    G, Gf = _G_full(orc, sysm)
    ref = _expected_full(sysm, G, (sysm.nts, sysm.ns), lambda k: sysm.ao_sizes[k], sysm.ao_block_offsets,
                         sysm.kslice, sysm.kao_off_irre(), sysm.orep, sysm.orep_irreps, H.kao_off_full(sysm))
    assert np.abs(Gf - ref).max() < 1e-13
    rng = np.random.default_rng(4)
    nts = sysm.nts
    Pi = rng.standard_normal(nts * sysm.flat_aux) + 1j * rng.standard_normal(nts * sysm.flat_aux)
    Pf = orc.rotate_P(Pi)
    # q_t_QP: [q][t][slice] -> per-q views with a leading t axis
    qoi, qof = sysm.qaux_off_irre(), H.qaux_off_full(sysm)
    Pi_t = np.concatenate([Pi[nts * qoi[i]:nts * (qoi[i] + sysm.qslice(int(q)))].reshape(nts, -1)
                           for i, q in enumerate(sysm.irre_list)], axis=1)
    Pf_t = np.concatenate([Pf[nts * qof[q]:nts * (qof[q] + sysm.qslice(q))].reshape(nts, -1)
                           for q in range(sysm.nk)], axis=1)
    ref = _expected_full(sysm, Pi_t, (nts,), lambda q: sysm.aux_sizes[q], sysm.aux_block_offsets, sysm.qslice,
                         qoi, sysm.auxrep, sysm.auxrep_irreps, qof)
    assert np.abs(Pf_t - ref).max() < 1e-13


def test_P0_lindhard_closed_form(tiny_big):
    """Non-interacting G = U g U^H: P0_QP(t) = c sum_k sum_ab g_a(beta-t) g_b(t) Vt_Qab conj(Vt_Pab),
    Vt_Q = U_k^H V_Q U_{k+q} (eigenbasis contraction, no GEMM chain)."""
    sysm = tiny_big
    orc = Oracle(sysm)
    V = sysm.host_V()
    G, eig = H.noninteracting_G_ibz(sysm)
    Gf = orc.rotate_G(G)
    got = orc.P0(V, Gf)
    # eigenbasis of every full-BZ block from the rotated closed form
    beta, tau, nts, nk = sysm.transforms.beta, sysm.transforms.tau, sysm.nts, sysm.nk
    koff = H.kao_off_full(sysm)
    ref = sysm.zeros_q_t_QP()
    qoffi = sysm.qaux_off_irre()
    offs = sysm.kpair_offsets()
    for iqr, q in enumerate(sysm.irre_list):
        q = int(q)
        for k in range(nk):
            k2 = int(sysm.kpq[k, q])
            p = k * nk + k2
            o = offs[p][0]
            for (iq, i1, i2), (qs, a1, a2) in zip(sysm.V_irreps[p], sysm.V_block_sizes(p)):
                Vb = V[o:o + qs * a1 * a2].reshape(qs, a1, a2)
                o += qs * a1 * a2
                # eigen-decompose the full-BZ blocks (G(t=0+) determines U, eigenvalues from F)
                def eigblock(kk, ib, sz):
                    ikr = int(sysm.irre_index[kk])
                    e, U0 = eig[(0, ikr, sysm.ao_block_offsets(int(sysm.irre_list[ikr]))[ib])]
                    if sysm.index[kk] == kk:
                        return e, U0
                    # rotated block: O U0
                    ro = 0
                    for kk2 in range(kk):
                        ro += sum(sysm.ao_sizes[kk2][x] * sysm.ao_sizes[int(sysm.irre_list[int(sysm.irre_index[kk2])])][y]
                                  for x, y in sysm.orep_irreps[kk2])
                    for (x, y) in sysm.orep_irreps[kk]:
                        m = sysm.ao_sizes[kk][x]; n = sysm.ao_sizes[int(sysm.irre_list[ikr])][y]
                        if x == ib:
                            return e, sysm.orep[ro:ro + m * n].reshape(m, n) @ U0
                        ro += m * n
                    raise AssertionError
                e1, U1 = eigblock(k, i1, a1)
                e2, U2 = eigblock(k2, i2, a2)
                Vt = np.einsum("pa,Qpm,mb->Qab", U1.conj(), Vb, U2)
                for t in range(nts // 2):
                    g1 = H.fermi_g(e1, [beta - tau[t]], beta)[0]
                    g2 = H.fermi_g(e2, [tau[t]], beta)[0]
                    blk = np.einsum("a,b,Qab,Pab->QP", g1, g2, Vt, Vt.conj())
                    base = nts * qoffi[iqr] + t * sysm.qslice(q) + sysm.aux_block_offsets(q)[iq]
                    ref[base:base + qs * qs] += blk.ravel()
    ref *= -(2.0 / sysm.ns) / nk
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()


def test_dyson_sigma0_closed_form(config1):
    """Sigma = 0: G(tau) = -U diag(exp(-(e - mu) tau)/(1 + exp(-beta (e - mu)))) U^H (S = I)."""
    orc = Oracle(config1)
    G = orc.dyson(config1.zeros_ts_kij(), 0.1)
    ref, _ = H.noninteracting_G_ibz(config1, mu=0.1)
    assert np.abs(G - ref).max() < 5e-10


def test_dyson_lih_block_equals_dense(lih, lih_npz):
    """Dyson.BlockDyson relation (test_dyson.cpp:17-83): block Dyson of input_flat
    equals the dense Dyson of input_blk at 1e-10."""
    orc = Oracle(lih)
    G = orc.dyson(lih.zeros_ts_kij(), 0.0)
    tr = lih.transforms
    Fd = lih_npz["blk__Fock-k"][..., 0] + 1j * lih_npz["blk__Fock-k"][..., 1]
    Sd = lih_npz["blk__S-k"][..., 0] + 1j * lih_npz["blk__S-k"][..., 1]
    ns, nk, nao = Fd.shape[0], Fd.shape[1], Fd.shape[2]
    koffi = lih.kao_off_irre()
    for s in range(ns):
        for k in range(nk):
            F = 0.5 * (Fd[s, k] + Fd[s, k].conj().T)
            Sm = 0.5 * (Sd[s, k] + Sd[s, k].conj().T)
            Gw = np.array([np.linalg.inv(iw * Sm - F) for iw in tr.iwn])
            Gt = np.einsum("tw,wij->tij", tr.Ttn, Gw)
            # compare every diagonal block
            boffs = lih_npz[f"sym__block__ao_block_offsets__{k}"]
            for ib, (o, b) in enumerate(zip(lih.ao_block_offsets(k), lih.ao_sizes[k])):
                bo = int(boffs[ib])
                got = G[:, s, koffi[k] + o:koffi[k] + o + b * b].reshape(-1, b, b)
                assert np.abs(got - Gt[:, bo:bo + b, bo:bo + b]).max() < 1e-10
            # and the off-block-diagonal part of the dense result vanishes
            mask = np.ones((nao, nao), bool)
            for ib, b in enumerate(lih.ao_sizes[k]):
                bo = int(boffs[ib]); mask[bo:bo + b, bo:bo + b] = False
            assert np.abs(Gt[:, mask]).max() < 1e-10


def test_W_block_equals_dense(lih):
    """CuGWP relation (test_cugw.cpp:210-304): block W equals the dense
    gw_solver::compute_P_from_P0 (gw.cpp:129-160) at 1e-9, on a Hermitian P0(t)."""
    sysm = lih
    orc = Oracle(sysm)
    G = orc.dyson(sysm.zeros_ts_kij(), 0.0)
    Gf = orc.rotate_G(G)
    P0 = orc.P0(sysm.host_V(), Gf)
    P = orc.W(P0)
    tr, nts = sysm.transforms, sysm.nts
    qoffi = sysm.qaux_off_irre()
    for iqr, q in enumerate(sysm.irre_list):
        q = int(q)
        qs = sysm.qslice(q)
        naux = sum(sysm.aux_sizes[q])
        blocks = list(zip(sysm.aux_block_offsets(q), sysm.aux_sizes[q]))
        dense = np.zeros((nts, naux, naux), dtype=np.complex128)
        for t in range(nts // 2):
            bo = 0
            for o, b in blocks:
                m = P0[nts * qoffi[iqr] + t * qs + o:nts * qoffi[iqr] + t * qs + o + b * b].reshape(b, b)
                dense[t, bo:bo + b, bo:bo + b] = m
                dense[nts - 1 - t, bo:bo + b, bo:bo + b] = m
                bo += b
        dense = 0.5 * (dense + dense.conj().transpose(0, 2, 1))
        Pw = np.einsum("wt,tij->wij", tr.Tnt_BF, dense[1:-1])
        I = np.eye(naux)
        for n in range(Pw.shape[0]):
            A = np.linalg.inv(I - Pw[n])
            A = 0.5 * (A + A.conj().T)
            Pw[n] = A @ Pw[n]
        Pt = np.einsum("tw,wij->tij", tr.Ttn_FB, Pw)
        Pt = 0.5 * (Pt + Pt.conj().transpose(0, 2, 1))
        for t in range(nts):
            bo = 0
            for o, b in blocks:
                got = P[nts * qoffi[iqr] + t * qs + o:nts * qoffi[iqr] + t * qs + o + b * b].reshape(b, b)
                assert np.abs(got - Pt[t, bo:bo + b, bo:bo + b]).max() < 1e-9
                bo += b


def test_inverses_against_numpy():
    rng = np.random.default_rng(1)
    orc = Oracle(S.tiny())
    for n in (1, 2, 7, 19):
        A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        assert np.abs(orc.inverse_lu(A) - np.linalg.inv(A)).max() < 1e-11
        Hm = A @ A.conj().T + n * np.eye(n)
        assert np.abs(orc.inverse_chol(Hm) - np.linalg.inv(Hm)).max() < 1e-12


def test_density_fermi_sum(config1):
    """Sigma = 0: N = (2/ns) sum_k w_k sum_a f(e_a - mu) / nk (Fermi-Dirac)."""
    orc = Oracle(config1)
    mu = 0.05
    G, eig = H.noninteracting_G_ibz(config1, mu=mu)
    beta = config1.transforms.beta
    ref = 0.0
    for (s, ikr, o), (e, U) in eig.items():
        ref += config1.kweights[ikr] * np.sum(1.0 / (np.exp(beta * (e - mu)) + 1.0))
    ref *= (2.0 / config1.ns) / config1.nk
    assert abs(orc.density(G) - ref) < 1e-12


def test_lih_table_pins(lih, lih_npz):
    """Derived offsets equal the tables symmetry_info.h5 stores."""
    z = lih_npz
    orc = Oracle(lih)
    sz = orc.flat_sizes()
    assert sz["flat_ao"] == 328 and sz["flat_aux"] == 7456 and sz["flat_V"] == 206688
    koff = H.kao_off_full(lih)
    qoff = H.qaux_off_full(lih)
    assert list(z["sym__block__kao_slice_offsets"]) == koff
    assert list(z["sym__block__qaux_slice_offsets"]) == qoff
    assert list(z["sym__block__kao_slice_sizes"]) == [lih.kslice(k) for k in range(8)]
    for k in range(8):
        assert list(z[f"sym__block__ao_offsets__{k}"]) == lih.ao_block_offsets(k)
        assert list(z[f"sym__block__aux_offsets__{k}"]) == lih.aux_block_offsets(k)
    offs = lih.kpair_offsets()
    for p in range(64):
        assert z["sym__block__kpair_slice_offsets"][p] == offs[p][0]
        assert z["sym__block__kpair_slice_sizes"][p] == offs[p][1]
        assert orc.kpair_offset(p // 8, p % 8) == offs[p][0]
        toff = np.concatenate([[0], np.cumsum([a * b * c for a, b, c in lih.V_block_sizes(p)])[:-1]])
        assert list(z[f"sym__block__tensor_offsets__{p}"]) == list(toff)


def test_transforms_single_pole(ir_npz):
    from mbsymmetry_amd.transforms import ir_from_fixture
    for tr, tol in ((ir_from_fixture(ir_npz, 5.0), 2e-12), (chebyshev(60, 50.0), 5e-11)):
        for eps in (-1.0, -0.2, 0.0, 0.3, 1.0):
            gt = H.fermi_g([eps], tr.tau, tr.beta)[:, 0]
            gw = 1.0 / (tr.iwn - eps)
            assert np.abs(tr.Ttn @ gw - gt).max() < tol
            assert np.abs(tr.Tnt @ gt[1:-1] - gw).max() < 1e-12
        # sum rule G(0+) + G(beta-) = -1 for S = 1
        gt = H.fermi_g([0.37], tr.tau, tr.beta)[:, 0]
        assert abs(gt[0] + gt[-1] + 1.0) < 1e-14


def test_ir_fixture_tau_mesh(ir_npz):
    from mbsymmetry_amd.transforms import ir_from_fixture
    tr = ir_from_fixture(ir_npz, 5.0)
    assert tr.nts == 74 and tr.nw == 72 and tr.nw_b == 71
    assert tr.tau[0] == 0.0 and tr.tau[-1] == 5.0 and np.all(np.diff(tr.tau) > 0)


@pytest.mark.parametrize("name", ["lih", "config1"])
def test_oracle_matches_golden(name, request):
    """The oracle's GW iteration still reproduces the committed regression
    vectors (tests/golden/make_oracle_golden.py)."""
    import json
    import os
    from tests.golden.make_oracle_golden import iterate, projections
    with open(os.path.join(os.path.dirname(__file__), "golden", "oracle_golden.json")) as f:
        gold = json.load(f)["systems"][name]
    res = iterate(request.getfixturevalue(name))
    for i, (k, v) in enumerate(sorted(res.items())):
        got = projections(v, seed=1000 + i)
        ref = gold[k]
        tol = 1e-11 * ref["norm"]
        assert abs(got["norm"] - ref["norm"]) < tol, (name, k)
        assert np.abs(np.array(got["proj"]) - np.array(ref["proj"])).max() < tol * 10, (name, k)


@pytest.mark.parametrize("name", ["tiny_rot", "tiny_tr"])
def test_sigma_blocks_hermitian(name, request):
    """With G(t) from a Dyson solve and P(t) from the W solve (both Hermitian
    per block), Sigma(t) = -sum_{QP} V_Q G P_QP V_P^H is Hermitian per ao block
    (the hermiticity pin of the Sigma contraction, gw_block.cpp:86-158)."""
    sysm = request.getfixturevalue(name)
    orc = Oracle(sysm)
    G0 = orc.dyson(sysm.zeros_ts_kij(), 0.0)
    assert H.sigma_hermiticity_error(sysm, G0) < 1e-13
    _, _, Sig, _ = orc.gw_iteration(G0, 0.0)
    assert np.abs(Sig).max() > 0
    assert H.sigma_hermiticity_error(sysm, Sig) < 1e-12


@pytest.mark.parametrize("name", ["tiny_rot", "tiny_tr"])
def test_sampled_slices_equal_full_sweeps(name, request):
    """orc_P0_at / orc_Sigma_at / orc_rotate_P_t (one output slice summed
    over the whole mesh, used for the parity of systems too large for full
    CPU sweeps) give the same numbers as the full sweeps."""
    sysm = request.getfixturevalue(name)
    orc = Oracle(sysm)
    G0 = orc.dyson(sysm.zeros_ts_kij(), 0.0)
    V = sysm.host_V()
    Gf = orc.rotate_G(G0)
    P0 = orc.P0(V, Gf)
    P = orc.W(P0)
    Pf = orc.rotate_P(P)
    Sg = orc.Sigma(V, Gf, Pf)
    offs = {p: o for p, (o, sz) in sysm.kpair_offsets().items()}
    qoff, koff = sysm.qaux_off_irre(), sysm.kao_off_irre()
    nfull = orc.flat_sizes()["flat_aux_full"]
    for iqr in range(sysm.ink):
        qs = sysm.qslice(int(sysm.irre_list[iqr]))
        for t in (0, sysm.nts // 2 - 1):
            base = sysm.nts * qoff[iqr] + t * qs
            assert np.abs(orc.P0_at(V, offs, Gf, iqr, t, nthreads=3) - P0[base:base + qs]).max() < 1e-13
    for t in (0, sysm.nts - 1):
        Pt = orc.rotate_P_t(P, t)
        for q in range(sysm.nk):
            o = sum(sysm.qslice(qq) for qq in range(q))
            ref = Pf[sysm.nts * o + t * sysm.qslice(q): sysm.nts * o + (t + 1) * sysm.qslice(q)]
            assert np.abs(Pt[o:o + sysm.qslice(q)] - ref).max() < 1e-15
        assert Pt.size == nfull
        for ikr in range(sysm.ink):
            ks = sysm.kslice(int(sysm.irre_list[ikr]))
            for sp in range(sysm.ns):
                got = orc.Sigma_at(V, offs, Gf, Pt, ikr, t, sp, nthreads=4)
                assert np.abs(got - Sg[t, sp, koff[ikr]:koff[ikr] + ks]).max() < 1e-13
