"""GPU parity of the HIP path against the FP64 CPU oracle, through the C ABI.

Tolerances follow the reference's own tests: P0 and Sigma 1e-10
(test_cugw.cpp:98, :203), P (W) 1e-9 (:301), Dyson 1e-10 (test_dyson.cpp:82).
"""
import numpy as np
import pytest

from mbsymmetry_amd import _lib
from mbsymmetry_amd import system as S
from tests import helpers as H
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def systems(lih, config1):
    return {
        "tiny_rot_ns2": S.tiny(ns=2, n=2, ink=4, ao=(3, 2), aux=(4, 3), ni=8, beta=10.0, seed=11),
        "tiny_3cube": S.tiny(ns=1, n=3, ink=4, ao=(2, 3, 1), aux=(3, 2), ni=6, beta=8.0, seed=5),
        "tr_square": S.time_reversal_square(ns=2, ao=(5, 3, 2), aux=(12, 8, 6), ni=10, beta=12.0, seed=13),
        # rotations with permuted block maps (a1 != a2), k-dependent blocks
        "permuted": S.permuted_blocks(n=3, ns=2),
        "config1": config1,
        "lih": lih,
    }


def _setup(sysm, precision=_lib.PREC_FP64, mu=0.0, device_V=False):
    orc = Oracle(sysm)
    ctx = _lib.Context(sysm, 0, precision)
    if device_V:
        ctx.generate_V(sysm.v_seed, sysm.v_scale, sysm.decay_table())
    else:
        ctx.set_V(sysm.host_V())
    ctx.set_reps(sysm.orep, sysm.auxrep)
    ctx.set_transforms(sysm.transforms)
    ctx.set_fock(sysm.F, sysm.S)
    G0 = orc.dyson(sysm.zeros_ts_kij(), mu)
    ctx.upload_G(G0)
    return orc, ctx, G0


def _rel(a, b):
    return np.abs(a - b).max() / max(1.0, np.abs(b).max())


@pytest.mark.parametrize("name", ["tiny_rot_ns2", "tiny_3cube", "tr_square", "permuted", "config1", "lih"])
def test_P0_parity(name, systems):
    sysm = systems[name]
    _say("P0", name)
    orc, ctx, G0 = _setup(sysm)
    ctx.solve_P0()
    got = ctx.copy_P0_to_host(sysm.zeros_q_t_QP())
    ref = orc.P0(sysm.host_V(), orc.rotate_G(G0))
    nts, h = sysm.nts, sysm.nts // 2
    got_s = got / sysm.nk
    qoff = sysm.qaux_off_irre()
    for iqr, q in enumerate(sysm.irre_list):
        qs = sysm.qslice(int(q))
        base = nts * qoff[iqr]
        a = got_s[base:base + h * qs]
        b = ref[base:base + h * qs]
        assert np.abs(a - b).max() < 1e-10 * max(1.0, np.abs(b).max()), (name, iqr)
        # rows t >= nts/2 are zero, as in the reference's device buffer
        assert np.all(got[base + h * qs:base + nts * qs] == 0)
    ctx.close()


@pytest.mark.parametrize("name", ["tiny_rot_ns2", "tiny_3cube", "tr_square", "permuted", "config1", "lih"])
def test_W_and_Sigma_parity(name, systems):
    sysm = systems[name]
    _say("W+Sigma", name)
    orc, ctx, G0 = _setup(sysm)
    V = sysm.host_V()
    Gf = orc.rotate_G(G0)
    P0 = orc.P0(V, Gf)
    Pref = orc.W(P0)
    ctx.solve_P0()
    ctx.solve_W()
    P = ctx.copy_P0_to_host(sysm.zeros_q_t_QP())
    assert _rel(P, Pref) < 1e-9, name
    # Sigma from the oracle's P (isolates the Sigma kernel)
    ctx.upload_P(Pref)
    ctx.solve_Sigma()
    Sg = ctx.copy_Sigma_to_host(sysm.zeros_ts_kij()) / sysm.nk
    Sref = orc.Sigma(V, Gf, orc.rotate_P(Pref))
    assert _rel(Sg, Sref) < 1e-10, name
    ctx.close()


@pytest.mark.parametrize("name", ["tiny_rot_ns2", "tr_square", "permuted", "config1", "lih"])
def test_dyson_parity(name, systems):
    sysm = systems[name]
    _say("Dyson", name)
    orc, ctx, G0 = _setup(sysm, mu=0.05)
    # Sigma = 0 path (G0 itself)
    assert _rel(ctx.download_G(G0.shape), G0) < 1e-13
    ctx.dyson(0.05, 0.0)
    assert _rel(ctx.download_G(G0.shape), G0) < 1e-10
    rng = np.random.default_rng(4)
    Sig = 0.05 * (rng.standard_normal(G0.shape) + 1j * rng.standard_normal(G0.shape))
    ctx.upload_Sigma(Sig)
    ctx.dyson(0.05, 1.0)
    assert _rel(ctx.download_G(G0.shape), orc.dyson(Sig, 0.05)) < 1e-10, name
    ctx.close()


@pytest.mark.parametrize("name", ["tiny_rot_ns2", "tr_square", "permuted", "config1"])
def test_gw_iteration_parity_and_graph(name, systems):
    sysm = systems[name]
    _say("iteration", name)
    mu = 0.0
    orc, ctx, G0 = _setup(sysm, mu=mu)
    ctx.gw_iteration(mu, use_graph=False)
    ctx.sync()
    G1 = ctx.download_G(G0.shape)
    _, _, _, G1ref = orc.gw_iteration(G0, mu)
    assert _rel(G1, G1ref) < 1e-10, name
    # graph replay is bitwise identical to the eager sequence
    ctx2 = _lib.Context(sysm)
    ctx2.set_V(sysm.host_V()); ctx2.set_reps(sysm.orep, sysm.auxrep)
    ctx2.set_transforms(sysm.transforms); ctx2.set_fock(sysm.F, sysm.S); ctx2.upload_G(G0)
    eager = []
    for _ in range(3):
        ctx.gw_iteration(mu, use_graph=False); ctx.sync(); eager.append(ctx.download_G(G0.shape))
    ctx2.gw_iteration(mu, use_graph=False); ctx2.sync()
    for i in range(3):
        ctx2.gw_iteration(mu, use_graph=True); ctx2.sync()
        assert np.array_equal(ctx2.download_G(G0.shape), eager[i]), (name, i)
    ctx.close(); ctx2.close()


def test_device_generator_matches_host_bits(systems):
    sysm = systems["config1"]
    _say("generator", "config1")
    _, c_host, _ = _setup(sysm)
    _, c_dev, _ = _setup(sysm, device_V=True)
    c_host.solve_P0(); c_dev.solve_P0()
    a = c_host.copy_P0_to_host(sysm.zeros_q_t_QP())
    b = c_dev.copy_P0_to_host(sysm.zeros_q_t_QP())
    assert np.array_equal(a, b)
    c_host.close(); c_dev.close()


def test_contractions_are_deterministic(systems):
    sysm = systems["config1"]
    _say("determinism", "config1")
    _, ctx, G0 = _setup(sysm)
    outs = []
    for _ in range(2):
        ctx.solve_P0()
        outs.append(ctx.copy_P0_to_host(sysm.zeros_q_t_QP()))
    assert np.array_equal(outs[0], outs[1])
    ctx.solve_W()
    s = []
    for _ in range(2):
        ctx.solve_Sigma()
        s.append(ctx.copy_Sigma_to_host(sysm.zeros_ts_kij()))
    assert np.array_equal(s[0], s[1])
    ctx.close()


def test_fp32_mode_error_is_bounded(systems):
    sysm = systems["config1"]
    _say("fp32", "config1")
    orc, ctx, G0 = _setup(sysm, precision=_lib.PREC_FP32)
    ctx.solve_P0()
    got = ctx.copy_P0_to_host(sysm.zeros_q_t_QP()) / sysm.nk
    ref = orc.P0(sysm.host_V(), orc.rotate_G(G0))
    h = sysm.nts // 2
    mask = np.zeros_like(ref, dtype=bool)
    qoff = sysm.qaux_off_irre()
    for iqr, q in enumerate(sysm.irre_list):
        qs = sysm.qslice(int(q))
        mask[sysm.nts * qoff[iqr]:sysm.nts * qoff[iqr] + h * qs] = True
    err = np.abs(got[mask] - ref[mask]).max() / np.abs(ref[mask]).max()
    assert err < 1e-5, err
    # Sigma from the oracle's P in FP32 arithmetic
    Pref = orc.W(ref)
    ctx.upload_P(Pref)
    ctx.solve_Sigma()
    Sg = ctx.copy_Sigma_to_host(sysm.zeros_ts_kij()) / sysm.nk
    Sref = orc.Sigma(sysm.host_V(), orc.rotate_G(G0), orc.rotate_P(Pref))
    err = np.abs(Sg - Sref).max() / np.abs(Sref).max()
    assert err < 1e-5, err
    ctx.close()


def test_density_and_mu_search(systems):
    sysm = systems["config1"]
    _say("density", "config1")
    orc, ctx, G0 = _setup(sysm, mu=0.1)
    n_dev = ctx.density(sysm.kweights)
    assert abs(n_dev - orc.density(G0)) < 1e-11
    target = n_dev
    mu, n = ctx.mu_search(target, sysm.kweights, 0.0, -1.0, 1.0, 1e-10)
    assert abs(n - target) < 1e-9 and abs(mu - 0.1) < 1e-6
    ctx.close()


@pytest.mark.parametrize("name", ["lih", "config1"])
def test_iteration_matches_golden(name, systems):
    """Sigma and the new G of one device GW iteration against the committed
    oracle regression vectors (tests/golden/oracle_golden.json)."""
    import json
    import os
    from tests.golden.make_oracle_golden import projections
    sysm = systems[name]
    _say("golden", name)
    with open(os.path.join(os.path.dirname(__file__), "golden", "oracle_golden.json")) as f:
        gold = json.load(f)
    mu = gold["mu"]
    order = sorted(["G1", "P", "P0", "Sigma"])
    _, ctx, G0 = _setup(sysm, mu=mu)
    ctx.gw_iteration(mu, use_graph=True)
    ctx.sync()
    got = {"Sigma": ctx.copy_Sigma_to_host(sysm.zeros_ts_kij()) / sysm.nk, "G1": ctx.download_G(G0.shape)}
    ctx.close()
    for k, v in got.items():
        g = projections(v, seed=1000 + order.index(k))
        ref = gold["systems"][name][k]
        tol = 1e-10 * ref["norm"]
        assert abs(g["norm"] - ref["norm"]) < tol, (name, k)
        assert np.abs(np.array(g["proj"]) - np.array(ref["proj"])).max() < 10 * tol, (name, k)


@pytest.mark.parametrize("name", ["tiny_rot_ns2", "config1"])
def test_self_consistent_loop_matches_oracle(name, systems):
    """Three damped iterations of the device self-consistency driver against
    the same loop on the oracle (G_{n+1} = d GW(G_n) + (1 - d) G_n)."""
    from mbsymmetry_amd.scloop import self_consistent
    sysm = systems[name]
    _say("sc loop", name)
    mu, damp = 0.02, 0.6
    orc, ctx, G0 = _setup(sysm, mu=mu)
    res = self_consistent(ctx, sysm, G0, mu, max_iter=3, tol=0.0, damp=damp, keep_sigma=True)
    G = G0.copy()
    for _ in range(3):
        _, _, Sig, Gn = orc.gw_iteration(G, mu)
        G = damp * Gn + (1 - damp) * G
    assert res.iterations == 3 and not res.converged
    # one iteration agrees to 1e-10 (test_gw_iteration_parity_and_graph); three compound it
    assert _rel(res.G, G) < 1e-9, name
    assert _rel(res.Sigma, Sig) < 1e-9, name
    assert [h["iteration"] for h in res.history] == [1, 2, 3]
    ctx.close()


def test_self_consistent_loop_holds_the_electron_count():
    """With a target electron count the driver re-finds mu after every
    iteration: the count stays at the target and mu matches a bisection on
    the oracle's Dyson solve of the same Sigma."""
    from mbsymmetry_amd.scloop import self_consistent
    from mbsymmetry_amd.transforms import chebyshev
    # the config-1 system with V scaled by 0.05: |Sigma| ~ 0.02 Ha, so N(mu)
    # rises monotonically from 0 to 2 nao across [-2, 2] (with the unscaled V,
    # |Sigma| ~ 400 Ha and N(mu) never reaches the Sigma = 0 count)
    sysm = S.synthetic("config1-weak", 2, 4, [8, 6, 6, 4, 2], [40, 32, 32, 24], 1, chebyshev(60, 50.0),
                       S.SEED_BASE + 1, v_scale=0.05, all_pairs=True)
    _say("sc loop nel", "config1-weak")
    orc, ctx, G0 = _setup(sysm, mu=0.0)
    nel = orc.density(G0)
    res = self_consistent(ctx, sysm, G0, 0.0, max_iter=2, tol=0.0, nel=nel)
    for h in res.history:
        assert abs(h["nel"] - nel) < 1e-8
    # oracle: one iteration at mu = 0, then bisection of the Dyson solve on its Sigma
    _, _, Sig, _ = orc.gw_iteration(G0, 0.0)
    lo, hi = -2.0, 2.0
    for _ in range(80):
        mid = 0.5 * (lo + hi)
        if orc.density(orc.dyson(Sig, mid)) < nel:
            lo = mid
        else:
            hi = mid
    assert abs(res.history[0]["mu"] - 0.5 * (lo + hi)) < 1e-7
    ctx.close()


@pytest.mark.parametrize("ao,aux", [([66, 10], [100, 24]), ([10, 6], [200, 20]), ([100, 10], [64, 24])])
def test_large_blocks_parity(ao, aux):
    """Blocks that exercise every chunking path of the contraction kernels:
    an ao block of 66 (> 48: the wide-block Sigma slabs, one time of one
    80-wide band in FP64; orbital chunks in P0), an ao block of 100 (two times
    of 64- and 36-wide bands: four band pairs per block)
    and an aux block of 100 (row bands in P0, aux-row chunks in Sigma); an aux
    block of 200 (more P0 column tiles than waves: the two times of a slab
    run one after the other).  P0, W and Sigma against the oracle."""
    import os
    from mbsymmetry_amd.transforms import chebyshev
    sysm = S.synthetic("big_blocks", 2, 4, ao, aux, 1, chebyshev(4, 6.0), 31, v_scale=0.1)
    _say("large blocks", "ao", ao, "aux", aux)
    nth = min(16, len(os.sched_getaffinity(0)))
    orc, ctx, G0 = _setup(sysm, mu=0.0)
    V = sysm.host_V()
    Gf = orc.rotate_G(G0)
    P0 = orc.P0(V, Gf, nthreads=nth)
    ctx.solve_P0()
    got = ctx.copy_P0_to_host(sysm.zeros_q_t_QP()) / sysm.nk
    h = sysm.nts // 2
    qoff = sysm.qaux_off_irre()
    for iqr, q in enumerate(sysm.irre_list):
        qs = sysm.qslice(int(q))
        base = sysm.nts * qoff[iqr]
        a, b = got[base:base + h * qs], P0[base:base + h * qs]
        assert np.abs(a - b).max() < 1e-10 * max(1.0, np.abs(b).max()), iqr
    Pref = orc.W(P0)
    ctx.solve_W()
    assert _rel(ctx.copy_P0_to_host(sysm.zeros_q_t_QP()), Pref) < 1e-9
    ctx.upload_P(Pref)
    ctx.solve_Sigma()
    Sg = ctx.copy_Sigma_to_host(sysm.zeros_ts_kij()) / sysm.nk
    Sref = orc.Sigma(V, Gf, orc.rotate_P(Pref), nthreads=nth)
    assert _rel(Sg, Sref) < 1e-10
    ctx.close()


def test_measured_peaks_are_sane():
    """The peak microbenchmarks run and land near the MI355X datasheet: FP64
    matrix 78.6 TFLOP/s, FP32 157.3, HBM 8 TB/s (a copy reaches about two
    thirds of the datasheet)."""
    _say("measure peaks")
    pk = _lib.measure_peaks(0)
    assert 0.6 * 78.6 < pk["fp64_mfma_tflops"] < 1.02 * 78.6, pk
    assert 0.6 * 157.3 < pk["fp32_mfma_tflops"] < 1.02 * 157.3, pk
    assert 3000 < pk["hbm_copy_gbps"] < 8200, pk


@pytest.mark.parametrize("name,tol", [("tr_square", 1e-12), ("config1", 1e-9)])
def test_device_sigma_blocks_hermitian(name, tol, systems):
    """The device Sigma of a physical iteration (G from Dyson, P from the W
    solve) is Hermitian per ao block, to the Hermiticity of its G0 (config 1's
    Chebyshev transforms hold G0 to ~1e-11)."""
    sysm = systems[name]
    _say("sigma hermitian", name)
    _, ctx, G0 = _setup(sysm, mu=0.0)
    ctx.gw_iteration(0.0, use_graph=True)
    ctx.sync()
    Sig = ctx.copy_Sigma_to_host(sysm.zeros_ts_kij())
    ctx.close()
    assert H.sigma_hermiticity_error(sysm, Sig) < tol, name


def test_density_is_monotonic_in_mu():
    """N(mu) of the device Dyson solve rises monotonically with mu (the
    mu-search's bracketing relies on it), with the Sigma of one GW iteration
    of a weakly interacting config-1 system held fixed."""
    from mbsymmetry_amd.transforms import chebyshev
    sysm = S.synthetic("config1-weak", 2, 4, [8, 6, 6, 4, 2], [40, 32, 32, 24], 1, chebyshev(60, 50.0),
                       S.SEED_BASE + 1, v_scale=0.05, all_pairs=True)
    _say("N(mu) monotonic", "config1-weak")
    _, ctx, _ = _setup(sysm, mu=0.0)
    ctx.gw_iteration(0.0, use_graph=True)
    ctx.sync()
    ns = []
    for mu in np.linspace(-1.5, 1.5, 13):
        ctx.dyson(float(mu), 1.0 / sysm.nk)
        ns.append(ctx.density(sysm.kweights))
    ctx.close()
    assert all(b > a for a, b in zip(ns, ns[1:])), ns
    assert 0.0 <= ns[0] < ns[-1] <= 2.0 * sum(sysm.ao_sizes[int(sysm.irre_list[0])]) + 1e-9


def test_uneven_blocks_1_to_256():
    """Irrep blocks of every size from 1 to 256 (ao {1, 3, 9, 27, 81, 256},
    aux {1, 4, 16, 64, 256}) in one launch per contraction, checked by
    size-independent properties (no oracle at this size): P0 and Sigma
    Hermitian per block for G from the Dyson solve, Sigma bitwise identical
    across runs, and the FP32 path within FP32 accuracy of the FP64 one."""
    import time
    from mbsymmetry_amd.transforms import chebyshev
    sysm = S.synthetic("uneven-1-256", 2, 4, [1, 3, 9, 27, 81, 256], [1, 4, 16, 64, 256], 1,
                       chebyshev(20, 20.0), S.SEED_BASE + 3)
    _say("uneven 1-256", "ao", sysm.ao_sizes[0], "aux", sysm.aux_sizes[0])
    out = {}
    for prec in (_lib.PREC_FP64, _lib.PREC_FP32):
        ctx = _lib.Context(sysm, 0, prec)
        ctx.generate_V(sysm.v_seed, sysm.v_scale, sysm.decay_table())
        ctx.set_reps(sysm.orep, sysm.auxrep)
        ctx.set_transforms(sysm.transforms)
        ctx.set_fock(sysm.F, sysm.S)
        ctx.upload_G(sysm.zeros_ts_kij())
        ctx.dyson(0.0, 0.0)
        ctx.sync()
        t0 = time.perf_counter()
        ctx.solve_P0()
        ctx.solve_W()
        ctx.solve_Sigma()
        ctx.sync()
        dt = time.perf_counter() - t0
        inf = ctx.info()
        _say("  prec", prec, "P0+W+Sigma %.3f s, %.1f TFLOP/s" % (dt, (inf.flop_P0 + inf.flop_Sigma) / dt / 1e12))
        P0 = ctx.copy_P0_to_host(sysm.zeros_q_t_QP())
        S1 = ctx.copy_Sigma_to_host(sysm.zeros_ts_kij())
        ctx.solve_Sigma()
        S2 = ctx.copy_Sigma_to_host(sysm.zeros_ts_kij())
        ctx.close()
        assert np.array_equal(S1, S2), prec
        assert np.abs(S1).max() > 0 and np.abs(P0).max() > 0
        out[prec] = (P0, S1)
    P0, Sig = out[_lib.PREC_FP64]
    assert H.p0_hermiticity_error(sysm, P0) < 1e-12
    assert H.sigma_hermiticity_error(sysm, Sig) < 1e-10
    assert _rel(out[_lib.PREC_FP32][0], P0) < 1e-4
    assert _rel(out[_lib.PREC_FP32][1], Sig) < 1e-4


def test_single_precision_with_double_cleanup():
    """The reference's single_prec = 2 mode: iterations on the FP32 path until
    the update reaches 1e-6, then FP64 iterations to 1e-11 converge to the same
    fixed point as the FP64-only loop (to 1e-9), with fewer FP64 iterations."""
    from mbsymmetry_amd.scloop import self_consistent, self_consistent_cleanup
    from mbsymmetry_amd.transforms import chebyshev
    sysm = S.synthetic("config1-weak", 2, 4, [8, 6, 6, 4, 2], [40, 32, 32, 24], 1, chebyshev(60, 50.0),
                       S.SEED_BASE + 1, v_scale=0.05, all_pairs=True)
    _say("sp + dp cleanup", "config1-weak")
    ctxs = {}
    for prec in (_lib.PREC_FP64, _lib.PREC_FP32):
        c = _lib.Context(sysm, 0, prec)
        c.generate_V(sysm.v_seed, sysm.v_scale, sysm.decay_table())
        c.set_reps(sysm.orep, sysm.auxrep); c.set_transforms(sysm.transforms); c.set_fock(sysm.F, sysm.S)
        ctxs[prec] = c
    c64, c32 = ctxs[_lib.PREC_FP64], ctxs[_lib.PREC_FP32]
    c64.upload_G(sysm.zeros_ts_kij()); c64.dyson(0.0, 0.0)
    G0 = c64.download_G((sysm.nts, sysm.ns, sysm.flat_ao))
    ref = self_consistent(c64, sysm, G0, 0.0, max_iter=20, tol=1e-11)
    mix = self_consistent_cleanup(c32, c64, sysm, G0, 0.0, max_iter=20, tol=1e-11, sp_tol=1e-6)
    c64.close(); c32.close()
    assert ref.converged and mix.converged
    n64 = sum(1 for h in mix.history if h["precision"] == "fp64")
    n32 = sum(1 for h in mix.history if h["precision"] == "fp32")
    assert n32 >= 1 and n64 < ref.iterations, (n32, n64, ref.iterations)
    assert np.abs(mix.G - ref.G).max() < 1e-9
