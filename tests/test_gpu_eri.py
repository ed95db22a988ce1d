"""GPU parity of the two-electron path against the CPU oracle (all through the C ABI)."""
import json

import numpy as np
import pytest

import oracle
from helpers import GOLDEN, bench_molecule, iqcp_energy, preset, random_density
from iqcp_amd import workloads
from iqcp_amd.engine import Engine, QcgError, ScfConfig

pytestmark = pytest.mark.gpu


def _eng(b, **kw):
    return Engine(b, device=0, **kw)


# ------------------------------------------------------------------ Boys / Rys
def test_gpu_boys_matches_oracle():
    rng = np.random.default_rng(0)
    T = np.concatenate([[0.0, 1e-15, 1e-12, 1e-6], rng.uniform(0, 2, 200), rng.uniform(0, 60, 200), [100.0, 1e3]])
    eng = _eng(preset("h2_sto3g_r1.4")[1])
    for m in (0, 2, 5, 10, 16):
        F = eng.boys_eval_all(m, T)
        for i, t in enumerate(T):
            ref = oracle.boys_all(m, float(t))
            # same algorithm on both sides; upward recurrence amplifies erf/exp ulp differences
            np.testing.assert_allclose(F[i], ref, rtol=1e-12, atol=1e-300)
    with pytest.raises(Exception):
        eng.boys_eval_all(51, T)
    with pytest.raises(Exception):
        eng.boys_eval_all(2, np.array([-1.0]))


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6])
def test_gpu_rys_moments_and_oracle(n):
    """GPU Rys rule: moments reproduce F_k(T), k < 2n, to 1e-14 relative; nodes/weights agree
    with the oracle's RDK restatement (rys/mod.rs:468) where that is well conditioned."""
    rng = np.random.default_rng(n)
    T = np.concatenate([[0.0, 1e-9, 0.5, 1.999, 2.0, 39.9, 40.0, 45.9, 46.0, 77.0, 78.0, 200.0],
                        rng.uniform(0, 80, 300)])
    eng = _eng(preset("h2_sto3g_r1.4")[1])
    x, w = eng.rys_roots(n, T)
    assert np.all((x >= 0) & (x < 1)) and np.all(w > 0)
    for i, t in enumerate(T):
        f = oracle.boys_all(2 * n - 1, float(t)) if n > 1 else oracle.boys_all(2, float(t))
        for k in range(2 * n):
            s = np.sum(w[i] * x[i] ** k)
            # highest moments are dominated by the largest node: ~k * (relative node error 1e-15)
            assert abs(s - f[k]) <= 2e-13 * f[k] + 1e-300, (n, t, k, abs(s - f[k]) / f[k])
        if n <= 4:
            r, ww = oracle.rys_roots(n, float(t))
            np.testing.assert_allclose(x[i], r, rtol=1e-9, atol=1e-14)
            np.testing.assert_allclose(w[i], ww, rtol=1e-9, atol=1e-300)


# ------------------------------------------------------------------ Schwarz / screening
@pytest.mark.parametrize("mol,bas", [("H2O", "6-31g*"), ("NH3", "6-31+g*"), ("C6H6", "sto-3g")])
def test_gpu_schwarz_and_screening_bit_exact(mol, bas):
    b = bench_molecule(mol, bas)
    eng = _eng(b)
    Q = eng.schwarz_bounds()
    Qo = oracle.schwarz(b)
    np.testing.assert_allclose(Q, Qo, rtol=1e-12, atol=1e-300)
    # integer screening decisions: the set of unique quartets with Q_ij Q_kl >= 1e-12
    iu = np.tril_indices(b.nshell)
    qg, qo = Q[iu], Qo[iu]
    keep_g = np.tril(np.outer(qg, qg) >= 1e-12)
    keep_o = np.tril(np.outer(qo, qo) >= 1e-12)
    assert np.array_equal(keep_g, keep_o)
    kept, _ = eng.screen_counts()
    assert kept == int(keep_o.sum())
    # the oracle's eri_compressed loop (si>=sj, all sk, sl<=sk) computes each pair-pair twice
    _, n_ref = oracle.eri_compressed(b)
    n_diag = int(np.sum(qo * qo >= 1e-12))
    assert n_ref == 2 * kept - n_diag


# ------------------------------------------------------------------ compressed tensor
@pytest.mark.parametrize("mol,bas", [("H2O", "6-31g*"), ("NH3", "6-31+g*"), ("CH4", "6-31g*")])
def test_gpu_eri_compressed(mol, bas):
    b = bench_molecule(mol, bas)
    ref, _ = oracle.eri_compressed(b)
    got = _eng(b).eri_compressed()
    assert np.abs(got - ref).max() < 1e-12 * max(1.0, np.abs(ref).max())


def test_gpu_eri_compressed_propane():
    """Full FP64 tensor of RHF/STO-3G propane (23 bf, 38 226 unique ERIs)."""
    b = workloads.propane_sto3g()
    ref, _ = oracle.eri_compressed(b)
    got = _eng(b).eri_compressed()
    assert got.size == 38226
    assert np.abs(got - ref).max() < 1e-12


def test_gpu_eri_presets():
    for name in ("h2_sto3g_r1.4", "heh_plus_sto3g", "lih_sto3g", "h2o_sto3g", "nh3_sto3g"):
        d, b = preset(name)
        got = _eng(b).eri_compressed()
        ref, _ = oracle.eri_compressed(b)
        assert np.abs(got - ref).max() < 1e-13
        assert np.abs(got - np.array(d["eri_compressed"])).max() < 2e-7


# ------------------------------------------------------------------ J / K / Fock
@pytest.mark.parametrize("mol,bas", [("H2O", "6-31g*"), ("NH3", "6-31+g*"), ("CH4", "6-31g*"), ("HF", "6-31g")])
def test_gpu_jk_fp64_vs_oracle(mol, bas):
    b = bench_molecule(mol, bas)
    eri, _ = oracle.eri_compressed(b)
    eng = _eng(b)
    for seed in (0, 1):
        D = random_density(b.nbf, seed=seed)
        Jo, Ko = oracle.build_jk(D, eri, b.nbf)
        jk = eng.build_jk(D)
        np.testing.assert_allclose(jk.J, Jo, atol=1e-11, rtol=0)
        np.testing.assert_allclose(jk.K, Ko, atol=1e-11, rtol=0)
        H = np.diag(np.linspace(-2, 1, b.nbf))
        F = eng.build_fock(H, D)
        np.testing.assert_allclose(F, oracle.build_fock(H, D, eri, b.nbf), atol=1e-11, rtol=0)
    # J only / K only
    jk = eng.build_jk(D, want_j=True, want_k=False)
    np.testing.assert_allclose(jk.J, Jo, atol=1e-11)
    jk = eng.build_jk(D, want_j=False, want_k=True)
    np.testing.assert_allclose(jk.K, Ko, atol=1e-11)


def test_gpu_jk_h2_pyscf():
    d, b = preset("h2_sto3g_r1.4")
    g = json.loads((GOLDEN / "h2_fock.json").read_text())
    jk = _eng(b).build_jk(np.array(g["D"]).reshape(2, 2))
    # raw (qc-core) coefficients vs PySCF's renormalised ones differ at ~1e-9 for H2
    np.testing.assert_allclose(jk.J.ravel(), g["J"], atol=5e-8)
    np.testing.assert_allclose(jk.K.ravel(), g["K"], atol=5e-8)


def test_gpu_jk_propane_vs_oracle_and_mixed_bound():
    b = workloads.propane_sto3g()
    eri, _ = oracle.eri_compressed(b)
    D = random_density(b.nbf, seed=5, nocc=b.n_occ)
    Jo, Ko = oracle.build_jk(D, eri, b.nbf)
    eng = _eng(b)
    jk = eng.build_jk(D)
    np.testing.assert_allclose(jk.J, Jo, atol=1e-11)
    np.testing.assert_allclose(jk.K, Ko, atol=1e-11)
    eng.set_options("mixed", 1e-6)
    jm = eng.build_jk(D)
    # mixed: FP32 quartets all have Q_ab Q_cd < 1e-6; each contributes < 1e-6 * 1e-6 relative error
    assert np.abs(jm.J - Jo).max() < 1e-8
    assert np.abs(jm.K - Ko).max() < 1e-8


@pytest.mark.parametrize("tile", [(64, 1), (128, 2), (0, 1), (0, 3)])
def test_gpu_jk_tiles_vs_oracle(tile):
    """Every quartet class through one launch configuration: per-lane kernel with several kets per
    thread, or (block 0) the wave-cooperative kernel for every class with >= 27 components.  NH3
    6-31+G* has every class (dd|dd) .. (ss|ss) with contracted s/p shells."""
    b = bench_molecule("NH3", "6-31+g*")
    eri, _ = oracle.eri_compressed(b)
    eng = _eng(b)
    eng.set_tiles([tile] * 21)
    assert eng.get_tiles() == [[tile, tile]] * 21
    D = random_density(b.nbf, seed=7)
    Jo, Ko = oracle.build_jk(D, eri, b.nbf)
    jk = eng.build_jk(D)
    np.testing.assert_allclose(jk.J, Jo, atol=1e-11, rtol=0)
    np.testing.assert_allclose(jk.K, Ko, atol=1e-11, rtol=0)
    eng.set_options("mixed", 1e-6)
    jm = eng.build_jk(D)
    assert np.abs(jm.J - Jo).max() < 1e-8
    assert np.abs(jm.K - Ko).max() < 1e-8


def test_gpu_fock_graph_replay():
    """The captured device Fock build, replayed several times with eager builds, host copies and
    synchronisations in between, equals the eager build every time (FP64 and mixed)."""
    import torch
    b = bench_molecule("CH4", "6-31g*")
    eng = _eng(b)
    dev = torch.device("cuda:0")
    H = np.diag(np.linspace(-2, 1, b.nbf))
    st = torch.cuda.Stream(device=dev)
    for prec in ("fp64", "mixed"):
        eng.set_options(prec, 1e-6)
        for seed in (0, 1):
            D = random_density(b.nbf, seed=seed)
            dH, dD = torch.tensor(H, device=dev), torch.tensor(D, device=dev)
            dF_e, dF_g = torch.empty_like(dD), torch.empty_like(dD)
            eng.set_graph(False)
            eng.build_fock_device(dH.data_ptr(), dD.data_ptr(), dF_e.data_ptr(), 0.5, st.cuda_stream)
            torch.cuda.synchronize()
            Fe = dF_e.cpu().numpy()
            eng.set_graph(True)
            for rep in range(4):
                dF_g.fill_(7.0)
                torch.cuda.synchronize()
                eng.build_fock_device(dH.data_ptr(), dD.data_ptr(), dF_g.data_ptr(), 0.5, st.cuda_stream)
                torch.cuda.synchronize()
                np.testing.assert_allclose(dF_g.cpu().numpy(), Fe, atol=1e-12, rtol=0)
                # eager work of the library between replays (same J/K buffers the graph writes)
                eng.build_jk(random_density(b.nbf, seed=10 + rep))
            eng.set_graph(False)


def test_gpu_fock_bitwise_reproducible():
    """J and K are accumulated in fixed point (order-independent integer sums), so with the tiles
    held fixed every build of the same density gives the same bits: 20 graph replays of the device
    Fock build (21 class kernels on four streams) and eager builds in between, FP64 and mixed.
    The reference merges its per-thread Fock parts in a fixed order for the same reason
    (integrals/eri/mod.rs:1366-1380)."""
    import torch
    b = bench_molecule("C6H6", "6-31g*")
    eng = _eng(b)
    dev = torch.device("cuda:0")
    H = np.diag(np.linspace(-2, 1, b.nbf))
    D = random_density(b.nbf, seed=3)
    eri, _ = oracle.eri_compressed(b)
    Fo = oracle.build_fock(H, D, eri, b.nbf)
    st = torch.cuda.Stream(device=dev)
    dH, dD = torch.tensor(H, device=dev), torch.tensor(D, device=dev)
    dF = torch.empty_like(dD)
    for prec, tol in (("fp64", 1e-11), ("mixed", 1e-8)):
        eng.set_options(prec, 1e-6)
        eng.set_graph(True)
        first = None
        for rep in range(20):
            dF.fill_(7.0)
            eng.build_fock_device(dH.data_ptr(), dD.data_ptr(), dF.data_ptr(), 0.5, st.cuda_stream)
            torch.cuda.synchronize()
            F = dF.cpu().numpy()
            if first is None:
                first = F
                assert np.abs(F - Fo).max() < tol, prec
                assert np.array_equal(F, F.T)  # integer words summed before the conversion
            assert np.array_equal(F, first), (prec, rep, np.abs(F - first).max())
            if rep % 5 == 4:
                eng.build_jk(random_density(b.nbf, seed=rep))
        eng.set_graph(False)
        for rep in range(3):
            assert np.array_equal(eng.build_fock(H, D), first), (prec, "eager", rep)


def test_gpu_launch_schedule_does_not_change_bits():
    """The class kernels' order over the four streams (set per-class costs, the plan's estimate,
    reversed costs) changes which blocks meet at the accumulators, not the result: every schedule
    gives the same J and K bits, and the plan's estimate is reported as such."""
    b = bench_molecule("C6H6", "6-31g*")
    eng = _eng(b)
    D = random_density(b.nbf, seed=5)
    eng.set_options("fp64")
    jk0 = eng.build_jk(D)
    cost, measured = eng.get_class_cost()
    assert not measured and len(cost) == 21 and max(cost) > 0 and min(cost) >= 0
    for c in (list(reversed(cost)), [1.0] * 21, [float(i + 1) for i in range(21)], None):
        eng.set_class_cost(c)
        got, m = eng.get_class_cost()
        assert m == (c is not None)
        if c is not None:
            assert got == c
        jk = eng.build_jk(D)
        assert np.array_equal(jk.J, jk0.J) and np.array_equal(jk.K, jk0.K), c
    with pytest.raises(QcgError):
        eng.set_class_cost([1.0] * 20 + [float("nan")])


def test_gpu_autotune_keeps_parity():
    """After the autotuner picks a tile per (class, precision part), J/K still match the oracle,
    and every chosen tile is one of the candidates."""
    import torch
    b = bench_molecule("NH3", "6-31+g*")
    eri, _ = oracle.eri_compressed(b)
    eng = _eng(b)
    D = random_density(b.nbf, seed=11)
    dD = torch.tensor(D, device=torch.device("cuda:0"))
    cands = {(64, 1), (64, 2), (64, 4), (128, 1), (256, 1), (64, 8), (128, 4), (256, 2), (0, 1), (0, 2), (0, 4),
             (0, 8)}
    Jo, Ko = oracle.build_jk(D, eri, b.nbf)
    for prec in ("fp64", "mixed"):
        eng.set_options(prec, 1e-6)
        tiles = eng.autotune(dD.data_ptr(), 1)
        assert len(tiles) == 21 and all(tuple(t) in cands for pair in tiles for t in pair)
        assert eng.get_tiles() == [[tuple(t) for t in pair] for pair in tiles]
        jk = eng.build_jk(D)
        tol = 1e-11 if prec == "fp64" else 1e-8
        np.testing.assert_allclose(jk.J, Jo, atol=tol, rtol=0)
        np.testing.assert_allclose(jk.K, Ko, atol=tol, rtol=0)
        # the whole-build refinement changes only FP64-part tiles, among the same candidates
        refined = eng.autotune_refine(dD.data_ptr(), 4)
        assert all(tuple(t) in cands for pair in refined for t in pair)
        assert [pair[1] for pair in refined] == [tuple(pair[1]) for pair in tiles]
        assert eng.get_tiles() == [[tuple(t) for t in pair] for pair in refined]
        jk = eng.build_jk(D)
        np.testing.assert_allclose(jk.J, Jo, atol=tol, rtol=0)
        np.testing.assert_allclose(jk.K, Ko, atol=tol, rtol=0)


def test_gpu_set_tiles_rejects_bad_tiles():
    b = bench_molecule("H2O", "sto-3g")
    eng = _eng(b)
    with pytest.raises(Exception):
        eng.set_tiles([(32, 1)] * 21)
    with pytest.raises(Exception):
        eng.set_tiles([(64, 0)] * 21)
    with pytest.raises(Exception):
        eng.set_tiles([(512, 1)] * 21)  # the kernels are built for at most 256 threads
    with pytest.raises(Exception):
        eng.set_prim_split(-1)


@pytest.mark.parametrize("split", [1, 3, 7])
def test_gpu_prim_split_keeps_parity(split):
    """Work items cut into bra primitive ranges give the same J/K (per-lane and wave-cooperative
    kernels, FP64 and mixed) and the same ERI gradient as whole bra pairs."""
    b = bench_molecule("NH3", "6-31+g*")
    eri, _ = oracle.eri_compressed(b)
    D = random_density(b.nbf, seed=13, nocc=b.n_occ)
    Jo, Ko = oracle.build_jk(D, eri, b.nbf)
    eng = _eng(b)
    g0 = eng.eri_gradient(D, 1.0)
    eng.set_prim_split(split)
    for tile in ((64, 2), (0, 2)):
        eng.set_tiles([tile] * 21)
        for prec, tol in (("fp64", 1e-11), ("mixed", 1e-8)):
            eng.set_options(prec, 1e-6)
            jk = eng.build_jk(D)
            np.testing.assert_allclose(jk.J, Jo, atol=tol, rtol=0, err_msg=f"{tile} {prec}")
            np.testing.assert_allclose(jk.K, Ko, atol=tol, rtol=0, err_msg=f"{tile} {prec}")
    eng.set_options("fp64")
    np.testing.assert_allclose(eng.eri_gradient(D, 1.0), g0, atol=1e-11, rtol=0)


# ------------------------------------------------------------------ SCF energies vs IQCP
@pytest.mark.parametrize("mol,bas", [("H2O", "sto-3g"), ("H2O", "6-31g*"), ("CH4", "6-31g*"),
                                     ("NH3", "6-31+g*"), ("HF", "6-31g"), ("C6H6", "sto-3g")])
def test_gpu_rhf_energy_vs_iqcp(mol, bas):
    """IQCP's own converged energies (tight, DIIS; iqcp_iteration_counts.json, printed to 1e-12)."""
    b = bench_molecule(mol, bas)
    out = _eng(b).rhf_scf(ScfConfig.tight())
    assert out["converged"] == 1
    assert abs(out["energy_total"] - iqcp_energy(mol, bas, "RHF")) < 1e-8


def test_gpu_one_electron_vs_presets():
    for name in ("h2_sto3g_r1.4", "lih_sto3g", "h2o_sto3g"):
        d, b = preset(name)
        S, T, V = _eng(b).one_electron()
        n = b.nbf
        assert np.abs(S - np.array(d["s_matrix"]).reshape(n, n)).max() < 1e-9
        assert np.abs(T + V - np.array(d["h_core"]).reshape(n, n)).max() < 5e-9


@pytest.mark.parametrize("mol,bas", [("H2O", "6-31+g*"), ("CH4", "6-31g*"), ("NH3", "sto-3g")])
def test_gpu_one_electron_vs_oracle(mol, bas):
    """Device S, T, V (onee_matrix_kernel: Obara-Saika tables, Rys moments) against the oracle's
    binomial Gaussian-moment restatement."""
    b = bench_molecule(mol, bas)
    S, T, V = _eng(b).one_electron()
    So, To, Vo = oracle.one_electron(b)
    for x, y in ((S, So), (T, To), (V, Vo)):
        np.testing.assert_allclose(x, y, atol=1e-12 * np.abs(y).max(), rtol=0)


def test_gpu_rhf_presets_heh_plus():
    """Charged system through qcg_set_charge: HeH+ STO-3G against the PySCF preset energy."""
    d, b = preset("heh_plus_sto3g")
    out = _eng(b).rhf_scf(ScfConfig.tight())
    assert abs(out["energy_total"] - d["reference"]["energy"]) < 1e-7


# ------------------------------------------------------------------ mixed-precision error statement
def _shell_of_bf(b):
    s = np.zeros(b.nbf, dtype=int)
    for sh, (off, l) in enumerate(zip(b.shell_bf_off, b.shell_l)):
        s[off:off + (l + 1) * (l + 2) // 2] = sh
    return s


def test_gpu_fp32_quartet_error_against_schwarz():
    """Single integrals in both precisions: with the FP32 cutoff above every Schwarz product all
    quartets run in FP32, and a density that is one symmetric unit pair (k, l) turns J into
    2 (mu nu|k l).  Each FP32 integral's error stays below 2^-20 Q_ab Q_cd (the bound of
    qcg_error_bound allows 2^-16).  H2O 6-31G* (s, p and d shells), k, l over the first function
    of every shell pair."""
    b = bench_molecule("H2O", "6-31g*")
    eng = _eng(b)
    Q = eng.schwarz_bounds()
    sh = _shell_of_bf(b)
    worst = 0.0
    for x in range(b.nshell):
        for y in range(x + 1):
            k, l = b.shell_bf_off[x], b.shell_bf_off[y]
            D = np.zeros((b.nbf, b.nbf))
            D[k, l] = D[l, k] = 1.0
            eng.set_options("fp64")
            J64 = eng.build_jk(D, want_k=False).J
            eng.set_options("mixed", 1e30)
            J32 = eng.build_jk(D, want_k=False).J
            scale = (1.0 if k == l else 2.0) * Q[np.ix_(sh, sh)] * Q[x, y]
            keep = scale > 1e-12
            worst = max(worst, float(np.max(np.abs(J32 - J64)[keep] / scale[keep])))
    assert worst < 2.0 ** -20, worst


@pytest.mark.parametrize("cutoff", [1e-6, 1e30])
def test_gpu_error_bound_holds(cutoff):
    """qcg_error_bound bounds the two-electron energy error of a mixed build against the FP64 build
    at the same density (RHF and hybrid exchange weights), for the default cutoff and with every
    quartet in FP32; zero in FP64 mode."""
    b = bench_molecule("C6H6", "6-31g*")
    eng = _eng(b)
    D = random_density(b.nbf, seed=7, nocc=b.n_occ)
    jk64 = eng.build_jk(D)
    assert eng.error_bound(D) == 0.0
    eng.set_options("mixed", cutoff)
    jk32 = eng.build_jk(D)
    for a in (1.0, 0.2):
        e64 = 0.5 * np.sum(D * (jk64.J - 0.5 * a * jk64.K))
        e32 = 0.5 * np.sum(D * (jk32.J - 0.5 * a * jk32.K))
        bound = eng.error_bound(D, a)
        assert abs(e32 - e64) <= bound, (a, abs(e32 - e64), bound)
        if cutoff == 1e-6:
            assert bound < 1e-6  # the default split keeps a 1 uHa statement


def test_gpu_scf_history():
    """The iteration trace mirrors ScfIteration (scf/mod.rs:1407-1531): iteration 0 is the guess
    build (no delta_e / rms, no DIIS), the last entry is the converged one with the returned
    energy; a mixed run's early iterations are FP32 builds with a positive error bound and its last
    ones FP64 builds."""
    b = bench_molecule("H2O", "6-31g*")
    for prec in ("fp64", "mixed"):
        eng = _eng(b)
        eng.set_options(prec, 1e-3)  # H2O has no quartet below the default 1e-6 split
        out = eng.rhf_scf(ScfConfig.tight())
        h = eng.scf_history()
        assert len(h) == out["iterations"] + 1
        assert [r["iteration"] for r in h] == list(range(len(h)))
        assert np.isnan(h[0]["delta_e"]) and np.isnan(h[0]["rms_density_change"]) and not h[0]["diis_applied"]
        assert h[-1]["converged"] == 1 and sum(r["converged"] for r in h) == 1
        assert h[-1]["energy_total"] == out["energy_total"]
        assert all(r["diis_applied"] for r in h[2:])
        for r in h[1:]:
            assert r["delta_e"] >= 0.0 and r["rms_density_change"] >= 0.0
        if prec == "fp64":
            assert not any(r["fp32_build"] for r in h) and all(r["error_bound"] == 0.0 for r in h)
        else:
            assert h[0]["fp32_build"] == 1 and h[0]["error_bound"] > 0.0
            assert h[-1]["fp32_build"] == 0
