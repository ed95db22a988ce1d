"""Batched molecules (64 n-alkanes, RHF/6-31G): one launch per quartet class for all molecules."""
import numpy as np
import pytest

from helpers import bench_molecule, random_density
from iqcp_amd import workloads
from iqcp_amd.engine import BatchEngine, Engine, ScfConfig

pytestmark = pytest.mark.gpu


def test_batch_jk_equals_single():
    bases = [bench_molecule(m, b) for m, b in (("H2O", "6-31g*"), ("CH4", "6-31g"), ("NH3", "6-31+g*"), ("HF", "sto-3g"))]
    be = BatchEngine(bases)
    Ds = [random_density(b.nbf, seed=i, nocc=b.n_occ) for i, b in enumerate(bases)]
    Js, Ks = be.build_jk(Ds)
    for b, D, J, K in zip(bases, Ds, Js, Ks):
        jk = Engine(b).build_jk(D)
        np.testing.assert_allclose(J, jk.J, atol=1e-12, rtol=0)
        np.testing.assert_allclose(K, jk.K, atol=1e-12, rtol=0)


@pytest.mark.parametrize("tile", [(0, 1), (256, 2)])
def test_batch_jk_tiles_equal_single(tile):
    """Batched J/K through the wave-cooperative kernel (block 0) and a wide per-lane tile: work
    items of different molecules share the launches."""
    bases = [bench_molecule(m, b) for m, b in (("NH3", "6-31+g*"), ("H2O", "6-31g*"), ("CH4", "6-31g*"))]
    be = BatchEngine(bases)
    be.set_tiles([tile] * 21)
    Ds = [random_density(b.nbf, seed=20 + i, nocc=b.n_occ) for i, b in enumerate(bases)]
    Js, Ks = be.build_jk(Ds)
    for b, D, J, K in zip(bases, Ds, Js, Ks):
        jk = Engine(b).build_jk(D)
        np.testing.assert_allclose(J, jk.J, atol=1e-12, rtol=0)
        np.testing.assert_allclose(K, jk.K, atol=1e-12, rtol=0)


def test_batch_autotune_keeps_parity():
    """Tiles autotuned on a packed batch density: batched J/K still equal the single-molecule ones."""
    import torch
    bases = [bench_molecule(m, b) for m, b in (("NH3", "6-31+g*"), ("H2O", "6-31g*"), ("CH4", "6-31g"))]
    be = BatchEngine(bases)
    Ds = [random_density(b.nbf, seed=30 + i, nocc=b.n_occ) for i, b in enumerate(bases)]
    dD = torch.tensor(be.pack(Ds), device=torch.device("cuda:0"))
    tiles = be.autotune(dD.data_ptr(), 1)
    assert len(tiles) == 21 and be.get_tiles() == [[tuple(t) for t in pair] for pair in tiles]
    Js, Ks = be.build_jk(Ds)
    for b, D, J, K in zip(bases, Ds, Js, Ks):
        jk = Engine(b).build_jk(D)
        np.testing.assert_allclose(J, jk.J, atol=1e-12, rtol=0)
        np.testing.assert_allclose(K, jk.K, atol=1e-12, rtol=0)


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_scf_equals_single(use_graph):
    bases = [bench_molecule(m, b) for m, b in (("H2O", "6-31g*"), ("CH4", "6-31g"), ("HF", "6-31g"))]
    res, Ds, _ = BatchEngine(bases).rhf_scf(ScfConfig.tight(), use_graph=use_graph)
    for b, r in zip(bases, res):
        single = Engine(b).rhf_scf(ScfConfig.tight())
        assert r["converged"] == 1
        assert abs(r["energy_total"] - single["energy_total"]) < 1e-9


def test_alkane_batch_64():
    """The 64-molecule workload: every molecule converges; four checked against single runs.

    Core-Hamiltonian guess + DIIS alone diverges for molecule 56 (C11H24): the same SCF driven by
    the oracle's Fock builds on the CPU diverges identically, so it is a property of the
    iteration, not of the GPU path.  The workload therefore uses the reference's level_shift
    option (scf/mod.rs:1038-1063) at 0.1 Ha, with which all 64 converge (a sweep of DIIS start
    2 / 3, subspace 6 and shift 0.3 left 4-8 molecules unconverged; molecule 56 converges to
    -430.3014749375 Ha, the CPU-oracle-driven value)."""
    bases = workloads.alkane_batch_631g()
    assert len(bases) == 64 and min(b.nbf for b in bases) == 56 and max(b.nbf for b in bases) == 160
    cfg = ScfConfig.tight(level_shift=0.1)
    res, Ds, ms = BatchEngine(bases).rhf_scf(cfg)
    assert all(r["converged"] == 1 for r in res)
    assert abs(res[56]["energy_total"] - (-430.3014749375)) < 1e-8
    for i in (0, 17, 56, 63):
        single = Engine(bases[i]).rhf_scf(cfg)
        assert abs(res[i]["energy_total"] - single["energy_total"]) < 1e-9
