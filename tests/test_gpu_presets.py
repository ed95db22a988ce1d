"""Preset systems on the GPU: build_fock and rhf_scf on a stored compressed tensor
(qcg_build_fock_stored, qcg_rhf_scf_stored), and presets written from a basis."""
import numpy as np
import pytest

import oracle
from helpers import GOLDEN, bench_molecule, random_density
from iqcp_amd import presets
from iqcp_amd.engine import Engine, QcgError, ScfConfig

pytestmark = pytest.mark.gpu

NAMES = ["h2_sto3g_r1.4", "heh_plus_sto3g", "lih_sto3g", "h2o_sto3g", "nh3_sto3g"]


def _preset(name):
    return presets.from_json((GOLDEN / "presets" / f"{name}.json").read_text())


@pytest.mark.parametrize("name", NAMES)
def test_gpu_rhf_scf_on_reference_presets(name):
    """rhf_scf on the PySCF-made S, H and tensor of the reference's presets reproduces PySCF's
    energy stored in the file."""
    p = _preset(name)
    out = presets.StoredEngine().rhf_scf(p, ScfConfig.tight())
    assert out["converged"] == 1
    assert abs(out["energy_total"] - p.reference_energy()) < 1e-8, (out["energy_total"], p.reference_energy())


@pytest.mark.parametrize("mol,bas", [("H2O", "6-31g*"), ("CH4", "6-31+g*")])
def test_gpu_build_fock_stored_vs_oracle(mol, bas):
    b = bench_molecule(mol, bas)
    eri, _ = oracle.eri_compressed(b)
    rng = np.random.default_rng(3)
    H = rng.normal(size=(b.nbf, b.nbf))
    H = H + H.T
    D = random_density(b.nbf, seed=2, nocc=b.n_occ)
    F = presets.StoredEngine().build_fock(H, D, eri)
    ref = oracle.build_fock(H, D, eri, b.nbf)
    np.testing.assert_allclose(F, ref, atol=1e-12 * np.abs(ref).max(), rtol=0)
    # the integral-direct build of the same molecule agrees to the Schwarz threshold
    Fd = Engine(b).build_fock(H, D)
    np.testing.assert_allclose(Fd, F, atol=1e-9, rtol=0)


def test_gpu_preset_from_basis_round_trip():
    """A preset written from a basis (S, H, tensor on the GPU) reads back, matches the oracle's
    tensor, and its stored-tensor SCF gives the integral-direct SCF energy."""
    b = bench_molecule("H2O", "6-31g*")
    eng = Engine(b)
    p = presets.from_basis(b, "h2o_631gs", "H2O (6-31G*)", engine=eng)
    q = presets.from_json(p.to_json())
    eri, _ = oracle.eri_compressed(b)
    np.testing.assert_allclose(q.eri_compressed, eri, atol=1e-12, rtol=0)
    direct = eng.rhf_scf(ScfConfig.tight())
    stored = presets.StoredEngine().rhf_scf(q, ScfConfig.tight())
    assert stored["converged"] == 1
    assert abs(stored["energy_total"] - direct["energy_total"]) < 1e-9


def test_gpu_stored_validation():
    p = _preset("h2_sto3g_r1.4")
    se = presets.StoredEngine()
    with pytest.raises(QcgError, match="Invalid ERI array size: expected 6, got 5"):
        se.build_fock(p.h_core, np.eye(2), p.eri_compressed[:5])
    p.nelec = 3
    with pytest.raises(QcgError, match="Invalid number of electrons: 3"):
        se.rhf_scf(p)
