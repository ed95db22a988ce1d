"""Analytic RHF nuclear gradient on the GPU: finite differences, translational invariance, PySCF."""
import json

import numpy as np
import pytest

from helpers import GOLDEN, bench_molecule, random_density
from iqcp_amd.basis import build_basis
from iqcp_amd.engine import Engine, ScfConfig

pytestmark = pytest.mark.gpu


def _atoms(b):
    return [(int(z), list(map(float, x))) for z, x in zip(b.Z, b.xyz)]


def _displaced(atoms, basis, a, x, h):
    at = [(z, list(p)) for z, p in atoms]
    at[a][1][x] += h
    return build_basis(at, basis)


def _distorted(mol, basis, seed=0, sigma=0.05):
    b = bench_molecule(mol, basis)
    rng = np.random.default_rng(seed)
    at = [(int(z), list(np.asarray(x) + rng.normal(0, sigma, 3))) for z, x in zip(b.Z, b.xyz)]
    return at


@pytest.mark.parametrize("mol,bas", [("H2O", "6-31+g*"), ("CH4", "6-31g*")])
def test_eri_gradient_vs_fd_of_jk(mol, bas):
    """d/dR of E2 = 1/2 Tr DJ - 1/4 Tr DK at fixed D equals the adjoint kernel's result."""
    atoms = _distorted(mol, bas)
    b = build_basis(atoms, bas)
    D = random_density(b.nbf, seed=4, nocc=b.n_occ)
    g = Engine(b).eri_gradient(D, 1.0)
    h = 1e-4
    for a in range(b.natom):
        for x in range(3):
            e = []
            for s in (+1, -1):
                jk = Engine(_displaced(atoms, bas, a, x, s * h)).build_jk(D)
                e.append(0.5 * np.sum(D * jk.J) - 0.25 * np.sum(D * jk.K))
            fd = (e[0] - e[1]) / (2 * h)
            assert abs(g[a, x] - fd) < 2e-7, (a, x, g[a, x], fd)
    np.testing.assert_allclose(g.sum(axis=0), 0.0, atol=1e-9)


def test_one_electron_gradient_vs_fd():
    atoms = _distorted("NH3", "6-31g*", seed=1)
    bas = "6-31g*"
    b = build_basis(atoms, bas)
    D = random_density(b.nbf, seed=5, nocc=b.n_occ)
    W = random_density(b.nbf, seed=6, nocc=b.n_occ) * -0.3
    g = Engine(b).one_electron_gradient(D, W)
    h = 1e-4
    for a in range(b.natom):
        for x in range(3):
            e = []
            for s in (+1, -1):
                bd = _displaced(atoms, bas, a, x, s * h)
                S, T, V = Engine(bd).one_electron()
                e.append(np.sum(D * (T + V)) - np.sum(W * S) + bd.e_nuc)
            fd = (e[0] - e[1]) / (2 * h)
            assert abs(g[a, x] - fd) < 2e-7, (a, x, g[a, x], fd)


def test_rhf_gradient_vs_fd_of_scf_energy():
    """Full analytic RHF gradient against central differences of converged energies (h = 1e-4)."""
    bas = "6-31g*"
    atoms = _distorted("H2O", bas, seed=3)
    b = build_basis(atoms, bas)
    eng = Engine(b)
    out = eng.rhf_scf(ScfConfig.tight())
    g = eng.rhf_gradient(out["D"], out["C"], out["eps"], b.n_occ)
    h = 1e-4
    cfg = ScfConfig.tight(energy_tol=1e-12, density_tol=1e-10)
    for a in range(b.natom):
        for x in range(3):
            e = [Engine(_displaced(atoms, bas, a, x, s * h)).rhf_scf(cfg, D_init=out["D"])["energy_total"]
                 for s in (+1, -1)]
            fd = (e[0] - e[1]) / (2 * h)
            assert abs(g[a, x] - fd) < 1e-5, (a, x, g[a, x], fd)
    np.testing.assert_allclose(g.sum(axis=0), 0.0, atol=1e-8)


@pytest.mark.parametrize("mol,bas", [("H2O", "sto-3g"), ("H2O", "6-31g*"), ("HF", "6-31g*"), ("NH3", "sto-3g")])
def test_rhf_gradient_vs_pyscf(mol, bas):
    """PySCF analytic gradients (tests/golden/dft/gradient_validation.json); the reference's own
    agreement with them is 1e-5 Ha/bohr (README.md:36); PySCF renormalises contractions."""
    ref = [r for r in json.loads((GOLDEN / "pyscf_gradients.json").read_text())["results"]
           if r["molecule"] == mol and r["basis"] == bas and r["method"] == "RHF"]
    if not ref:
        pytest.skip("no PySCF row")
    b = bench_molecule(mol, bas)
    eng = Engine(b)
    out = eng.rhf_scf(ScfConfig.tight())
    g = eng.rhf_gradient(out["D"], out["C"], out["eps"], b.n_occ)
    np.testing.assert_allclose(g, np.array(ref[0]["gradient"]), atol=2e-5, rtol=0)
