"""Parity at the sizes of the benchmark configurations, on the GPU through the C ABI.

- RHF/6-31G* n-decane (194 bf): J, K and F at the converged density against the oracle's
  integral-direct restatement (or_jk_direct: the integrals of eri_compressed, integrals/eri/
  mod.rs:992, contracted as build_jk, dft/ks_scf.rs:1369), integer screening counts against the
  oracle's Schwarz bounds, and the mixed-precision energy error.
- RHF/6-31+G* n-octane (188 bf) analytic gradient: translational invariance, central finite
  differences of converged SCF energies (h = 1e-4 bohr), screening count.
- B3LYP/6-31G* (H2O)16 (304 bf) on the literal 75 x 302 unpruned grid: kept-point set against the
  oracle's build_becke_grid (dft/grid.rs:103), the integral of rho, V_xc / E_xc against the oracle
  on whole 512-point blocks, and the mixed-precision J + K + XC energy error.
- IQCP's own C6H6/6-31G* and 6-31+G* RHF and B3LYP energies (iqcp_iteration_counts.json).
"""
import numpy as np
import pytest

import oracle
from helpers import bench_molecule, iqcp_energy
from iqcp_amd import workloads
from iqcp_amd.basis import build_basis
from iqcp_amd.engine import Engine, ScfConfig

pytestmark = pytest.mark.gpu


def _energy(D, H, F):
    return 0.5 * float(np.sum(D * (H + F)))


@pytest.mark.timeout(600)
def test_decane_fock_parity_at_config_size():
    b = workloads.decane_631gs()
    assert b.nbf == 194
    eng = Engine(b)
    scf = eng.rhf_scf(ScfConfig.tight())
    assert scf["converged"] == 1
    D = scf["D"]
    S, T, V = eng.one_electron()
    H = T + V
    # integer screening: unique shell quartets kept (1e-12) and the FP32 share (1e-6), bit-exact
    kept_o, low_o = oracle.screen_counts(b, 1e-12, 1e-6)
    eng.set_options("mixed", 1e-6)
    assert eng.screen_counts() == (kept_o, low_o)
    Jo, Ko, kept_d = oracle.jk_direct(b, D)
    assert kept_d == kept_o
    Fo = H + Jo - 0.5 * Ko
    Eo = _energy(D, H, Fo)
    E = {}
    for prec, tol in (("fp64", 1e-11), ("mixed", 1e-8)):
        eng.set_options(prec, 1e-6)
        jk = eng.build_jk(D)
        assert np.abs(jk.J - Jo).max() < tol, prec
        assert np.abs(jk.K - Ko).max() < tol, prec
        F = eng.build_fock(H, D)
        assert np.abs(F - Fo).max() < tol, prec
        E[prec] = _energy(D, H, F)
    assert abs(E["fp64"] - Eo) < 1e-9
    assert abs(E["mixed"] - E["fp64"]) < 1e-6
    # the stated error of the mixed build bounds the measured one, and is itself below 1 uHa
    eng.set_options("mixed", 1e-6)
    bound = eng.error_bound(D, 1.0)
    assert abs(E["mixed"] - E["fp64"]) <= bound < 1e-6, (E["mixed"] - E["fp64"], bound)


def _displaced(b, a, x, h, basis):
    atoms = [(int(z), list(map(float, p))) for z, p in zip(b.Z, b.xyz)]
    atoms[a][1][x] += h
    return build_basis(atoms, basis)


@pytest.mark.timeout(600)
def test_octane_gradient_at_config_size():
    b = workloads.octane_631pgs()
    assert b.nbf == 188
    eng = Engine(b)
    out = eng.rhf_scf(ScfConfig.tight())
    assert out["converged"] == 1
    assert eng.screen_counts()[0] == oracle.screen_counts(b)
    g = eng.rhf_gradient(out["D"], out["C"], out["eps"], b.n_occ)
    assert np.abs(g.sum(axis=0)).max() < 1e-8
    h = 1e-4
    cfg = ScfConfig.tight(energy_tol=1e-12, density_tol=1e-10)
    # carbons at both ends and in the middle, hydrogens on both ends: all three directions
    for a, x in ((0, 0), (3, 1), (7, 2), (8, 0), (17, 1), (25, 2)):
        e = [Engine(_displaced(b, a, x, s * h, "6-31+g*")).rhf_scf(cfg, D_init=out["D"])["energy_total"]
             for s in (+1, -1)]
        fd = (e[0] - e[1]) / (2 * h)
        assert abs(g[a, x] - fd) < 1e-5, (a, x, g[a, x], fd)


@pytest.mark.timeout(900)
def test_water16_b3lyp_at_config_size():
    b = workloads.water16_631gs()
    assert b.nbf == 304
    eng = Engine(b)
    n = eng.build_grid(75, "standard", False)
    pts, w, at = eng.get_grid()
    po, wo, ato = oracle.becke_grid(b.Z, b.xyz, 75, 0, 0)
    assert n == len(wo)  # kept-point count: integer output, bit-exact
    assert np.array_equal(at, ato)
    np.testing.assert_allclose(pts, po, rtol=0, atol=1e-13)
    np.testing.assert_allclose(w, wo, rtol=1e-12, atol=1e-300)
    out = eng.ks_scf(ScfConfig.tight())
    assert out["converged"] == 1
    D = out["D"]
    # the number of electrons from the grid (radial-rule error ~3e-6 relative)
    nel = 0.0
    for i in range(0, n, 100000):
        chi, _ = oracle.basis_on_grid(b, pts[i:i + 100000])
        nel += float(np.sum(w[i:i + 100000] * np.einsum("gi,gi->g", chi @ D, chi)))
    assert abs(nel - 2 * b.n_occ) < 2e-3, nel
    # J + K + XC energy at the converged density: mixed precision against all-FP64
    e2, exc = {}, {}
    for prec in ("fp64", "mixed"):
        eng.set_options(prec, 1e-6)
        jk = eng.build_jk(D)
        V, exc[prec] = eng.xc(D)
        e2[prec] = 0.5 * np.sum(D * jk.J) - 0.25 * 0.20 * np.sum(D * jk.K)
    assert abs(e2["mixed"] + exc["mixed"] - e2["fp64"] - exc["fp64"]) < 1e-6, (e2, exc)
    # two-electron part: within the stated bound of the mixed build
    bound = eng.error_bound(D, 0.20)
    assert abs(e2["mixed"] - e2["fp64"]) <= bound < 1e-6, (e2["mixed"] - e2["fp64"], bound)
    # V_xc / E_xc against the oracle on a sample of whole 512-point blocks of the grid
    rng = np.random.default_rng(0)
    blocks = rng.choice(n // 512, size=24, replace=False)
    idx = np.concatenate([np.arange(k * 512, (k + 1) * 512) for k in np.sort(blocks)])
    eng.set_options("fp64")
    eng.set_grid(pts[idx], w[idx])
    Vs, excs = eng.xc(D)
    chi, grad = oracle.basis_on_grid(b, pts[idx])
    Vo, exco, _ = oracle.vxc(chi, grad, w[idx], D)
    # the oracle drops functions below 1e-10 per block (dft/ks_scf.rs:958), the GPU keeps them
    assert abs(excs - exco) < 1e-9 * max(1.0, abs(exco))
    np.testing.assert_allclose(Vs, Vo, atol=2e-9, rtol=0)


@pytest.mark.timeout(600)
@pytest.mark.parametrize("bas", ["6-31g*", "6-31+g*"])
@pytest.mark.parametrize("method", ["RHF", "B3LYP"])
def test_c6h6_energy_vs_iqcp(bas, method):
    """IQCP's converged C6H6 energies (tight, DIIS; B3LYP on 75/Standard/SG-1), printed to
    1e-12: FP64 within 1e-7 Ha and mixed precision within 1e-6 Ha."""
    ref = iqcp_energy("C6H6", bas, method)
    b = bench_molecule("C6H6", bas)
    for prec, tol in (("fp64", 1e-7), ("mixed", 1e-6)):
        eng = Engine(b)
        eng.set_options(prec, 1e-6)
        if method == "RHF":
            out = eng.rhf_scf(ScfConfig.tight())
        else:
            eng.build_grid(75, "standard", True)
            out = eng.ks_scf(ScfConfig.tight())
        assert out["converged"] == 1
        assert abs(out["energy_total"] - ref) < tol, (prec, out["energy_total"] - ref)
