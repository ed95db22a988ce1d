"""Pin the DFT part of the oracle: libxc vectors, Lebedev exactness, grid integration."""
import json
import math

import numpy as np
import pytest

import oracle
from helpers import GOLDEN, bench_molecule

XC = json.loads((GOLDEN / "xc_libxc.json").read_text())


@pytest.mark.parametrize("name", ["slater", "vwn5", "b88", "lyp"])
def test_xc_components_vs_libxc(name):
    ref = XC[name]
    exc, vr, vs = oracle.xc_component(name, XC["rho"], XC["sigma"])
    np.testing.assert_allclose(exc, ref["exc"], atol=ref["tol"], rtol=0)
    np.testing.assert_allclose(vr, ref["vrho"], atol=ref["tol"], rtol=0)
    if "vsigma" in ref:
        np.testing.assert_allclose(vs, ref["vsigma"], atol=ref["tol"], rtol=0)


def test_b3lyp5_vs_libxc():
    ref = XC["b3lyp5"]
    exc, vr, vs = oracle.b3lyp(XC["rho"], XC["sigma"])
    np.testing.assert_allclose(exc, ref["exc"], atol=ref["tol"], rtol=0)
    np.testing.assert_allclose(vr, ref["vrho"], atol=ref["tol"], rtol=0)
    np.testing.assert_allclose(vs, ref["vsigma"], atol=ref["tol"], rtol=0)


def test_b3lyp5_vs_mpmath():
    """40-digit mpmath evaluation of the published forms, derivatives by mpmath differentiation
    (tests/golden/make_b3lyp_mpmath.py): independent of both floating-point implementations."""
    d = json.loads((GOLDEN / "b3lyp_mpmath.json").read_text())
    got = oracle.b3lyp(np.array(d["rho"]), np.array(d["sigma"]))
    for k, a in zip(("exc", "vrho", "vsigma"), got):
        np.testing.assert_allclose(a, d[k], rtol=1e-13, atol=0, err_msg=k)


def test_b3lyp_derivatives_by_finite_difference():
    """vrho = d(rho exc)/d rho, vsigma = d(rho exc)/d sigma (independent of the libxc vectors)."""
    rng = np.random.default_rng(1)
    rho = rng.uniform(0.01, 3.0, 50)
    sig = rng.uniform(0.0, 2.0, 50)
    h = 1e-6
    e0, vr, vs = oracle.b3lyp(rho, sig)
    ep, _, _ = oracle.b3lyp(rho * (1 + h), sig)
    em, _, _ = oracle.b3lyp(rho * (1 - h), sig)
    fd_r = ((rho * (1 + h)) * ep - (rho * (1 - h)) * em) / (2 * h * rho)
    np.testing.assert_allclose(vr, fd_r, rtol=1e-6, atol=1e-9)
    sp, sm = sig + 1e-5, np.maximum(sig - 1e-5, 0)
    ep, _, _ = oracle.b3lyp(rho, sp)
    em, _, _ = oracle.b3lyp(rho, sm)
    fd_s = rho * (ep - em) / (sp - sm)
    np.testing.assert_allclose(vs, fd_s, rtol=1e-5, atol=1e-9)


def test_zero_density_threshold():
    exc, vr, vs = oracle.b3lyp([0.0, 1e-21, 1e-20], [0.0, 1e-30, 1e-25])
    assert np.all(exc == 0) and np.all(vr == 0) and np.all(vs == 0)


def _gauss_integral(b, pts, w, alpha=1.3):
    """integral of sum_A (2 alpha/pi)^(3/2) exp(-2 alpha |r-A|^2) = natom."""
    s = 0.0
    for A in b.xyz:
        r2 = np.sum((pts - A) ** 2, axis=1)
        s += np.sum(w * (2 * alpha / math.pi) ** 1.5 * np.exp(-2 * alpha * r2))
    return s


@pytest.mark.parametrize("pruning", [1, 0])
def test_becke_grid_integrates_gaussians(pruning):
    b = bench_molecule("H2O", "6-31g*")
    pts, w, at = oracle.becke_grid(b.Z, b.xyz, 75, 0, pruning)
    assert np.all(np.abs(w) > 1e-30)
    assert np.all(np.diff(at) >= 0)  # atom by atom
    # 75-point Mura-Knowles radial rule: ~1e-6 for this Gaussian; a wrong partition is off by ~1e-2
    assert abs(_gauss_integral(b, pts, w) - b.natom) < 1e-5


def test_grid_sizes():
    """Point counts of the two configurations (integer outputs)."""
    b = bench_molecule("H2O", "6-31g*")
    p1, _, _ = oracle.becke_grid(b.Z, b.xyz, 75, 0, 1)
    p0, _, _ = oracle.becke_grid(b.Z, b.xyz, 75, 0, 0)
    assert len(p0) <= 3 * 75 * 302
    assert len(p1) < len(p0)
    # single atom: no partition, every raw point kept (Mura-Knowles weights are positive)
    pa, wa, _ = oracle.becke_grid(np.array([8]), np.zeros((1, 3)), 75, 0, 0)
    assert len(pa) == 75 * 302


# ------------------------------------------------------------------ XC gradient (scf/gradient.rs:1687)
def _displaced(b, atom, axis, h):
    """The basis with every shell on `atom` moved by h along `axis` (the grid is NOT moved)."""
    import copy
    c = copy.deepcopy(b)
    sc = np.array(c.shell_center, dtype=float).copy()
    sc[np.asarray(c.shell_atom) == atom, axis] += h
    c.shell_center = sc
    return c


def _exc_fixed_grid(b, pts, w, D):
    chi, g = oracle.basis_on_grid(b, pts)
    X = chi @ D
    rho = np.maximum(np.einsum("gm,gm->g", chi, X), 0.0)
    grho = 2.0 * np.einsum("gmd,gm->gd", g, X)
    exc, _, _ = oracle.b3lyp(rho, np.einsum("gd,gd->g", grho, grho))
    return float(np.sum(w * exc * rho))


def test_basis_hessian_vs_fd():
    """d2 chi/dr_a dr_b (gradient.rs:1938) against central differences of the analytic first
    derivatives (ks_scf.rs:115), for s, p and Cartesian d shells with contracted primitives."""
    b = bench_molecule("H2O", "6-31g*")
    rng = np.random.default_rng(3)
    pts = np.array(b.xyz)[rng.integers(0, b.natom, 40)] + rng.normal(scale=0.8, size=(40, 3))
    hess = oracle.basis_hessian_on_grid(b, pts)
    h = 1e-5
    idx = {(0, 0): 0, (0, 1): 1, (0, 2): 2, (1, 1): 3, (1, 2): 4, (2, 2): 5}
    for bdir in range(3):
        dp = pts.copy(); dp[:, bdir] += h
        dm = pts.copy(); dm[:, bdir] -= h
        _, gp = oracle.basis_on_grid(b, dp)
        _, gm = oracle.basis_on_grid(b, dm)
        fd = (gp - gm) / (2 * h)  # d/dr_b of d chi/dr_a
        for a in range(3):
            got = hess[:, :, idx[tuple(sorted((a, bdir)))]]
            np.testing.assert_allclose(got, fd[:, :, a], atol=2e-7 * max(1.0, np.abs(fd).max()))


def test_xc_gradient_vs_fixed_grid_fd():
    """The analytic XC gradient (no grid-weight response) is the derivative of E_xc with the grid
    held fixed while the basis functions move with their atoms: central differences, h = 1e-4."""
    from helpers import random_density
    b = bench_molecule("H2O", "6-31g*")
    pts, w, _ = oracle.becke_grid(b.Z, b.xyz, 40, 0, 1)
    D = random_density(b.nbf, seed=2, nocc=b.n_occ)
    g = oracle.xc_gradient(b, pts, w, D)
    h = 1e-4
    fd = np.zeros_like(g)
    for a in range(b.natom):
        for x in range(3):
            fd[a, x] = (_exc_fixed_grid(_displaced(b, a, x, h), pts, w, D)
                        - _exc_fixed_grid(_displaced(b, a, x, -h), pts, w, D)) / (2 * h)
    np.testing.assert_allclose(g, fd, atol=2e-7)
    # a dropped Hessian term or a wrong sign shows up at the 1e-2 level
    assert np.abs(fd).max() > 1e-2
