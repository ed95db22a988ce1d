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
