"""Write b3lyp_mpmath.json: closed-shell B3LYP5 energy density and its two derivatives at
seeded (rho, sigma) points, from 40-digit mpmath evaluations of the published functionals
(Slater; VWN 1980 paramagnetic "5" fit; Becke 1988 per spin channel; LYP in the Miehlich et al.
1989 closed form), with vrho and vsigma by mpmath's numerical differentiation of rho * exc.
Independent of the product's and the oracle's floating-point code; run from the repo root:
    python tests/golden/make_b3lyp_mpmath.py
"""
import json
from pathlib import Path

import numpy as np
from mpmath import asinh, atan, cbrt, diff, exp, log, mp, mpf, pi, sqrt

mp.dps = 40


def f_b3lyp(rho, sig):
    """rho * exc of B3LYP5 (0.08 Slater + 0.72 B88 + 0.19 VWN5 + 0.81 LYP)."""
    rho, sig = mpf(rho), mpf(sig)
    cx = mpf(3) / 4 * cbrt(3 / pi)
    f_slater = -cx * rho ** (mpf(4) / 3)
    a, x0, b, c = mpf("0.0310907"), mpf("-0.10498"), mpf("3.72744"), mpf("12.9352")
    x = sqrt(cbrt(3 / (4 * pi * rho)))
    X, X0, Q = x * x + b * x + c, x0 * x0 + b * x0 + c, sqrt(4 * c - b * b)
    at = atan(Q / (2 * x + b))
    f_vwn = rho * a * (log(x * x / X) + 2 * b / Q * at - b * x0 / X0 * (log((x - x0) ** 2 / X) + 2 * (b + 2 * x0) / Q * at))
    beta, rs, gs = mpf("0.0042"), rho / 2, sig / 4
    xs = sqrt(gs) / rs ** (mpf(4) / 3)
    f_b88 = 2 * (-cbrt(2) * cx * rs ** (mpf(4) / 3) - beta * rs ** (mpf(4) / 3) * xs ** 2 / (1 + 6 * beta * xs * asinh(xs)))
    la, lb, lc, ld = mpf("0.04918"), mpf("0.132"), mpf("0.2533"), mpf("0.349")
    cf = mpf(3) / 10 * (3 * pi ** 2) ** (mpf(2) / 3)
    u = rho ** (-mpf(1) / 3)
    om = exp(-lc * u) / (1 + ld * u) * rho ** (-mpf(11) / 3)
    dl = lc * u + ld * u / (1 + ld * u)
    f_lyp = -la * rho / (1 + ld * u) - la * lb * om * (cf * rho ** (mpf(14) / 3) - (3 + 7 * dl) / 72 * rho ** 2 * sig)
    return mpf("0.08") * f_slater + mpf("0.72") * f_b88 + mpf("0.19") * f_vwn + mpf("0.81") * f_lyp


def main():
    rng = np.random.default_rng(2026)
    rho = np.concatenate([rng.uniform(1e-4, 50, 20), 10.0 ** rng.uniform(-9, 1, 20)])
    sig = np.concatenate([rng.uniform(1e-6, 20, 20), 10.0 ** rng.uniform(-14, 2, 20)])
    out = {"source": __doc__.strip().splitlines()[0], "rho": rho.tolist(), "sigma": sig.tolist(),
           "exc": [], "vrho": [], "vsigma": []}
    for r, s in zip(rho.tolist(), sig.tolist()):
        out["exc"].append(float(f_b3lyp(r, s) / mpf(r)))
        out["vrho"].append(float(diff(lambda t: f_b3lyp(t, s), r)))
        out["vsigma"].append(float(diff(lambda t: f_b3lyp(r, t), s)))
    Path(__file__).with_name("b3lyp_mpmath.json").write_text(json.dumps(out, indent=1) + "\n")


if __name__ == "__main__":
    main()
