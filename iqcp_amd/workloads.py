"""Synthetic molecular workloads (seeded, reproducible), coordinates in angstrom.

- n-alkanes CnH2n+2: ideal all-trans geometry, C-C 1.54 A, C-H 1.09 A, tetrahedral
  angles, plus per-component Gaussian jitter N(0, sigma) from numpy default_rng(seed).
- (H2O)n clusters: rigid waters (O-H 0.957 A, H-O-H 104.5 deg) on an O lattice with
  seeded uniform random orientations.
- the 64-alkane batch: n_C drawn uniformly from 4..12 with a seeded generator.
"""
from __future__ import annotations

import numpy as np

from .basis import build_basis

TETRA = np.arccos(-1.0 / 3.0)  # 109.4712 deg


def _unit(v):
    return v / np.linalg.norm(v)


def alkane_ideal(n_c: int, r_cc: float = 1.54, r_ch: float = 1.09):
    """All-trans CnH2n+2 in angstrom: list of (symbol, xyz)."""
    if n_c < 1:
        raise ValueError("n_c >= 1")
    half = TETRA / 2.0
    dx = r_cc * np.sin(half)
    dy = r_cc * np.cos(half)
    # zigzag backbone in the xy plane; virtual atoms at both ends continue the zigzag
    cs = [np.array([i * dx, (dy if i % 2 else 0.0), 0.0]) for i in range(-1, n_c + 1)]
    atoms = [("C", cs[i + 1]) for i in range(n_c)]
    hs = []
    for i in range(n_c):
        c = cs[i + 1]
        prev_, next_ = cs[i], cs[i + 2]
        u1 = _unit(prev_ - c)
        u2 = _unit(next_ - c)
        if n_c == 1:
            # methane: tetrahedron
            for v in ([1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]):
                hs.append(c + r_ch * _unit(np.array(v, float)))
            continue
        b = _unit(-(u1 + u2))
        nrm = _unit(np.cross(u1, u2))
        h_dirs = [np.cos(half) * b + np.sin(half) * nrm, np.cos(half) * b - np.sin(half) * nrm]
        if i == 0:
            h_dirs.append(u1)          # in-plane H where the previous carbon would be
        if i == n_c - 1:
            h_dirs.append(u2)
        for d in h_dirs:
            hs.append(c + r_ch * d)
    atoms += [("H", h) for h in hs]
    return atoms


def jitter(atoms, sigma: float, seed: int):
    rng = np.random.default_rng(seed)
    out = []
    for sym, xyz in atoms:
        out.append((sym, np.asarray(xyz, float) + rng.normal(0.0, sigma, 3)))
    return out


def alkane(n_c: int, sigma: float, seed: int):
    return jitter(alkane_ideal(n_c), sigma, seed)


def _random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def water_cluster(shape=(4, 2, 2), spacing: float = 2.9, seed: int = 0,
                  r_oh: float = 0.957, angle_deg: float = 104.5):
    rng = np.random.default_rng(seed)
    th = np.radians(angle_deg) / 2.0
    h1 = r_oh * np.array([np.sin(th), 0.0, np.cos(th)])
    h2 = r_oh * np.array([-np.sin(th), 0.0, np.cos(th)])
    atoms = []
    for ix in range(shape[0]):
        for iy in range(shape[1]):
            for iz in range(shape[2]):
                o = spacing * np.array([ix, iy, iz], float)
                r = _random_rotation(rng)
                atoms.append(("O", o))
                atoms.append(("H", o + r @ h1))
                atoms.append(("H", o + r @ h2))
    return atoms


# ---- the named workloads (sizes stated in their names) ----

def propane_sto3g(seed: int = 1):
    return build_basis(alkane(3, 0.02, seed), "sto-3g", units="angstrom")


def decane_631gs(seed: int = 2):
    return build_basis(alkane(10, 0.02, seed), "6-31g*", units="angstrom")


def water16_631gs(seed: int = 3):
    return build_basis(water_cluster((4, 2, 2), 2.9, seed), "6-31g*", units="angstrom")


def octane_631pgs(seed: int = 5):
    return build_basis(alkane(8, 0.05, seed), "6-31+g*", units="angstrom")


def alkane_batch_631g(n_mol: int = 64, seed: int = 4):
    rng = np.random.default_rng(seed)
    sizes = rng.integers(4, 13, size=n_mol)
    return [build_basis(alkane(int(n), 0.03, seed * 1000 + k), "6-31g", units="angstrom")
            for k, n in enumerate(sizes)]
