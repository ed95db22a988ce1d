"""Basis-set construction (host side, FP64, plain data).

Mirrors qc-core's ``BasisSet::build(atoms, name)`` (crates/qc-core/src/basis/mod.rs:164-193):
shells are emitted atom by atom in the built-in table order, each shell keeps its
raw contraction coefficients (no contracted renormalisation, basis/primitives.rs:232-237),
and Cartesian components follow ``cartesian_components`` (integrals/cartesian.rs:113-141):
for i = l..0, for j = l-i..0, k = l-i-j  (xx, xy, xz, yy, yz, zz for d).
Coordinates are in bohr.  The flat structure-of-arrays layout is what the C ABI
(`qcg_set_basis`) takes.
"""
from __future__ import annotations

import json
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

ANGSTROM_TO_BOHR = 1.8897261245650618  # crates/qc-core/src/constants.rs:63
SYMBOL_TO_Z = {"H": 1, "He": 2, "Li": 3, "Be": 4, "B": 5, "C": 6, "N": 7, "O": 8, "F": 9, "Ne": 10}
Z_TO_SYMBOL = {v: k for k, v in SYMBOL_TO_Z.items()}

_DATA = None


def _tables():
    global _DATA
    if _DATA is None:
        _DATA = json.loads(Path(__file__).with_name("data").joinpath("basis_sets.json").read_text())
    return _DATA["bases"]


def normalize_basis_name(name: str) -> str:
    """Accept the same spellings as get_element_basis (basis/builtin.rs:97-116)."""
    s = name.lower().replace(" ", "").replace("_", "")
    alias = {
        "sto-3g": "sto-3g", "sto3g": "sto-3g",
        "6-31g": "6-31g", "631g": "6-31g",
        "6-31g*": "6-31g*", "631g*": "6-31g*", "6-31gs": "6-31g*", "631gs": "6-31g*",
        "6-31g(d)": "6-31g*", "631g(d)": "6-31g*",
        "6-31+g*": "6-31+g*", "631+g*": "6-31+g*", "6-31+gs": "6-31+g*", "631+gs": "6-31+g*",
        "6-31+g(d)": "6-31+g*", "631+g(d)": "6-31+g*",
    }
    if s not in alias:
        raise ValueError(f"unknown basis set: {name}")
    return alias[s]


def ncart(l: int) -> int:
    return (l + 1) * (l + 2) // 2


def cartesian_components(l: int):
    out = []
    for i in range(l, -1, -1):
        for j in range(l - i, -1, -1):
            out.append((i, j, l - i - j))
    return out


@dataclass
class BasisSet:
    """Flat shell list in bohr (structure of arrays)."""
    name: str
    Z: np.ndarray              # (natom,) int32
    xyz: np.ndarray            # (natom, 3) float64, bohr
    shell_l: np.ndarray        # (nshell,) int32
    shell_nprim: np.ndarray    # (nshell,) int32
    shell_prim_off: np.ndarray  # (nshell,) int32
    shell_atom: np.ndarray     # (nshell,) int32
    shell_center: np.ndarray   # (nshell, 3) float64
    shell_bf_off: np.ndarray   # (nshell,) int32
    exps: np.ndarray           # (nprim_total,) float64
    coefs: np.ndarray          # (nprim_total,) float64
    nbf: int = 0
    nelec: int = 0
    e_nuc: float = 0.0
    meta: dict = field(default_factory=dict)

    @property
    def nshell(self) -> int:
        return len(self.shell_l)

    @property
    def natom(self) -> int:
        return len(self.Z)

    @property
    def n_occ(self) -> int:
        return self.nelec // 2


def nuclear_repulsion(Z, xyz) -> float:
    """compute_nuclear_repulsion (basis/mod.rs:375-391): pairs closer than 1e-10 bohr are skipped."""
    e = 0.0
    n = len(Z)
    for i in range(n):
        for j in range(i + 1, n):
            r = float(np.sqrt(((xyz[i] - xyz[j]) ** 2).sum()))
            if r > 1e-10:
                e += float(Z[i]) * float(Z[j]) / r
    return e


def build_basis(atoms, basis_name: str, units: str = "bohr", charge: int = 0) -> BasisSet:
    """atoms: list of (symbol_or_Z, (x, y, z)).  Returns the flat BasisSet."""
    name = normalize_basis_name(basis_name)
    table = _tables()[name]
    Z = np.array([a[0] if isinstance(a[0], int) else SYMBOL_TO_Z[a[0]] for a in atoms], dtype=np.int32)
    xyz = np.array([a[1] for a in atoms], dtype=np.float64).reshape(-1, 3)
    if units.lower().startswith("ang"):
        xyz = xyz * ANGSTROM_TO_BOHR
    sl, snp, spo, sat, sc, sbo, ex, co = [], [], [], [], [], [], [], []
    bf = 0
    for ia, z in enumerate(Z):
        key = str(int(z))
        if key not in table:
            raise ValueError(f"element Z={z} not in basis {name}")
        for l, prims in table[key]:
            sl.append(l)
            snp.append(len(prims))
            spo.append(len(ex))
            sat.append(ia)
            sc.append(xyz[ia])
            sbo.append(bf)
            bf += ncart(l)
            for e, c in prims:
                ex.append(e)
                co.append(c)
    nelec = int(Z.sum()) - charge
    return BasisSet(
        name=name, Z=Z, xyz=xyz,
        shell_l=np.array(sl, np.int32), shell_nprim=np.array(snp, np.int32),
        shell_prim_off=np.array(spo, np.int32), shell_atom=np.array(sat, np.int32),
        shell_center=np.array(sc, np.float64).reshape(-1, 3), shell_bf_off=np.array(sbo, np.int32),
        exps=np.array(ex, np.float64), coefs=np.array(co, np.float64),
        nbf=bf, nelec=nelec, e_nuc=nuclear_repulsion(Z, xyz),
    )
