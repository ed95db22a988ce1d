"""Shared helpers for the tests (molecules from the committed fixtures)."""
import json
from pathlib import Path

import numpy as np

from iqcp_amd.basis import build_basis

GOLDEN = Path(__file__).resolve().parent / "golden"


def preset(name):
    d = json.loads((GOLDEN / "presets" / f"{name}.json").read_text())
    atoms = [(a["symbol"], a["xyz"]) for a in d["geometry"]["atoms"]]
    zsum = sum({"H": 1, "He": 2, "Li": 3, "N": 7, "O": 8}[a[0]] for a in atoms)
    return d, build_basis(atoms, "sto-3g", charge=zsum - d["nelec"])


def bench_molecule(mol, basis):
    geo = json.loads((GOLDEN / "benchmark_geometries.json").read_text())
    return build_basis([(z, x) for z, x in geo[mol]], basis)


def iqcp_energy(mol, basis, method):
    for r in json.loads((GOLDEN / "iqcp_energies.json").read_text())["results"]:
        if r["molecule"] == mol and r["basis"] == basis and r["method"] == method:
            return r["energy"]
    raise KeyError((mol, basis, method))


def random_density(nbf, seed=0, nocc=None):
    """A symmetric positive semi-definite 'density' 2 C C^T from random orthonormal columns."""
    rng = np.random.default_rng(seed)
    nocc = nocc or max(1, nbf // 3)
    q, _ = np.linalg.qr(rng.normal(size=(nbf, nbf)))
    c = q[:, :nocc]
    return 2.0 * c @ c.T
