"""The reference's preset-system format and the GPU path on it.

A preset system carries S, H = T + V and the 8-fold compressed two-electron tensor of a molecule
instead of its basis (``PresetSystem``, crates/qc-core/src/scf/mod.rs:284; JSON form
``PresetSystemJson``, :464-583; files under content/presets/systems/).  This module

- reads and validates that JSON exactly as ``PresetSystemJson::from_json`` +
  ``to_preset_system`` do (same required fields, same checks in the same order, same messages);
- writes it, from a basis, with S, H and the tensor computed on the GPU (``from_basis``), so the
  reference can load what this package produces;
- runs ``build_fock`` and ``rhf_scf`` on a preset's stored tensor on the GPU (``StoredEngine``,
  C ABI ``qcg_build_fock_stored`` / ``qcg_rhf_scf_stored``).
"""
from __future__ import annotations

import ctypes
import json
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from .basis import Z_TO_SYMBOL, BasisSet
from .engine import Engine, QcgError, ScfConfig, ScfResult, _d, lib

FORMAT_VERSION = 1  # PresetSystemJson::FORMAT_VERSION (scf/mod.rs:501)
ERI_INDEXING = "8-fold symmetry: pair(i,j)=i*(i+1)/2+j, idx(P,Q)=P*(P+1)/2+Q"

# fields without #[serde(default)] in PresetSystemJson (scf/mod.rs:464-495)
_REQUIRED = ("format_version", "system_id", "label", "basis_id", "nbf", "nelec", "e_nuc", "s_matrix", "h_core",
             "eri_compressed")


def eri_length(nbf: int) -> int:
    npair = nbf * (nbf + 1) // 2
    return npair * (npair + 1) // 2


@dataclass
class PresetSystem:
    """PresetSystemJson (scf/mod.rs:464): matrices row-major nbf x nbf, the tensor flat."""
    system_id: str
    label: str
    basis_id: str
    nbf: int
    nelec: int
    e_nuc: float
    s_matrix: np.ndarray
    h_core: np.ndarray
    eri_compressed: np.ndarray
    format_version: int = FORMAT_VERSION
    description: str = ""
    geometry: Optional[dict] = None
    eri_indexing: str = ""
    reference: Optional[dict] = None
    extra: dict = field(default_factory=dict)

    def validate(self):
        """to_preset_system (scf/mod.rs:526-569): format version, S, H, ERI sizes; ValueError."""
        if self.format_version != FORMAT_VERSION:
            raise ValueError(f"Unsupported format version: {self.format_version} (expected {FORMAT_VERSION})")
        n2 = self.nbf * self.nbf
        if self.s_matrix.size != n2:
            raise ValueError(f"S matrix size mismatch: expected {n2}, got {self.s_matrix.size}")
        if self.h_core.size != n2:
            raise ValueError(f"H_core matrix size mismatch: expected {n2}, got {self.h_core.size}")
        if self.eri_compressed.size != eri_length(self.nbf):
            raise ValueError(f"ERI array size mismatch: expected {eri_length(self.nbf)}, "
                             f"got {self.eri_compressed.size}")
        return self

    def reference_energy(self) -> Optional[float]:
        return None if self.reference is None else float(self.reference["energy"])

    def to_dict(self) -> dict:
        d = {"format_version": self.format_version, "system_id": self.system_id, "label": self.label,
             "description": self.description}
        if self.geometry is not None:
            d["geometry"] = self.geometry
        d.update({"basis_id": self.basis_id, "nbf": self.nbf, "nelec": self.nelec, "e_nuc": self.e_nuc,
                  "s_matrix": self.s_matrix.ravel().tolist(), "h_core": self.h_core.ravel().tolist(),
                  "eri_compressed": self.eri_compressed.tolist(), "eri_indexing": self.eri_indexing})
        if self.reference is not None:
            d["reference"] = self.reference
        return d

    def to_json(self) -> str:
        return json.dumps(self.to_dict())


def from_json(text: str) -> PresetSystem:
    """PresetSystemJson::from_json + to_preset_system (scf/mod.rs:515-569)."""
    try:
        d = json.loads(text)
    except json.JSONDecodeError as e:
        raise ValueError(f"JSON parse error: {e}") from None
    if not isinstance(d, dict):
        raise ValueError("JSON parse error: expected an object")
    for k in _REQUIRED:
        if k not in d:
            raise ValueError(f"JSON parse error: missing field `{k}`")
    nbf = int(d["nbf"])
    if nbf < 0 or int(d["nelec"]) < 0:
        raise ValueError("JSON parse error: nbf and nelec are unsigned")
    known = set(_REQUIRED) | {"description", "geometry", "eri_indexing", "reference"}
    p = PresetSystem(
        system_id=str(d["system_id"]), label=str(d["label"]), basis_id=str(d["basis_id"]), nbf=nbf,
        nelec=int(d["nelec"]), e_nuc=float(d["e_nuc"]),
        s_matrix=np.asarray(d["s_matrix"], dtype=np.float64).ravel(),
        h_core=np.asarray(d["h_core"], dtype=np.float64).ravel(),
        eri_compressed=np.asarray(d["eri_compressed"], dtype=np.float64).ravel(),
        format_version=int(d["format_version"]), description=str(d.get("description", "")),
        geometry=d.get("geometry"), eri_indexing=str(d.get("eri_indexing", "")), reference=d.get("reference"),
        extra={k: v for k, v in d.items() if k not in known})
    p.validate()
    if p.s_matrix.size == nbf * nbf:
        p.s_matrix = p.s_matrix.reshape(nbf, nbf)
        p.h_core = p.h_core.reshape(nbf, nbf)
    return p


def from_basis(basis: BasisSet, system_id: str, label: str, description: str = "", reference=None,
               engine: Optional[Engine] = None) -> PresetSystem:
    """A preset of the molecule of ``basis`` with S, H and the compressed tensor computed on the GPU
    (``overlap_matrix`` / ``hcore_matrix`` / ``eri_compressed``), geometry in bohr."""
    eng = engine if engine is not None else Engine(basis)
    S, T, V = eng.one_electron()
    eri = eng.eri_compressed()
    atoms = [{"symbol": Z_TO_SYMBOL[int(z)], "xyz": [float(v) for v in x]} for z, x in zip(basis.Z, basis.xyz)]
    return PresetSystem(system_id=system_id, label=label, basis_id=basis.name, nbf=basis.nbf, nelec=basis.nelec,
                        e_nuc=float(basis.e_nuc), s_matrix=S, h_core=T + V, eri_compressed=eri,
                        description=description, geometry={"atoms": atoms, "units": "bohr"},
                        eri_indexing=ERI_INDEXING, reference=reference)


class StoredEngine:
    """A qcg context without a basis, for the stored-tensor path of preset systems."""

    def __init__(self, device: int = 0):
        self.L = lib()
        self.ctx = ctypes.c_void_p()
        rc = self.L.qcg_create(ctypes.byref(self.ctx), device)
        if rc:
            raise QcgError(rc, "qcg_create failed (no GPU?)")

    def _check(self, rc):
        if rc:
            raise QcgError(rc, self.L.qcg_last_error(self.ctx).decode())

    def close(self):
        if self.ctx:
            self.L.qcg_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def build_fock(self, H, D, eri) -> np.ndarray:
        """build_fock(h, d, eri, nbf) (scf/mod.rs:825) on the GPU."""
        H = np.ascontiguousarray(H, dtype=np.float64)
        D = np.ascontiguousarray(D, dtype=np.float64)
        eri = np.ascontiguousarray(eri, dtype=np.float64).ravel()
        nbf = H.shape[0]
        if H.shape != (nbf, nbf) or D.shape != (nbf, nbf):
            raise ValueError("H and D must be nbf x nbf")
        F = np.zeros_like(H)
        self._check(self.L.qcg_build_fock_stored(self.ctx, nbf, _d(H), _d(D), _d(eri), ctypes.c_int64(eri.size),
                                                 _d(F)))
        return F

    def rhf_scf(self, preset: PresetSystem, cfg: ScfConfig = None, D_init=None) -> dict:
        """rhf_scf(&PresetSystem, &ScfConfig) (scf/mod.rs:1363) with the Fock builds on the GPU."""
        cfg = cfg or ScfConfig.tight()
        n = preset.nbf
        S = np.ascontiguousarray(preset.s_matrix, dtype=np.float64).ravel()
        H = np.ascontiguousarray(preset.h_core, dtype=np.float64).ravel()
        eri = np.ascontiguousarray(preset.eri_compressed, dtype=np.float64).ravel()
        if S.size != n * n or H.size != n * n:
            raise ValueError("S and H must hold nbf * nbf values")
        D, C, F = np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, n))
        eps = np.zeros(n)
        res = ScfResult()
        Di = None if D_init is None else _d(np.ascontiguousarray(D_init, dtype=np.float64))
        self._check(self.L.qcg_rhf_scf_stored(self.ctx, n, int(preset.nelec), ctypes.c_double(preset.e_nuc), _d(S),
                                              _d(H), _d(eri), ctypes.c_int64(eri.size), ctypes.byref(cfg), Di,
                                              ctypes.byref(res), _d(D), _d(C), _d(eps), _d(F)))
        out = res.as_dict()
        out.update(D=D, C=C, eps=eps, F=F)
        return out
