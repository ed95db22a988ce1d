"""Preset-system JSON (PresetSystemJson, scf/mod.rs:464-583): parsing and validation as the
reference does them, on its own preset files (content/presets/systems, copied as fixtures)."""
import json

import numpy as np
import pytest

from helpers import GOLDEN
from iqcp_amd import presets

NAMES = ["h2_sto3g_r1.4", "heh_plus_sto3g", "lih_sto3g", "h2o_sto3g", "nh3_sto3g"]


def _text(name):
    return (GOLDEN / "presets" / f"{name}.json").read_text()


@pytest.mark.parametrize("name", NAMES)
def test_reference_presets_parse_and_round_trip(name):
    p = presets.from_json(_text(name))
    d = json.loads(_text(name))
    assert p.nbf == d["nbf"] and p.nelec == d["nelec"] and p.e_nuc == d["e_nuc"]
    assert p.s_matrix.shape == (p.nbf, p.nbf)
    assert p.eri_compressed.size == presets.eri_length(p.nbf)
    assert p.reference_energy() == d["reference"]["energy"]
    q = presets.from_json(p.to_json())
    for k in ("system_id", "label", "basis_id", "nbf", "nelec", "e_nuc", "description", "geometry", "eri_indexing",
              "reference"):
        assert getattr(q, k) == getattr(p, k), k
    for k in ("s_matrix", "h_core", "eri_compressed"):
        assert np.array_equal(getattr(q, k), getattr(p, k)), k


def _bad(mutate):
    d = json.loads(_text("h2_sto3g_r1.4"))
    mutate(d)
    with pytest.raises(ValueError) as e:
        presets.from_json(json.dumps(d))
    return str(e.value)


def test_validation_messages_follow_to_preset_system():
    """to_preset_system (scf/mod.rs:526-569): version, then S, H and ERI sizes, in that order."""
    assert _bad(lambda d: d.update(format_version=2)) == "Unsupported format version: 2 (expected 1)"
    assert _bad(lambda d: d.update(s_matrix=[1.0, 0.5, 0.5])) == "S matrix size mismatch: expected 4, got 3"
    assert _bad(lambda d: d.update(h_core=[1.0])) == "H_core matrix size mismatch: expected 4, got 1"
    assert _bad(lambda d: d.update(eri_compressed=[0.0] * 5)) == "ERI array size mismatch: expected 6, got 5"
    # the version is checked before the sizes
    assert _bad(lambda d: d.update(format_version=0, h_core=[])).startswith("Unsupported format version")
    assert _bad(lambda d: d.pop("e_nuc")).startswith("JSON parse error")
    with pytest.raises(ValueError, match="JSON parse error"):
        presets.from_json("{not json")


def test_optional_fields_default():
    d = json.loads(_text("h2_sto3g_r1.4"))
    for k in ("description", "geometry", "eri_indexing", "reference"):
        d.pop(k, None)
    p = presets.from_json(json.dumps(d))
    assert p.description == "" and p.geometry is None and p.reference is None and p.reference_energy() is None


def test_eri_length():
    assert [presets.eri_length(n) for n in (1, 2, 7)] == [1, 6, 406]
