"""CPU checks of the product library: it is built, loads, and exports every C-ABI symbol."""
import re
from pathlib import Path

import pytest

from iqcp_amd import engine

ROOT = Path(__file__).resolve().parent.parent


def test_header_declares_exports():
    hdr = (ROOT / "include" / "qcg.h").read_text()
    declared = set(re.findall(r"\b(qcg_[a-z0-9_]+)\s*\(", hdr))
    assert declared == set(engine.EXPORTS), declared ^ set(engine.EXPORTS)


def test_library_exports_all_symbols():
    L = engine.lib()
    for name in engine.EXPORTS:
        assert hasattr(L, name), name


def test_no_fallback_when_library_missing(monkeypatch, tmp_path):
    monkeypatch.setattr(engine, "LIB_PATH", tmp_path / "missing.so")
    monkeypatch.setattr(engine, "_LIB", None)
    with pytest.raises(RuntimeError):
        engine.lib()
