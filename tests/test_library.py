"""The C-ABI library: builds for gfx950, loads, and exports every declared symbol.

No compute is called here (no GPU needed).
"""

import ctypes
import re
from pathlib import Path

import pytest

from jaxiga_amd import _lib

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "jaxiga_hip.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|void)\s+(iga_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_bound_symbols():
    assert declared_symbols() == _lib.exported_symbols()


def test_library_loads_and_exports_every_declared_symbol():
    if not _lib.LIB_PATH.exists():
        pytest.skip("library not built")
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "_LIB", None)
    monkeypatch.setenv("JAXIGA_HIP_LIB", str(tmp_path / "absent.so"))
    with pytest.raises(RuntimeError, match="missing"):
        _lib.load()
