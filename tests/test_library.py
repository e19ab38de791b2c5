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
    return sorted(set(re.findall(r"^\s*(?:int|int32_t|int64_t|void)\s+(iga_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_bound_symbols():
    assert declared_symbols() == _lib.exported_symbols()


def test_library_loads_and_exports_every_declared_symbol():
    if not _lib.LIB_PATH.exists():
        pytest.skip("library not built")
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_batched_block_plan():
    """The segmented batched solve's block map (a host-only function): rows cut
    into blocks of at most rows_per_block rows, an empty instance keeps one
    block."""
    if not _lib.LIB_PATH.exists():
        pytest.skip("library not built")
    import numpy as np

    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    lib.iga_batched_cg_seg_plan.restype = ctypes.c_int32
    n = np.array([0, 1, 256, 257, 1000], np.int32)
    nb = lib.iga_batched_cg_seg_plan(5, ctypes.c_void_p(n.ctypes.data), 256, None, None, None)
    assert nb == 1 + 1 + 1 + 2 + 4
    bi, br, b0 = np.full(nb, -1, np.int32), np.full(nb, -1, np.int32), np.full(6, -1, np.int32)
    assert lib.iga_batched_cg_seg_plan(5, ctypes.c_void_p(n.ctypes.data), 256, ctypes.c_void_p(bi.ctypes.data),
                                       ctypes.c_void_p(br.ctypes.data), ctypes.c_void_p(b0.ctypes.data)) == nb
    np.testing.assert_array_equal(b0, [0, 1, 2, 3, 5, 9])
    np.testing.assert_array_equal(bi, [0, 1, 2, 3, 3, 4, 4, 4, 4])
    np.testing.assert_array_equal(br, [0, 0, 0, 0, 256, 0, 256, 512, 768])
    assert lib.iga_batched_cg_seg_plan(1, ctypes.c_void_p(n.ctypes.data), 0, None, None, None) < 0


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "_LIB", None)
    monkeypatch.setenv("JAXIGA_HIP_LIB", str(tmp_path / "absent.so"))
    with pytest.raises(RuntimeError, match="missing"):
        _lib.load()
