"""The C-ABI library loads and exports every symbol include/mcvslam_gpu.h
declares (no compute calls: runs without a GPU)."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def declared():
    text = (ROOT / "include" / "mcvslam_gpu.h").read_text()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mcv_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert "mcv_orb_extract" in names and "mcv_match_knn2" in names and "mcv_ba_linearize" in names


def test_library_exports_all(tmp_path):
    from mcvslam_amd import _native
    if not _native.LIB_PATH.exists():
        pytest.fail(f"{_native.LIB_PATH} not built: run `python -m mcvslam_amd.build`")
    lib = ctypes.CDLL(str(_native.LIB_PATH))
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding covers every declared entry point
    assert set(declared()) <= set(_native.SIGNATURES), set(declared()) - set(_native.SIGNATURES)


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from mcvslam_amd import _native
    monkeypatch.setattr(_native, "LIB_PATH", tmp_path / "nope.so")
    monkeypatch.setattr(_native, "_lib", None)
    with pytest.raises(RuntimeError, match="missing"):
        _native.lib()
