"""The engine library loads and binds every entry point include/meep_fdtd.h declares (CPU:
no compute call is made)."""

import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "meep_fdtd.h")).read()
    return sorted(set(re.findall(r"\b(fdtd_[a-z_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported_and_bound():
    from meep_gpu import _native

    lib = _native.load()
    declared = _declared()
    assert declared, "no declarations found"
    for name in declared:
        assert hasattr(lib, name), name
    assert set(declared) == set(_native.SIGNATURES), (
        set(declared) ^ set(_native.SIGNATURES))


def test_missing_library_fails_loudly(tmp_path):
    import pytest
    from meep_gpu import _native

    saved = _native._lib
    _native._lib = None
    try:
        with pytest.raises(_native.EngineError, match="not built"):
            _native.load(str(tmp_path / "absent.so"))
    finally:
        _native._lib = saved
