"""The engine library loads and binds every entry point include/meep_fdtd.h declares (CPU:
no compute call is made)."""

import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "meep_fdtd.h")).read()
    return sorted(set(re.findall(r"\b(fdtd_[a-z_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported_and_bound():
    """Both builds (subnormals kept, flushed) export every declaration and report their
    policy (fdtd_subnormal_mode is a build-time constant: no GPU call)."""
    from meep_gpu import _native

    declared = _declared()
    assert declared, "no declarations found"
    assert set(declared) == set(_native.SIGNATURES), (
        set(declared) ^ set(_native.SIGNATURES))
    for mode in _native.SUBNORMAL_MODES:
        lib = _native.load(subnormals=mode)
        for name in declared:
            assert hasattr(lib, name), (mode, name)
        assert lib.fdtd_subnormal_mode() == _native.SUBNORMAL_MODES.index(mode)
    assert _native.load(subnormals="flush") is not _native.load()


def test_missing_library_fails_loudly(tmp_path):
    import pytest
    from meep_gpu import _native

    with pytest.raises(_native.EngineError, match="not built"):
        _native.load(str(tmp_path / "absent.so"))
    with pytest.raises(ValueError, match="subnormals"):
        _native.load(subnormals="ftz")
