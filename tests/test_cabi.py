"""The C-ABI library loads and exports every symbol include/magtense_hip.h declares (no GPU calls)."""

import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "magtense_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b((?:mt|icl)_[A-Za-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_symbols()
    for n in ("mt_run_micromag", "mt_prism_tensor", "mt_demag_apply_bf16x2", "icl_MVMult_GetH"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from magtense_amd import _lib

    lib = _lib.load()
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(declared_symbols()) <= set(_lib.EXPORTS)


def test_no_gpu_fails_loudly():
    """Without a visible device the product path raises instead of falling back to the CPU."""
    from magtense_amd import _lib

    lib = _lib.load()
    if lib.mt_device_count() > 0:
        return
    import pytest

    with pytest.raises(_lib.MagTenseHipError):
        _lib.lib()
    rc = lib.mt_demag_apply_f32(None, 4, None, None, 1, ctypes.c_float(-1.0))
    assert rc != 0
    assert b"GPU" in lib.mt_last_error()
