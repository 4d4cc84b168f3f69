"""The C-ABI library loads and exports every symbol include/magtense_hip.h declares (no GPU calls)."""

import ctypes
import os
import re

import pytest

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


def test_split_row_offsets():
    """Byte offset of a row block of the fragment-tiled split tensor: 24 bytes
    (6 components x hi/lo x 2 B) per padded source per receiver row; row blocks
    start on 32-row tiles (host arithmetic only)."""
    from magtense_amd import _lib

    lib = _lib.load()
    lib.mt_split_tensor_row_offset.restype = ctypes.c_longlong
    lib.mt_split_tensor_bytes.restype = ctypes.c_longlong
    n = 32768
    assert lib.mt_split_tensor_row_offset(n, 0) == 0
    assert lib.mt_split_tensor_row_offset(n, 4096) * 8 == lib.mt_split_tensor_bytes(n)
    assert lib.mt_split_tensor_row_offset(540, 512) == 512 * 544 * 24
    assert lib.mt_split_tensor_bytes(540) == 544 * 544 * 24
    for bad in ((540, 33), (540, 544), (540, -32), (0, 0)):
        assert lib.mt_split_tensor_row_offset(*bad) == -1


def test_status_codes_match_library():
    """The header's MT_ERR_* values equal the library's internal status enum."""
    hdr = open(os.path.join(ROOT, "include", "magtense_hip.h")).read()
    src = open(os.path.join(ROOT, "magtense_amd", "csrc", "mt_common.h")).read()
    declared = {k: int(v) for k, v in re.findall(r"#define MT_(OK|ERR_\w+) \(?(-?\d+)\)?", hdr)}
    enum = {k: int(v) for k, v in re.findall(r"\bk(Ok|Err\w+) = (-?\d+)", src)}
    names = {"Ok": "OK", "ErrArg": "ERR_ARG", "ErrHip": "ERR_HIP", "ErrSolver": "ERR_SOLVER",
             "ErrAlloc": "ERR_ALLOC", "ErrState": "ERR_STATE", "ErrUnsupported": "ERR_UNSUPPORTED"}
    assert set(enum) == set(names)
    for k, v in enum.items():
        assert declared[names[k]] == v, k


@pytest.mark.parametrize("cname,pyname", [("mt_micromag_problem", "MtProblem"), ("mt_stats", "MtStats"),
                                          ("mt_tiles", "MtTiles")])
def test_struct_layouts_match_header(tmp_path, cname, pyname):
    """The ctypes mirrors have the size and field offsets the C compiler gives
    the header's structs (fields in the same order under the same names)."""
    import ctypes as C
    import shutil
    import subprocess

    from magtense_amd import _lib

    cc = shutil.which("gcc")
    if cc is None:
        pytest.skip("no C compiler")
    py = getattr(_lib, pyname)
    names = [f[0] for f in py._fields_]
    src = tmp_path / "layout.c"
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "magtense_hip.h"', "int main(void) {",
             f'  printf("%zu\\n", sizeof({cname}));']
    lines += [f'  printf("%zu\\n", offsetof({cname}, {nm}));' for nm in names]
    lines += ["  return 0;", "}"]
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    assert out[0] == C.sizeof(py), (cname, out[0], C.sizeof(py))
    for nm, off in zip(names, out[1:]):
        assert getattr(py, nm).offset == off, (cname, nm, off, getattr(py, nm).offset)
