"""CPU checks of the C-ABI boundary: libpdot_hip.so loads and exports every
entry point include/pdot.h declares, with the struct layouts the Python
binding uses (no compute call; there is no GPU here)."""
import ctypes
import os
import re

from pdot_amd import _lib

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    text = open(os.path.join(REPO, "include", "pdot.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pdot_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    declared = _declared_functions()
    assert len(declared) >= 15
    for name in declared:
        assert hasattr(lib, name), name


def test_binding_covers_header():
    assert set(_declared_functions()) == set(_lib.SIGNATURES)


def test_struct_sizes_match_header_layout():
    # pdot_solver_state: 4 doubles, 4 int64, 8 doubles, 8 int32, 1 int64, 2 int32, 2 doubles
    assert ctypes.sizeof(_lib.SolverState) == 4 * 8 + 4 * 8 + 8 * 8 + 8 * 4 + 8 + 2 * 4 + 2 * 8
    assert _lib.SolverState.pend_aw.offset == 184
    assert ctypes.sizeof(_lib.PointStats) == 11 * 8


def test_missing_library_fails_loudly(tmp_path):
    import pytest
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        _lib.load_library(str(tmp_path / "absent.so"))
