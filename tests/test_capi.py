"""The C-ABI library loads and exports every entry point include/karman.h
declares (no compute calls: this runs without a GPU)."""
import ctypes
import re
from pathlib import Path

import pytest

from karman_taichi_amd import _lib

ROOT = Path(__file__).resolve().parents[1]


def header_functions():
    text = (ROOT / "include" / "karman.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^[A-Za-z_][\w\s\*]*?\b(km_\w+)\s*\(", text, flags=re.M)))


def test_header_parses():
    names = header_functions()
    assert "km_bicgstab" in names and "km_assemble_u" in names and len(names) >= 20


def test_library_exports_every_declared_symbol():
    if not _lib.LIB_PATH.exists():
        pytest.fail("libkarman_hip.so is not built (python -m karman_taichi_amd.build)")
    h = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in header_functions() if not hasattr(h, n)]
    assert not missing, missing


def test_binding_matches_header():
    assert sorted(_lib.exported_symbols()) == header_functions()


def test_struct_layouts_match_header():
    # km_grid: 4 int32 + 8 doubles + 8 doubles = 16 + 64 + 64 bytes
    assert ctypes.sizeof(_lib.KmGrid) == 16 + 8 * 8 + 8 * 8
    assert _lib.KmGrid.bc.offset == 16 + 8 * 8
    assert ctypes.sizeof(_lib.KmSolveInfo) == 16 + 3 * 8


def test_case_struct_layout_matches_header():
    # km_case: km_grid (144 bytes) + 16 pointers
    assert ctypes.sizeof(_lib.KmCase) == ctypes.sizeof(_lib.KmGrid) + 16 * 8
    assert _lib.KmCase.mask.offset == ctypes.sizeof(_lib.KmGrid) + 15 * 8


def test_noalloc_limit_matches_header():
    text = (ROOT / "include" / "karman.h").read_text()
    m = re.search(r"#define KM_NOALLOC_MAX_CELLS (\d+)", text)
    assert m and int(m.group(1)) == _lib.NOALLOC_MAX_CELLS
