"""The C-ABI library loads and exports every symbol its header declares (CPU)."""
import ctypes
import re
from pathlib import Path

import pytest

from tccg_amd import build
from tccg_amd.plan import TILES

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    text = (ROOT / "include" / "tccg_gfx950.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(tccg_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ["tccg_gett_f32", "tccg_gett_f64", "tccg_gett_bf16x3", "tccg_gett_fp8", "tccg_amax_f32",
              "tccg_tile_shape", "tccg_num_tiles", "tccg_last_error"]:
        assert n in names


def test_library_exports_declared_symbols():
    path = build.library_path()
    if not path.exists():
        pytest.skip("library not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(str(path))
    for n in declared_functions():
        assert hasattr(lib, n), n


def test_tile_tables_agree_and_errors_reported():
    path = build.library_path()
    if not path.exists():
        pytest.skip("library not built")
    from tccg_amd.runtime import load_library
    lib = load_library()   # checks the host tile table against the library's
    for prec, tiles in TILES.items():
        assert lib.tccg_num_tiles(prec) == len(tiles)
    bm = ctypes.c_int()
    assert lib.tccg_tile_shape(0, 99, ctypes.byref(bm), None, None) != 0
    assert b"invalid tile" in lib.tccg_last_error()
    # a NULL plan is rejected before any launch (no GPU needed)
    assert lib.tccg_gett_f32(None, None, None, 1.0, 0.0, None, None) != 0
    assert b"plan" in lib.tccg_last_error()
