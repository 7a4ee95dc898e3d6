"""The C-ABI library loads and exports every symbol its header declares (CPU)."""
import ctypes
import re
from pathlib import Path

import pytest

from tccg_amd import build
from tccg_amd.plan import PACKED_TILES, TILES

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    text = (ROOT / "include" / "tccg_gfx950.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|size_t|const char\*)\s+(tccg_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ["tccg_gett_f32", "tccg_gett_f64", "tccg_gett_bf16x3", "tccg_gett_fp8", "tccg_amax_f32",
              "tccg_tile_shape", "tccg_num_tiles", "tccg_last_error", "tccg_packed_workspace_bytes",
              "tccg_gett_packed_bf16x3", "tccg_gett_packed_f32", "tccg_num_packed_tiles"]:
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
    for prec, tiles in PACKED_TILES.items():
        assert lib.tccg_num_packed_tiles(prec) == len(tiles)
    assert lib.tccg_packed_tile_shape(2, 99, ctypes.byref(bm), None, None) != 0
    assert b"invalid packed tile" in lib.tccg_last_error()
    assert lib.tccg_gett_packed_bf16x3(None, None, None, 1.0, 0.0, None, None, 0, None) != 0
    assert b"plan" in lib.tccg_last_error()


def test_packed_workspace_bytes_agree():
    """tccg_packed_workspace_bytes() and GettPlan.workspace_bytes() agree (no GPU
    call: the size is computed on the host from the plan's lengths)."""
    path = build.library_path()
    if not path.exists():
        pytest.skip("library not built")
    from tccg_amd import Contraction
    from tccg_amd.plan import build_packed_plan
    from tccg_amd.runtime import _CPlan, load_library
    from tccg_amd.workloads import get
    lib = load_library()
    for name in ("abcd-aebf-dfce", "abcdef-dgfa-gceb"):
        c = Contraction.from_tccg(get(name))
        for prec, tiles in PACKED_TILES.items():
            for t in range(len(tiles)):
                p = build_packed_plan(c.problem, prec, t)
                cp = _CPlan()
                cp.Mp, cp.Np, cp.Kp, cp.tile = p.Mp, p.Np, p.Kp, p.tile
                assert lib.tccg_packed_workspace_bytes(prec, ctypes.byref(cp)) == p.workspace_bytes() > 0
