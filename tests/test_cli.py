"""The tccg command line and the gfx950 code generator (CPU: generate-only)."""
import os
import subprocess
import sys

import pytest

from tccg_amd import Contraction, codegen
from tccg_amd.cli import candidate_name, main
from tccg_amd.plan import F32, F64, FP8
from tccg_amd.workloads import WORKLOADS

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def decode_name(variant):
    """The reference's Gett.decodeName (tccg/gett.py:561-576), restated."""
    tokens = variant.split("_")
    return (int(tokens[1][3:]), int(tokens[2][2:]), int(tokens[3][2:]), int(tokens[4][2:]), int(tokens[5][3:]),
            int(tokens[6][3:]), int(tokens[7][2:]), int(tokens[8][2:]), "_".join(tokens[9:]))


def test_candidate_names_decode():
    c = Contraction.from_tccg(WORKLOADS["abcd-aebf-dfce"])
    p = c.plan(F32)
    var, mc, nc, kc, mc1, nc1, mr, nr, idx = decode_name(candidate_name(p))
    assert (mc, nc, kc) == (p.BM, p.BN, p.BK)
    assert idx == "aebf_dfce_abcd"


def test_cli_generate_only(tmp_path):
    f = tmp_path / "example.tccg"
    f.write_text(WORKLOADS["example"])
    out = tmp_path / "impl"
    rc = main(["--arch=gfx950", "--floatType=s", "--generateOnly", "--output", str(out), str(f)])
    assert rc == 0
    assert (out / "gett.hip").exists() and (out / "gett.hpp").exists() and (out / "libgett.so").exists()
    src = (out / "gett.hip").read_text()
    assert 'extern "C"' in src and "void gett(float* A, float* B, float* C, const float alpha, const float beta)" in src
    syms = subprocess.run(["nm", "-D", str(out / "libgett.so")], stdout=subprocess.PIPE, text=True).stdout
    for s in ("gett", "gett_stream", "gett_info", "gett_contraction", "gett_last_error"):
        assert (" T %s\n" % s) in syms


def test_cli_double_header(tmp_path):
    f = tmp_path / "c.tccg"
    f.write_text("C[a,b] = A[a,k] * B[k,b]\na=8\nb=8\nk=8\n")
    rc = main(["--floatType=d", "--generateOnly", "--output", str(tmp_path / "o"), str(f)])
    assert rc == 0
    assert "void gett(double* A, double* B, double* C, const double alpha, const double beta);" in \
        (tmp_path / "o" / "gett.hpp").read_text()


@pytest.mark.parametrize("argv,err", [
    (["--arch=avx2"], "architecture"),
    (["--floatType=z"], "floatType"),
    (["--noGETT"], "nothing to emit"),
])
def test_cli_rejects(tmp_path, capsys, argv, err):
    f = tmp_path / "e.tccg"
    f.write_text(WORKLOADS["example"])
    assert main(argv + ["--generateOnly", str(f)]) != 0
    assert err in capsys.readouterr().err


def test_cli_syntax_error_exit_status(tmp_path):
    f = tmp_path / "bad.tccg"
    f.write_text("C[a,b] += A[a,k] * B[k,b]\na=2\nb=3\nk=4\n")
    proc = subprocess.run([sys.executable, "-m", "tccg_amd", "--generateOnly", str(f)], cwd=ROOT,
                          stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    assert proc.returncode != 0 and "+=" in proc.stderr


def test_cli_missing_file(tmp_path, capsys):
    assert main(["--generateOnly", str(tmp_path / "nope.tccg")]) != 0


def test_cli_batch_index_rejected(tmp_path, capsys):
    f = tmp_path / "l.tccg"
    f.write_text("C[a,b,l] = A[a,k,l] * B[k,b,l]\na=2\nb=3\nk=4\nl=5\n")
    assert main(["--generateOnly", str(f)]) != 0
    assert "loop indices" in capsys.readouterr().err


def test_emit_source_tables_and_swap():
    # B holds C's stride-one index: the generated entry swaps the operands internally
    c = Contraction.from_labels(["b", "a"], ["a", "k"], ["k", "b"], dict(a=8, b=12, k=4))
    p = c.plan(F32)
    src = codegen.emit_source(p, "C[b,a] = A[a,k] * B[k,b]")
    assert "const float* opA = 1 ? B : A;" in src
    assert "tccg_gett_offAm[%d]" % p.Mp in src and "tccg_gett_offCn[%d]" % p.Np in src


def test_emit_fp8_reduces_amax():
    c = Contraction.from_tccg(WORKLOADS["example"])
    src = codegen.emit_source(c.plan(FP8), WORKLOADS["example"])
    assert "amax_reduce" in src and "t->amax" in src


def test_generated_tables_are_exported_device_symbols(tmp_path):
    """The embedded offset tables must be exported symbols of the device code
    object (a table dropped by the device compile, or kept as a local symbol,
    leaves the launch without valid table addresses)."""
    import re
    from tccg_amd import build
    c = Contraction.from_tccg(WORKLOADS["example"])
    name = "gett_regress"
    src = tmp_path / (name + ".hip")
    src.write_text(codegen.emit_source(c.plan(F32), WORKLOADS["example"], name))
    asm = tmp_path / "dev.s"
    cmd = [build.HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-I", str(build.CSRC), "--cuda-device-only",
           "-S", str(src), "-o", str(asm)]
    subprocess.run(cmd, check=True, stdout=subprocess.PIPE, stderr=subprocess.PIPE)
    text = asm.read_text()
    for key in ("offAm", "offAk", "offBn", "offBk", "offCm", "offCn", "ordA", "ordB"):
        sym = "tccg_%s_%s" % (name, key)
        assert re.search(r"^\s*\.globl\s+%s\s*$" % sym, text, flags=re.M), sym


def test_plan_kind_round_trip():
    """Every candidate plan is rebuilt exactly from its tuning-database entry."""
    from tccg_amd import Contraction
    from tccg_amd.plan import BF16X3, F32, candidate_plans, plan_from_kind, plan_kind
    from tccg_amd.workloads import get
    for wl in ("abc-akb-kc", "abcdef-dgfa-gceb"):
        c = Contraction.from_tccg(get(wl))
        for prec in (F32, BF16X3):
            kinds = set()
            for p in candidate_plans(c.problem, prec):
                k = plan_kind(p)
                kinds.add(k)
                q = plan_from_kind(c.problem, prec, p.tile, not p.lin_m.is_plain(), k)
                assert q.describe() == p.describe()
                assert all((a == b).all() for a, b in ((p.offAm, q.offAm), (p.offAk, q.offAk), (p.offCn, q.offCn)))
                assert candidate_name(q) == candidate_name(p)
            assert any(k & 1 for k in kinds)   # packed plans are candidates


def test_tuning_db_kind_and_migration(tmp_path):
    import sqlite3
    from tccg_amd import Contraction
    from tccg_amd.tuning_db import TuningDB
    path = tmp_path / "old.db"
    con = sqlite3.connect(str(path))   # schema without the kind column
    con.executescript("""
CREATE TABLE measurements (measurement_id INTEGER PRIMARY KEY AUTOINCREMENT, tensorA TEXT NOT NULL,
  tensorB TEXT NOT NULL, tensorC TEXT NOT NULL, device TEXT NOT NULL, prec TEXT NOT NULL,
  beta_nonzero INTEGER NOT NULL, created REAL NOT NULL, UNIQUE(tensorA, tensorB, tensorC, device, prec, beta_nonzero));
CREATE TABLE gett (gett_id INTEGER PRIMARY KEY AUTOINCREMENT, measurement_id INTEGER NOT NULL,
  name TEXT NOT NULL, tile INTEGER NOT NULL, m_blocked INTEGER NOT NULL, gflops REAL NOT NULL,
  estimated_gflops REAL NOT NULL);
""")
    con.close()
    db = TuningDB(path)
    c = Contraction.from_tccg(WORKLOADS["example"])
    mid = db.measurement_id(c, "gfx950", "bf16x3", 0.0, create=True)
    db.insert_gett(mid, "a", 1, False, 10.0, 5.0, kind=0)
    db.insert_gett(mid, "b", 2, True, 30.0, 5.0, kind=3, shape=(128, 128, 16))
    assert db.fastest(c, "gfx950", "bf16x3", 0.0) == ("b", 2, True, 3, 30.0, (128, 128, 16))
    db.close()


def test_tuning_db_entries_check_the_tile_shape():
    """A database row names its tile by id and records the shape: a row whose
    id names another shape today (a reordered or shortened table) is not
    reused, and the device key carries a digest of the tile tables."""
    from tccg_amd import plan as P
    assert P.tile_shape_of(P.BF16X3, 2, P.KIND_PACKED) == tuple(P.PACKED_TILES[P.BF16X3][2])
    assert P.tile_shape_of(P.F32, 0, P.KIND_VECTOR) == tuple(P.VECTOR_TILES[P.F32][0])
    assert P.tile_shape_of(P.F32, 999, 0) is None
    d = P.tile_tables_digest()
    saved = list(P.PACKED_TILES[P.F32])
    try:
        P.PACKED_TILES[P.F32].reverse()
        assert P.tile_tables_digest() != d
    finally:
        P.PACKED_TILES[P.F32][:] = saved
    assert P.tile_tables_digest() == d
    from tccg_amd import Contraction
    c = Contraction.from_tccg(WORKLOADS["example"])
    with pytest.raises(ValueError):
        P.plan_from_kind(c.problem, P.F32, 999, False, P.KIND_PACKED)


def test_emit_packed_source():
    from tccg_amd import Contraction
    from tccg_amd.plan import BF16X3, build_packed_plan
    c = Contraction.from_tccg(WORKLOADS["example"])
    plan = build_packed_plan(c.problem, BF16X3, 0)
    src = codegen.emit_source(plan, WORKLOADS["example"], "gett")
    assert "gett_packed_kernel<tccg::BF16X3, 128, 128, 16" in src
    # both operands packed in one launch
    assert src.count("tccg::pack_pair_kernel<tccg::BF16X3, 128, 128, 16>") == 1
    assert "hipMalloc(&t.ws, %d)" % plan.workspace_bytes() in src
