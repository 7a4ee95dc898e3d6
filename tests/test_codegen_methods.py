"""Emitted TTGT and LoG libraries (tccg_amd/codegen_methods.py), host side:
the reference's method entry points and the uniform gett() boundary are
exported, the baked-in tables are the plans' own, and the LoG records
(relative to the operand bases) list every GEMM instance once."""
import re
import subprocess

import numpy as np
import pytest

from tccg_amd import Contraction
from tccg_amd.build import hipcc_available
from tccg_amd.codegen_methods import emit_log_source, emit_ttgt_source, generate_method
from tccg_amd.loop_gemm import best_log_candidates, log_candidates, log_layout
from tccg_amd.plan import F32, F64
from tccg_amd.ttgt import best_candidates, ttgt_candidates
from tccg_amd.workloads import scaled

SMALL = scaled("abcd-aebf-dfce", 0.34)


def _array(src, name):
    m = re.search(r"%s\[(\d+)\] = \{([^}]*)\}" % name, src)
    return int(m.group(1)), np.array([int(x) for x in m.group(2).replace("\n", "").split(",")], dtype=np.int64)


def test_log_source_records():
    c = Contraction.from_tccg(SMALL)
    cand = best_log_candidates(log_candidates(c.problem), 1)[0]
    src = emit_log_source(c, cand, F32, SMALL)
    lay = log_layout(c.problem, cand, F32)
    n, probC = _array(src, "kProbC")
    assert n == len(lay.free_bases)
    # C offsets are the instances' bases in bytes, each once
    assert sorted(probC.tolist()) == sorted(4 * fc for _, _, fc in lay.free_bases)
    n_ops, opx = _array(src, "kOpX")
    assert n_ops == len({fa for fa, _, _ in lay.free_bases}) + len({fb for _, fb, _ in lay.free_bases})
    assert "int %s(const float* A, const float* B, float* C, float alpha, float beta)" % cand.name in src
    assert "tccg_gett_group(" in src and "int gett_stream(" in src


def test_ttgt_source_calls():
    c = Contraction.from_tccg(SMALL)
    cand = best_candidates(ttgt_candidates(c.problem, 4), 1)[0]
    sizes = {i.label: i.size for X in (c.problem.opA, c.problem.opB, c.problem.C) for i in X.indices}
    gemm = Contraction.from_labels(cand.idxC, cand.idxA, cand.idxB, sizes)
    plan = gemm.plan(F32)
    src = emit_ttgt_source(c, cand, F32, plan, SMALL)
    assert src.count("tccg_transpose_f32(") == int(cand.transA) + int(cand.transB) + int(cand.transC)
    assert ("tccg_gett_packed_f32(" in src) == plan.packed
    assert "int ttgemmt_%s(const float* A, const float* B, float* C, float alpha, float beta, float* work_)" % cand.name in src
    # the GEMM plan's m table is baked in at the offset the plan struct names
    n, words = _array(src, "ttgt_tables")
    off = int(re.search(r"gp.offAm = W\((\d+)\)", src).group(1)) // 4
    assert np.array_equal(words[off:off + plan.Mp].astype(np.uint32), plan.offAm)


@pytest.mark.skipif(not hipcc_available(), reason="hipcc is not available")
@pytest.mark.parametrize("kind,prec", [("ttgt", F32), ("log", F64)])
def test_generated_method_exports(kind, prec, tmp_path):
    c = Contraction.from_tccg(SMALL)
    if kind == "ttgt":
        cand = best_candidates(ttgt_candidates(c.problem, 8 if prec == F64 else 4), 1)[0]
        entry = "ttgemmt_" + cand.name
    else:
        cand = best_log_candidates(log_candidates(c.problem), 1)[0]
        entry = cand.name
    res = generate_method(kind, c, cand, prec, SMALL, tmp_path)
    syms = subprocess.run(["nm", "-D", "--defined-only", str(res["library"])], capture_output=True, text=True).stdout
    for name in (entry, "gett", "gett_stream", "gett_info", "gett_contraction", "gett_last_error"):
        assert re.search(r" T %s$" % name, syms, re.M), name
    assert entry in res["header"].read_text()


def test_ttgt_transposition_tables_baked_in():
    """Every transposition plan's six tables sit in the baked-in blob at the
    offsets its plan struct names (tables of plans built one after the other
    must not be confused)."""
    from tccg_amd.transpose import plan_transpose
    c = Contraction.from_tccg(SMALL)
    cand = best_candidates(ttgt_candidates(c.problem, 4), 1)[0]
    pb = c.problem
    sizes = {i.label: i.size for X in (pb.opA, pb.opB, pb.C) for i in X.indices}
    gemm = Contraction.from_labels(cand.idxC, cand.idxA, cand.idxB, sizes)
    src = emit_ttgt_source(c, cand, F32, gemm.plan(F32), SMALL)
    _, words = _array(src, "ttgt_tables")
    X, Y = (pb.opB, pb.opA) if cand.swap else (pb.opA, pb.opB)
    plans = {"pA": (cand.transA, X.labels(), cand.idxA), "pB": (cand.transB, Y.labels(), cand.idxB),
             "pC": (cand.transC, cand.idxC, pb.C.labels())}
    checked = 0
    for var, (need, src_l, dst_l) in plans.items():
        if not need:
            continue
        tp = plan_transpose(src_l, dst_l, sizes, elem_bytes=4)
        for field in ("inX", "outX", "inY", "outY", "inZ", "outZ"):
            off = int(re.search(r"%s\.%s = W\((\d+)\)" % (var, field), src).group(1)) // 4
            arr = getattr(tp, field)
            assert np.array_equal(words[off:off + arr.size].astype(np.uint32), arr.astype(np.uint32)), (var, field)
            checked += 1
    assert checked >= 6
