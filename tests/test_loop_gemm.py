"""LoG planning: the reference's candidate rules and selection, and a CPU
model of the batched launches (shared GEMM tables, per-problem base offsets,
accumulating contracted loops) checked against the FP64 einsum oracle, with
every access asserted in bounds."""
import numpy as np
import pytest

from oracle.einsum_ref import contract_ref, rel_frobenius
from tccg_amd import Contraction
from tccg_amd.loop_gemm import best_log_candidates, log_candidates, log_layout
from tccg_amd.plan import FREE4, INVALID, KDIM4
from tccg_amd.workloads import WORKLOADS, scaled


def test_known_candidates():
    c = Contraction.from_tccg(WORKLOADS["abcd-aebf-dfce"])
    cands = log_candidates(c.problem)
    # m = a (C's stride-one), n = d (B's stride-one), k free among e, f; batch among b, c
    assert sorted(x.name for x in cands) == sorted([
        "LoG_free_c_contracted_f_batch_b_ade", "LoG_free_b_contracted_f_batch_c_ade",
        "LoG_free_c_contracted_e_batch_b_adf", "LoG_free_b_contracted_e_batch_c_adf"])
    assert all(x.gemm_size() == 72 * 72 * 48 and x.sizes[0] == 72 for x in cands)
    # A's stride-one index d is an m index other than C's: no single GEMM holds both
    assert log_candidates(Contraction.from_tccg(WORKLOADS["abcdef-dgfa-gceb"]).problem) is None
    # largest GEMM first, then largest batch
    t = Contraction.from_labels(["a", "b", "c"], ["a", "k", "b"], ["k", "c"], dict(a=8, b=4, c=16, k=3))
    best = best_log_candidates(log_candidates(t.problem), 8)
    assert [x.name for x in best] == ["LoG_free_dummy_contracted_dummy_batch_b_ack"]


def emulate(c, cand, A, B, C, alpha, beta):
    pb = c.problem
    lay = log_layout(pb, cand)
    opA, opB = (B, A) if pb.swapped else (A, B)
    a, b = np.asarray(opA, np.float64).reshape(-1), np.asarray(opB, np.float64).reshape(-1)
    out = np.asarray(C, np.float64).reshape(-1).copy()
    offAm, offAk, offBn, offBk, offCm, offCn = (t.astype(np.int64) for t in lay.tables[:6])
    vm, vn, vk = offAm != INVALID, offBn != INVALID, offAk != INVALID
    assert np.array_equal(vm, offCm != INVALID) and np.array_equal(vn, offCn != INVALID)
    assert np.array_equal(vk, offBk != INVALID)
    Am, Ak, Bn, Bk, Cm, Cn = offAm[vm], offAk[vk], offBn[vn], offBk[vk], offCm[vm], offCn[vn]
    written = np.zeros(out.size, dtype=np.int64)
    for j, (ca, cb, _) in enumerate(lay.contracted_bases):
        bj = beta if j == 0 else 1.0
        for (fa, fb, fc) in lay.free_bases:
            ia = fa + ca + Am[:, None] + Ak[None, :]
            ib = fb + cb + Bk[:, None] + Bn[None, :]
            ic = fc + Cm[:, None] + Cn[None, :]
            assert ia.min() >= 0 and ia.max() < a.size and ib.max() < b.size and ic.max() < out.size
            # four-wide load modes: groups of four consecutive, 16-byte aligned elements
            for mode, idx, free_axis in ((lay.modeA, ia, 0), (lay.modeB, ib.T, 0)):
                if mode in (FREE4, KDIM4):
                    g = idx if mode == FREE4 else idx.T
                    assert g.shape[0] % 4 == 0 and np.all(g[::4] % 4 == 0)
                    for v in range(1, 4):
                        assert np.array_equal(g[v::4], g[::4] + v)
            prod = alpha * (a[ia] @ b[ib])
            out[ic] = prod + (bj * out[ic] if bj != 0 else 0)
            if j == 0:
                written[ic.reshape(-1)] += 1
    assert np.all(written == 1)     # every C element belongs to exactly one problem
    return out.reshape(np.shape(C))


CASES = [("abcd-aebf-dfce", 0.12), ("abc-akb-kc", 0.06), ("abcd-aklb-lckd", 0.12), ("cpu48", 0.2),
         ("abcd-aebf-dfce", 0.34), ("abc-akb-kc", 0.125), ("cpu48", 0.34)]   # the last three: four-wide loads


@pytest.mark.parametrize("wl,f", CASES)
def test_model_matches_oracle(wl, f):
    c = Contraction.from_tccg(scaled(wl, f))
    rng = np.random.default_rng(7)
    sa, sb, sc = c.shapes()
    A, B, C = rng.standard_normal(sa), rng.standard_normal(sb), rng.standard_normal(sc)
    lab = lambda T: [i.label for i in T.indices]
    ref = contract_ref(lab(c.C0), lab(c.A0), lab(c.B0), A, B, C, 1.3, -0.7)
    cands = log_candidates(c.problem)
    assert cands
    for cand in cands[:6]:
        assert rel_frobenius(emulate(c, cand, A, B, C, 1.3, -0.7), ref) < 1e-13, cand.name


def test_problem_records_match_layout():
    """The batched launch's problem records: pointers are the bases the CPU
    model checked, in FP32 elements, and the record layout is the C ABI's."""
    import ctypes
    from tccg_amd.batch import _CProblem
    from tccg_amd.loop_gemm import PROBLEM_DTYPE, problem_records
    assert PROBLEM_DTYPE.itemsize == ctypes.sizeof(_CProblem)
    for name in PROBLEM_DTYPE.names:
        assert PROBLEM_DTYPE.fields[name][1] == getattr(_CProblem, name).offset
    c = Contraction.from_tccg(scaled("abcd-aebf-dfce", 0.12))
    cand = log_candidates(c.problem)[0]
    lay = log_layout(c.problem, cand)
    recs = problem_records(lay, 1 << 20, 2 << 20, 3 << 20, list(range(8)), 2.0, 0.5)
    assert len(recs) == len(lay.contracted_bases)
    for j, r in enumerate(recs):
        ca, cb, _ = lay.contracted_bases[j]
        fb = np.array(lay.free_bases)
        assert np.array_equal(r["A"], (1 << 20) + 4 * (fb[:, 0] + ca))
        assert np.array_equal(r["B"], (2 << 20) + 4 * (fb[:, 1] + cb))
        assert np.array_equal(r["C"], (3 << 20) + 4 * fb[:, 2])
        assert np.all(r["beta"] == (0.5 if j == 0 else 1.0)) and np.all(r["alpha"] == 2.0)
        assert np.all(r["tilesM"] == lay.tiles[0]) and np.all(r["offCn"] == 5)
