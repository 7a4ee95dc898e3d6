"""TTGT planning: the reference's candidate enumeration and cost model, and
a CPU run of each chosen candidate (transpositions by the oracle, the GEMM by
einsum) that must equal the contraction."""
import numpy as np
import pytest

from oracle.einsum_ref import contract_ref, rel_frobenius, transpose_ref
from tccg_amd import Contraction
from tccg_amd.ttgt import best_candidates, ttgt_candidates
from tccg_amd.workloads import WORKLOADS, scaled


def test_enumeration_size_and_names():
    c = Contraction.from_tccg(WORKLOADS["abcd-aebf-dfce"])
    cands = ttgt_candidates(c.problem)
    # m!, n!, k! orders x A/B exchange x 2 x 2 orientations (tccg/ttgemmt.py:168-203)
    assert len(cands) == 2 * 2 * 2 * 2 * 4
    assert len({x.name for x in cands}) == len(cands)
    assert cands[0].name == "ab_cd_ef_0NN"


def test_best_candidates_known_answer():
    """C[a,b,c,d] = A[a,e,b,f] B[d,f,c,e]: A and B always need a transposition,
    C none when m = (a,b), n = (c,d); A keeps its stride-one index a first
    (full copy bandwidth), B's moves (0.71).  Four candidates tie: both k
    orders x both orientations of B."""
    c = Contraction.from_tccg(WORKLOADS["abcd-aebf-dfce"])
    best = best_candidates(ttgt_candidates(c.problem), 16)
    assert sorted(x.name for x in best) == ["ab_cd_ef_0NN", "ab_cd_ef_0NT", "ab_cd_fe_0NN", "ab_cd_fe_0NT"]
    for x in best:
        assert x.transA and x.transB and not x.transC
    n = 72 * 48 * 72 * 48 * 4
    assert best[0].cost == pytest.approx(n / 8e12 + n / (8e12 * 0.71))


def _cpu_ttgt(c, cand, A, B, C, alpha, beta):
    pb = c.problem
    opA, opB = (B, A) if pb.swapped else (A, B)
    X, Y = (opB, opA) if cand.swap else (opA, opB)
    lX, lY = (pb.opB.labels(), pb.opA.labels()) if cand.swap else (pb.opA.labels(), pb.opB.labels())
    sizes = {i.label: i.size for T in (pb.opA, pb.opB, pb.C) for i in T.indices}
    shape = lambda labels: [sizes[l] for l in reversed(labels)]
    X = X.reshape(shape(lX))
    Y = Y.reshape(shape(lY))
    if cand.transA:
        X = transpose_ref(lX, cand.idxA, X)
    if cand.transB:
        Y = transpose_ref(lY, cand.idxB, Y)
    Cm = C.reshape(shape(pb.C.labels()))
    if cand.transC:
        Ct = contract_ref(cand.idxC, cand.idxA, cand.idxB, X, Y, None, alpha, 0.0)
        out = transpose_ref(cand.idxC, pb.C.labels(), Ct, 1.0, beta, Cm)
    else:
        out = contract_ref(cand.idxC, cand.idxA, cand.idxB, X, Y, Cm, alpha, beta)
    return out.reshape(C.shape)


@pytest.mark.parametrize("wl", ["example", "abcd-aebf-dfce", "abc-akb-kc", "abcdef-dgfa-gceb"])
def test_candidates_compute_the_contraction(wl):
    c = Contraction.from_tccg(scaled(wl, 0.15) if wl != "example" else WORKLOADS[wl])
    rng = np.random.default_rng(1)
    sa, sb, sc = c.shapes()
    A, B, C = rng.standard_normal(sa), rng.standard_normal(sb), rng.standard_normal(sc)
    labels = lambda T: [i.label for i in T.indices]
    ref = contract_ref(labels(c.C0), labels(c.A0), labels(c.B0), A, B, C, 1.7, 0.3)
    cands = ttgt_candidates(c.problem)
    for cand in best_candidates(cands, 16) + cands[:: max(1, len(cands) // 12)]:
        assert rel_frobenius(_cpu_ttgt(c, cand, A, B, C, 1.7, 0.3), ref) < 1e-13, cand.name
