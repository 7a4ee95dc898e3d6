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


@pytest.mark.parametrize("seed", range(48))
def test_random_cases_transpositions_and_gemm(seed):
    """The random cases of the GPU sweep (tests/random_cases.py): for the
    candidates the GPU test runs (the model's best three, one that transposes
    C, one that exchanges A and B), every transposition plan through the
    transposition kernels' addressing model (every access in bounds, every
    output written once) against the transposition oracle, and the GEMM's
    default plan through the GETT addressing model, in FP32 and FP64 layouts."""
    from tccg_amd.plan import F32, F64
    from tccg_amd.transpose import plan_transpose
    from .random_cases import random_contraction
    from .test_plan_model import emulate as gemm_emulate
    from .test_transpose import emulate as tr_emulate
    c = random_contraction(seed)
    pb = c.problem
    sizes = {i.label: i.size for T in (pb.opA, pb.opB, pb.C) for i in T.indices}
    rng = np.random.default_rng(seed)
    for prec, esz in ((F32, 4), (F64, 8)):
        cands = ttgt_candidates(pb, 8 if prec == F64 else 4)
        chosen = best_candidates(cands, 3)
        for pred in (lambda x: x.transC, lambda x: x.swap == 1):
            extra = next((x for x in cands if pred(x)), None)
            if extra is not None and extra not in chosen:
                chosen.append(extra)
        for cand in chosen:
            X, Y = (pb.opB, pb.opA) if cand.swap else (pb.opA, pb.opB)
            for need, src, dst in ((cand.transA, X.labels(), cand.idxA), (cand.transB, Y.labels(), cand.idxB),
                                   (cand.transC, cand.idxC, pb.C.labels())):
                if not need:
                    continue
                plan = plan_transpose(src, dst, sizes, elem_bytes=esz)
                shape = [sizes[l] for l in reversed(list(src))]
                T = rng.standard_normal(shape)
                OUT = rng.standard_normal([sizes[l] for l in reversed(list(dst))])
                got = tr_emulate(plan, T, OUT, 1.0, 0.0)
                np.testing.assert_array_equal(got.reshape(-1), transpose_ref(src, dst, T).reshape(-1))
            gemm = Contraction.from_labels(cand.idxC, cand.idxA, cand.idxB, sizes)
            sa, sb, sc = gemm.shapes()
            A, B, C = rng.standard_normal(sa), rng.standard_normal(sb), rng.standard_normal(sc)
            ref = contract_ref(cand.idxC, cand.idxA, cand.idxB, A, B, C, 1.5, -0.5)
            out = gemm_emulate(gemm.plan(prec), A, B, C, 1.5, -0.5)
            assert rel_frobenius(out, ref) < 1e-12, (cand.name, prec)
